// lbm_mea.hip -- the momentum-exchange force on bounce-back bodies
// (MEA_CL / MEA_CD, lbm.cpp:3083-3163) as one FP64 reduction on gfx950.
//
// The reference evaluates it after Collision, Streaming and MacroProp: over
// the TYPE_F cells x of a box and the directions l whose source x - c_l is
// TYPE_S, it sums (f*_l(x) + f_opp(l)(x)) with f* the collision's output and
// f the streamed populations, weighted by -c_lx for the drag and +c_ly for
// the lift (the two signs differ in the reference; kept).  After a streaming
// launch the two A/B buffers hold exactly those two arrays: `oth` the f* the
// launch pulled from and `cur` the streamed f, so the kernel reads both in
// place.  The bounce-back links of each wall cell are the bits 0..18 of its
// link word (source x - c_m is solid), built with the wall tables.
//
// One thread per box cell; per-block FP64 tree sums over fixed 4096-cell
// chunks, then one block sums the partials (the partition is independent of
// the device, so the result is reproducible).
#include <hip/hip_runtime.h>

#include "lbm_ctx.h"
#include "lbm_step.h"

using namespace lbm;

namespace {

constexpr int MEA_THREADS = 256;
constexpr uint32_t MEA_CHUNK = 4096;

__device__ __forceinline__ double mea_block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = MEA_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

struct MeaBox {
  int x0, y0, z0, nx, ny, nz;
};

// Block b sums the wall links of box cells [b*chunk, (b+1)*chunk); the
// partials are stored component-major: partial[c * nblocks + b].
template <class S>
__global__ __launch_bounds__(MEA_THREADS) void k_mea(const S *__restrict__ fstar, const S *__restrict__ fcur,
                                                     const uint8_t *__restrict__ mask,
                                                     const uint32_t *__restrict__ links, Geo g, MeaBox B,
                                                     int nblocks, double *__restrict__ partial) {
  __shared__ double sh[MEA_THREADS];
  const uint32_t n = (uint32_t)B.nx * B.ny * B.nz;
  const uint32_t c0 = blockIdx.x * MEA_CHUNK, c1 = min(n, c0 + MEA_CHUNK);
  double px = 0, py = 0, pz = 0;
  for (uint32_t b = c0 + threadIdx.x; b < c1; b += MEA_THREADS) {
    const int i = B.x0 + (int)(b % B.nx), j = B.y0 + (int)((b / B.nx) % B.ny), k = B.z0 + (int)(b / (B.nx * B.ny));
    const uint32_t idx = ((uint32_t)k * g.ny + (uint32_t)j) * g.px + (uint32_t)i;
    const uint8_t mk = mask[idx];
    if (!(mk & M_FLUID)) continue;
    // the wall links: the link word's and those folded into the pull
    const uint32_t lk = ((mk & M_WALL) ? links[idx] : 0u) | fold_mask(g, (uint32_t)i, (uint32_t)j, (uint32_t)k);
    if (!lk) continue;
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      if constexpr (m > 0) {
        if ((lk >> m) & 1u) {
          double fs = (double)Store<S>::ld(fstar + (size_t)m * g.ps + idx);
          double fo = (double)Store<S>::ld(fcur + (size_t)OPP[m] * g.ps + idx);
          if constexpr (Store<S>::shifted) {
            fs = fs / (double)Store<S>::K + W[m];
            fo = fo / (double)Store<S>::K + W[OPP[m]];
          }
          const double t = fs + fo;
          px += -t * CX[m];
          py += t * CY[m];
          pz += t * CZ[m];
        }
      }
    });
  }
  const double sx = mea_block_sum(px, sh), sy = mea_block_sum(py, sh), sz = mea_block_sum(pz, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = sx;
    partial[nblocks + blockIdx.x] = sy;
    partial[2 * nblocks + blockIdx.x] = sz;
  }
}

__global__ __launch_bounds__(MEA_THREADS) void k_mea_sum(const double *__restrict__ partial, int nblocks,
                                                         double *__restrict__ out) {
  __shared__ double sh[MEA_THREADS];
  for (int c = 0; c < 3; c++) {
    double s = 0.0;
    for (int q = threadIdx.x; q < nblocks; q += MEA_THREADS) s += partial[(size_t)c * nblocks + q];
    const double t = mea_block_sum(s, sh);
    if (threadIdx.x == 0) out[c] = t;
  }
}

template <class S>
void launch_mea(const lbm_ctx *c, const MeaBox &B, int nb) {
  hipLaunchKernelGGL(k_mea<S>, dim3((unsigned)nb), dim3(MEA_THREADS), 0, c->stream, (const S *)c->oth,
                     (const S *)c->cur, c->mask, c->links, c->g, B, nb, c->red_partial);
}

}  // namespace

extern "C" {

// MEA_CD / MEA_CL's momentum sums (lbm.cpp:3085-3163) over the box
// [x0,x1]x[y0,y1]x[z0,z1] (reference indices, inclusive, within 0..N+1):
// out[0] = Px = -sum (f*_l + f_opp) c_lx, out[1] = Py = +sum (f*_l + f_opp)
// c_ly, out[2] = Pz = +sum (f*_l + f_opp) c_lz (the reference has no z; its
// y form).  A TYPE_F cell on the ghost shell takes its wrapped neighbour
// (the reference's unwrapped x - c_l would leave its array there).  dx =
// dt = 1, so the reference's pow(dx,3)/dt factor is 1; C_D and C_L are
// these over 1/2 rho0 u^2 times the body's reference length or area.
int lbm_momentum_exchange(lbm_ctx *c, const int32_t box[6], double out[3]) {
  if (!c || !box || !out) return LBM_ERR_ARG;
  if (c->state != ST_F || !c->fstar_ready) return LBM_ERR_STATE;
  const int Nx = (int)c->g.nx - 2, Ny = (int)c->g.ny - 2, Nz = (int)c->g.nz - 2;
  if (box[0] < 0 || box[1] > Nx + 1 || box[0] > box[1] || box[2] < 0 || box[3] > Ny + 1 || box[2] > box[3] ||
      box[4] < 0 || box[5] > Nz + 1 || box[4] > box[5])
    return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  const MeaBox B{box[0], box[2], box[4], box[1] - box[0] + 1, box[3] - box[2] + 1, box[5] - box[4] + 1};
  const uint32_t n = (uint32_t)B.nx * B.ny * B.nz;
  const int nb = (int)((n + MEA_CHUNK - 1) / MEA_CHUNK);
  if (!c->links) {  // no wall cell anywhere: no link to sum
    out[0] = out[1] = out[2] = 0.0;
    return LBM_OK;
  }
  switch (c->storage) {
    case LBM_STORE_F64: launch_mea<double>(c, B, nb); break;
    case LBM_STORE_F32: launch_mea<float>(c, B, nb); break;
    case LBM_STORE_F16S: launch_mea<__half>(c, B, nb); break;
    default: launch_mea<f16c_t>(c, B, nb); break;
  }
  hipLaunchKernelGGL(k_mea_sum, dim3(1), dim3(MEA_THREADS), 0, c->stream, c->red_partial, nb, c->red_out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, c->red_out, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

}  // extern "C"
