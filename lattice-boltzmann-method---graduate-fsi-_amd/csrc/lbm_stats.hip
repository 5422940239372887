// lbm_stats.hip -- the reference's MORE_DETAILS fields on gfx950: the
// viscous stress and pressure of MacroProp (lbm.cpp:681-705) and the running
// averages and rms of Sup_MacroProp (lbm.cpp:865-939).
//
// The stress is computed on demand from the state the last MacroProp saw
// (the streamed f in the current buffer, rho and u in the macro arrays), so
// the fused step carries no extra traffic for it.  The statistics live in
// 26 FP64 planes of the lattice (average and rms of u, p and the
// fluctuation products), updated in place by one elementwise kernel.  The
// reference's vorticity is never computed (its block in MacroProp is
// commented out, lbm.cpp:707-762), so its averages stay zero; they are not
// stored.
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "lbm_ctx.h"
#include "lbm_step.h"

#pragma clang fp contract(off)

using namespace lbm;

namespace {

// statistics planes
enum {
  ST_AVG_U = 0,   // 3: avg_velocity[0..2]
  ST_RMS_U = 3,   // 3: rms_velocity[0..2]
  ST_AVG_P = 6,   // avg_pressure
  ST_RMS_P = 7,   // rms_pressure
  ST_AVG_RS = 8,  // 9: avg_ReynoldStress[a][b], row-major
  ST_RMS_RS = 17, // 9: rms_ReynoldStress[a][b]
  STAT_PLANES = 26
};
static_assert(STAT_PLANES == LBM_STAT_PLANES, "the context's statistics planes");

// MacroProp's MORE_DETAILS block (lbm.cpp:681-705): pressure = rho/3 and
// sigma_ab = -(tau - 1/2)/tau sum_l f_l (c_la - u_a)(c_lb - u_b), for every
// cell.  TYPE_F cells (and TYPE_S cells unless their slots carry
// bounce-back values) read f from the buffer; other cells use feq of their
// stored rho, u (their Init populations, which the device never changes).
template <class S>
__global__ __launch_bounds__(256) void k_stress(const S *__restrict__ buf, const uint8_t *__restrict__ mask, Geo g,
                                                const typename Store<S>::C *__restrict__ rho,
                                                const typename Store<S>::C *__restrict__ vx,
                                                const typename Store<S>::C *__restrict__ vy,
                                                const typename Store<S>::C *__restrict__ vz, int wall_slots,
                                                const int32_t *__restrict__ fslot, const double *__restrict__ fof,
                                                int nslot,
                                                double tau, double *__restrict__ sig, double *__restrict__ pres) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell || idx % g.px >= g.nx) return;
  const uint8_t mk = mask[idx];
  const C r = rho[idx], cux = vx[idx], cuy = vy[idx], cuz = vz[idx];
  const C uu = cux * cux + cuy * cuy + cuz * cuz;
  const double ux = (double)cux, uy = (double)cuy, uz = (double)cuz;
  const bool from_buf = (mk & M_FLUID) || ((mk & M_SOLID) && !wall_slots);
  // inlet/outlet and the other types: their own populations
  const int32_t k = (fslot && !(mk & M_FLUID)) ? fslot[idx] : -1;
  double s[6] = {0, 0, 0, 0, 0, 0};  // xx xy xz yy yz zz
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    double f;
    if (k >= 0) {
      f = fof[(size_t)m * nslot + k];
    } else if (from_buf) {
      f = (double)Store<S>::ld(buf + (size_t)m * g.ps + idx);
      if constexpr (Store<S>::shifted) f = f / (double)Store<S>::K + W[m];
    } else {
      const C cu = C(CX[m]) * cux + C(CY[m]) * cuy + C(CZ[m]) * cuz;
      f = (double)(r * C(W[m]) * (C(1) + eq_A(cu, uu)));
    }
    const double dx = CX[m] - ux, dy = CY[m] - uy, dz = CZ[m] - uz;
    s[0] += f * dx * dx;
    s[1] += f * dx * dy;
    s[2] += f * dx * dz;
    s[3] += f * dy * dy;
    s[4] += f * dy * dz;
    s[5] += f * dz * dz;
  });
  double *o = sig + (size_t)idx * 9;
  const double xx = -(tau - 0.5) * s[0] / tau, xy = -(tau - 0.5) * s[1] / tau, xz = -(tau - 0.5) * s[2] / tau,
               yy = -(tau - 0.5) * s[3] / tau, yz = -(tau - 0.5) * s[4] / tau, zz = -(tau - 0.5) * s[5] / tau;
  o[0] = xx;
  o[1] = xy;
  o[2] = xz;
  o[3] = xy;
  o[4] = yy;
  o[5] = yz;
  o[6] = xz;
  o[7] = yz;
  o[8] = zz;
  pres[idx] = (double)r / 3.0;
}

// Sup_MacroProp(n), lbm.cpp:865-939, with told = (n - Nt) dt and dt = 1: the
// averages and rms of u, p (= rho/3, MacroProp's pressure) and the
// fluctuation products, each as the reference updates it (the new average
// enters the rms; rms_pressure weights pressure^2, not its old rms, by told;
// the rms of the Reynolds stress compares the raw products u_a u_b with the
// averaged fluctuation products).
template <class C>
__global__ __launch_bounds__(256) void k_stats(Geo g, const C *__restrict__ rho, const C *__restrict__ vx,
                                               const C *__restrict__ vy, const C *__restrict__ vz, double told,
                                               double *__restrict__ st) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell || idx % g.px >= g.nx) return;
  const size_t n = g.ncell;
  double *P[STAT_PLANES];
  for (int q = 0; q < STAT_PLANES; q++) P[q] = st + (size_t)q * n + idx;
  const double dt = 1.0;
  const double u[3] = {(double)vx[idx], (double)vy[idx], (double)vz[idx]};
  const double p = (double)rho[idx] / 3.0;
  double avg[3];
  for (int a = 0; a < 3; a++) {
    avg[a] = (*P[ST_AVG_U + a] * told + u[a] * dt) / (told + dt);
    *P[ST_AVG_U + a] = avg[a];
  }
  const double avg_p = (*P[ST_AVG_P] * told + p * 1) / (told + 1);
  *P[ST_AVG_P] = avg_p;
  double ars[9];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      ars[3 * a + b] = (*P[ST_AVG_RS + 3 * a + b] * told + (u[a] - avg[a]) * (u[b] - avg[b]) * 1) / (told + 1);
      *P[ST_AVG_RS + 3 * a + b] = ars[3 * a + b];
    }
  for (int a = 0; a < 3; a++) {
    const double r = *P[ST_RMS_U + a];
    *P[ST_RMS_U + a] = sqrt((r * r * told + (u[a] - avg[a]) * (u[a] - avg[a]) * 1) / (told + 1));
  }
  *P[ST_RMS_P] = sqrt((p * p * told + (p - avg_p) * (p - avg_p) * 1) / (told + 1));
  // uu uv uw / uv vv vw / uw vw ww against the averaged products
  const double prod[9] = {u[0] * u[0], u[0] * u[1], u[0] * u[2], u[0] * u[1], u[1] * u[1],
                          u[1] * u[2], u[0] * u[2], u[1] * u[2], u[2] * u[2]};
  for (int q = 0; q < 9; q++) {
    const double r = *P[ST_RMS_RS + q];
    const double d = prod[q] - ars[q];
    *P[ST_RMS_RS + q] = sqrt((r * r * told + d * d * 1) / (told + 1));
  }
}

inline dim3 grid_of(uint32_t n) { return dim3((n + 255) / 256); }

template <class S>
void launch_stress(const lbm_ctx *c, double *sig, double *pres) {
  using C = typename Store<S>::C;
  const LaunchCtx L = launch_ctx(c);
  hipLaunchKernelGGL(k_stress<S>, grid_of(c->g.ncell), dim3(256), 0, c->stream, (const S *)c->cur, c->mask, c->g,
                     (const C *)c->mac[0], (const C *)c->mac[1], (const C *)c->mac[2], (const C *)c->mac[3],
                     L.wall_slots, c->fslot, c->fof, c->fnslot, c->d.tau, sig, pres);
}

// reference order fluid[i][j][k] <- device order, `comp` values per cell
void to_reference(const lbm_ctx *c, const std::vector<double> &h, int comp, double *out) {
  const uint32_t nx = c->g.nx, ny = c->g.ny, nz = c->g.nz, px = c->g.px;
  for (uint32_t i = 0; i < nx; i++)
    for (uint32_t j = 0; j < ny; j++)
      for (uint32_t k = 0; k < nz; k++) {
        const size_t r = ((size_t)i * ny + j) * nz + k, d = ((size_t)k * ny + j) * px + i;
        memcpy(out + r * comp, h.data() + d * comp, comp * sizeof(double));
      }
}

}  // namespace

extern "C" {

int lbm_download_stress(lbm_ctx *c, double *sigma, double *pressure) {
  if (!c || (!sigma && !pressure)) return LBM_ERR_ARG;
  if (c->state != ST_F) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = c->g.ncell;
  double *d = nullptr;
  HIPCHK(hipMalloc((void **)&d, n * 10 * sizeof(double)));
  switch (c->storage) {
    case LBM_STORE_F64: launch_stress<double>(c, d, d + 9 * n); break;
    case LBM_STORE_F32: launch_stress<float>(c, d, d + 9 * n); break;
    case LBM_STORE_F16S: launch_stress<__half>(c, d, d + 9 * n); break;
    default: launch_stress<f16c_t>(c, d, d + 9 * n); break;
  }
  std::vector<double> h(n * 10);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, n * 10 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return LBM_ERR_HIP;
  if (sigma) to_reference(c, std::vector<double>(h.begin(), h.begin() + 9 * n), 9, sigma);
  if (pressure) to_reference(c, std::vector<double>(h.begin() + 9 * n, h.end()), 1, pressure);
  return LBM_OK;
}

int lbm_stats_update(lbm_ctx *c, int32_t n, int32_t nt) {
  if (!c || nt < 1 || n < nt) return LBM_ERR_ARG;
  if (c->state != ST_F) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  const size_t bytes = (size_t)STAT_PLANES * c->g.ncell * sizeof(double);
  if (!c->stats) {
    HIPCHK(hipMalloc((void **)&c->stats, bytes));
    HIPCHK(hipMemsetAsync(c->stats, 0, bytes, c->stream));  // the LATTICE defaults, LBM.h:576-587
  }
  const double told = (double)(n - nt) * 1.0;
  if (c->csize == 8)
    hipLaunchKernelGGL(k_stats<double>, grid_of(c->g.ncell), dim3(256), 0, c->stream, c->g,
                       (const double *)c->mac[0], (const double *)c->mac[1], (const double *)c->mac[2],
                       (const double *)c->mac[3], told, c->stats);
  else
    hipLaunchKernelGGL(k_stats<float>, grid_of(c->g.ncell), dim3(256), 0, c->stream, c->g,
                       (const float *)c->mac[0], (const float *)c->mac[1], (const float *)c->mac[2],
                       (const float *)c->mac[3], told, c->stats);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int lbm_download_stats(lbm_ctx *c, double *avg_u, double *rms_u, double *avg_p, double *rms_p, double *avg_rs,
                       double *rms_rs) {
  if (!c) return LBM_ERR_ARG;
  if (!c->stats) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = c->g.ncell;
  std::vector<double> h((size_t)STAT_PLANES * n);
  HIPCHK(hipMemcpyAsync(h.data(), c->stats, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // planes -> [cell][comp]
  auto gather = [&](int q0, int comp, double *out) {
    if (!out) return;
    std::vector<double> t(n * comp);
    for (size_t d = 0; d < n; d++)
      for (int q = 0; q < comp; q++) t[d * comp + q] = h[(size_t)(q0 + q) * n + d];
    to_reference(c, t, comp, out);
  };
  gather(ST_AVG_U, 3, avg_u);
  gather(ST_RMS_U, 3, rms_u);
  gather(ST_AVG_P, 1, avg_p);
  gather(ST_RMS_P, 1, rms_p);
  gather(ST_AVG_RS, 9, avg_rs);
  gather(ST_RMS_RS, 9, rms_rs);
  return LBM_OK;
}

}  // extern "C"
