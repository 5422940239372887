// lbm_stream.h -- the fused stream kernels (one lattice, ensembles) and
// their launchers, templated on the cell body (FOLD: lattices with folded or
// kept faces).  Included by lbm_kernels.hip, which launches the plain body,
// and by lbm_stream_fold.hip, which launches the FOLD body: the two bodies
// compile in parallel as separate translation units.
#pragma once

#include "lbm_device.h"
#include "lbm_launch.h"
#include "lbm_step.h"

namespace lbm {

// CPT cells per thread, one per grid-stride pass: pass r of the grid covers
// the contiguous cells [r G, (r + 1) G), G = gridDim.x * blockDim.x, so every
// pass keeps the one-contiguous-run-per-plane access of one cell per thread.
// (A run-time pass count makes the compiler duplicate the whole cell body,
// which costs the one-cell kernel instruction-cache room: the count is a
// template argument, and CPT > 1 exists for the fused launch only.)
// COLLIDE: relax and store f*; otherwise store f (post-streaming).
// EMIT   : write rho, u (u = (sum f c + F/2)/rho, lbm.cpp:672-675).
template <class S, int FORCE, bool COLLIDE, bool EMIT, int CPT, bool FOLD>
__device__ __forceinline__ void stream_cells(const Planes<S> &P, const uint8_t *__restrict__ mask,
                                             const KArgs<typename Store<S>::C> &a) {
  if constexpr (CPT == 1) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.g.ncell) return;
    stream_cell<S, FORCE, COLLIDE, EMIT, FOLD>(P, mask, a, idx);
  } else {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t idx0 = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int r = 0; r < CPT; r++) {
      const uint32_t idx = idx0 + (uint32_t)r * stride;
      if (idx < a.g.ncell) stream_cell<S, FORCE, COLLIDE, EMIT, FOLD>(P, mask, a, idx);
    }
  }
}
// FOLD: the cell body for lattices with folded or kept faces (Geo::fold),
// chosen on the host; the two bodies are separate kernels (in one function
// the compiler interleaves them into far more code than either).
template <class S, int FORCE, bool COLLIDE, bool EMIT, int CPT, bool FOLD>
__global__ __launch_bounds__(256) void k_stream(const Planes<S> P, const uint8_t *__restrict__ mask,
                                                KArgs<typename Store<S>::C> a) {
  stream_cells<S, FORCE, COLLIDE, EMIT, CPT, FOLD>(P, mask, a);
}

// ---- ensembles: one launch steps many independent lattices ----------------
// Per lattice and buffer parity: its plane pointers, kernel arguments and mask.
template <class S>
struct LatArgs {
  Planes<S> P;
  KArgs<typename Store<S>::C> a;
  const uint8_t *mask;
};

// Workgroup b takes tiles b*per .. b*per + per - 1 of the tile table; tile t
// is the blockDim.x cells [tile_first[t], tile_first[t] + blockDim.x) of
// lattice tile_lat[t].  per = 1 with one workgroup per tile, or a contiguous
// run of tiles per workgroup for a persistent grid of a few workgroups per CU.
// FOLD: some lattice of the batch has folded or kept faces; the body then
// also serves the lattices without (the host checks their offsets fit).
template <class S, int FORCE, bool COLLIDE, bool EMIT, bool FOLD>
__global__ __launch_bounds__(256) void k_stream_batch(const LatArgs<S> *__restrict__ la,
                                                      const uint32_t *__restrict__ tile_lat,
                                                      const uint32_t *__restrict__ tile_first,
                                                      uint32_t ntile, uint32_t per, int parity) {
  const uint32_t t0 = blockIdx.x * per;
  const uint32_t t1 = min(ntile, t0 + per);
  for (uint32_t t = t0; t < t1; t++) {
    const LatArgs<S> *A = la + 2 * tile_lat[t] + parity;
    const uint32_t idx = tile_first[t] + threadIdx.x;
    if (idx < A->a.g.ncell) stream_cell<S, FORCE, COLLIDE, EMIT, FOLD>(A->P, A->mask, A->a, idx);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 grid_for(uint32_t n, int bs) { return dim3((n + bs - 1) / bs); }

template <class S>
static KArgs<typename Store<S>::C> make_args(const LaunchCtx &L) {
  using C = typename Store<S>::C;
  KArgs<C> a;
  a.g = L.g;
  a.omega = (C)L.omega;
  a.Fx = (C)L.Fu[0];
  a.Fy = (C)L.Fu[1];
  a.Fz = (C)L.Fu[2];
  a.Fxa = (const C *)L.Fx;
  a.Fya = (const C *)L.Fy;
  a.Fza = (const C *)L.Fz;
  a.links = L.links;
  a.uwx = (const C *)L.uwx;
  a.uwy = (const C *)L.uwy;
  a.uwz = (const C *)L.uwz;
  a.rho = (C *)L.rho;
  a.ux = (C *)L.ux;
  a.uy = (C *)L.uy;
  a.uz = (C *)L.uz;
  a.status = L.status;
  a.first_bad = L.first_bad;
  a.box = L.box;
  a.wall_slots = L.wall_slots;
  a.slot_links = L.slot_links;
  return a;
}

template <class S>
static Planes<S> make_planes(const LaunchCtx &L, const S *in, S *out) {
  Planes<S> P;
  const size_t ps = L.g.ps;
  for (int m = 0; m < Q; m++) {
    const ptrdiff_t off = (ptrdiff_t)CX[m] + (ptrdiff_t)CY[m] * L.g.px + (ptrdiff_t)CZ[m] * L.g.plane;
    P.in[m] = in + m * ps;
    P.src[m] = in + (ptrdiff_t)(m * ps) - off;
    P.out[m] = out + m * ps;
  }
  return P;
}

template <class S, int FORCE, bool COLLIDE, bool EMIT, bool FOLD>
static void stream_launch_f(const LaunchCtx &L, const S *in, S *out) {
  auto a = make_args<S>(L);
  const Planes<S> P = make_planes(L, in, out);
  const int cpt = (COLLIDE && !EMIT && (L.cpt == 2 || L.cpt == 4)) ? L.cpt : 1;
  const dim3 grid = grid_for(L.g.ncell, L.block * cpt);
  if constexpr (COLLIDE && !EMIT) {
    if (cpt == 2) {
      hipLaunchKernelGGL((k_stream<S, FORCE, COLLIDE, EMIT, 2, FOLD>), grid, dim3(L.block), 0, L.stream, P, L.mask, a);
      return;
    }
    if (cpt == 4) {
      hipLaunchKernelGGL((k_stream<S, FORCE, COLLIDE, EMIT, 4, FOLD>), grid, dim3(L.block), 0, L.stream, P, L.mask, a);
      return;
    }
  }
  hipLaunchKernelGGL((k_stream<S, FORCE, COLLIDE, EMIT, 1, FOLD>), grid, dim3(L.block), 0, L.stream, P, L.mask, a);
}

template <class S, int FORCE, bool FOLD>
static hipError_t stream_dispatch(const LaunchCtx &L, const void *fin, void *fout, int mode) {
  const S *in = (const S *)fin;
  S *out = (S *)fout;
  switch (mode) {
    case STREAM_ONLY: stream_launch_f<S, FORCE, false, false, FOLD>(L, in, out); break;
    case STREAM_MACRO: stream_launch_f<S, FORCE, false, true, FOLD>(L, in, out); break;
    case STREAM_COLLIDE: stream_launch_f<S, FORCE, true, false, FOLD>(L, in, out); break;
    case STREAM_COLLIDE_MACRO: stream_launch_f<S, FORCE, true, true, FOLD>(L, in, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class S, bool FOLD>
static hipError_t stream_s(const LaunchCtx &L, const void *fin, void *fout, int mode) {
  switch (L.force_mode) {
    case 0: return stream_dispatch<S, 0, FOLD>(L, fin, fout, mode);
    case 1: return stream_dispatch<S, 1, FOLD>(L, fin, fout, mode);
    case 2: return stream_dispatch<S, 2, FOLD>(L, fin, fout, mode);
    case 3: return stream_dispatch<S, 3, FOLD>(L, fin, fout, mode);
  }
  return hipErrorInvalidValue;
}

// Streaming launch of the FOLD or the plain cell body
template <bool FOLD>
static hipError_t stream_body(const LaunchCtx &L, const void *fin, void *fout, int mode) {
  switch (L.storage) {
    case 0: return stream_s<double, FOLD>(L, fin, fout, mode);
    case 1: return stream_s<float, FOLD>(L, fin, fout, mode);
    case 2: return stream_s<__half, FOLD>(L, fin, fout, mode);
    case 3: return stream_s<f16c_t, FOLD>(L, fin, fout, mode);
  }
  return hipErrorInvalidValue;
}

template <class S, int FORCE, bool FOLD>
static void batch_mode_f(const BatchLaunch &B) {
  const auto *la = (const LatArgs<S> *)B.la;
  const dim3 grid(B.grid), block(B.block);
  switch (B.mode) {
    case STREAM_COLLIDE:
      hipLaunchKernelGGL((k_stream_batch<S, FORCE, true, false, FOLD>), grid, block, 0, B.stream, la, B.tile_lat,
                         B.tile_first, B.ntile, B.per, B.parity);
      break;
    case STREAM_COLLIDE_MACRO:
      hipLaunchKernelGGL((k_stream_batch<S, FORCE, true, true, FOLD>), grid, block, 0, B.stream, la, B.tile_lat,
                         B.tile_first, B.ntile, B.per, B.parity);
      break;
    case STREAM_MACRO:
      hipLaunchKernelGGL((k_stream_batch<S, FORCE, false, true, FOLD>), grid, block, 0, B.stream, la, B.tile_lat,
                         B.tile_first, B.ntile, B.per, B.parity);
      break;
    default:
      hipLaunchKernelGGL((k_stream_batch<S, FORCE, false, false, FOLD>), grid, block, 0, B.stream, la, B.tile_lat,
                         B.tile_first, B.ntile, B.per, B.parity);
      break;
  }
}

template <class S, bool FOLD>
static void batch_s(const BatchLaunch &B) {
  switch (B.force_mode) {
    case 0: batch_mode_f<S, 0, FOLD>(B); break;
    case 1: batch_mode_f<S, 1, FOLD>(B); break;
    default: batch_mode_f<S, 2, FOLD>(B); break;
  }
}

// Ensemble launch of the FOLD or the plain cell body
template <bool FOLD>
static hipError_t batch_body(const BatchLaunch &B) {
  switch (B.storage) {
    case 0: batch_s<double, FOLD>(B); break;
    case 1: batch_s<float, FOLD>(B); break;
    case 2: batch_s<__half, FOLD>(B); break;
    default: batch_s<f16c_t, FOLD>(B); break;
  }
  return hipGetLastError();
}

}  // namespace lbm
