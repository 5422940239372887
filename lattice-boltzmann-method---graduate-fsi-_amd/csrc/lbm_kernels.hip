// lbm_kernels.hip -- gfx950 kernels of the D3Q19 BGK step.
//
// The step is HBM-bound: per fluid cell it reads 19 populations and a 1-byte
// mask and writes 19 populations.  One thread per cell of the padded lattice,
// 1D launch over the x-fastest index, so the lanes of a wave touch one
// contiguous run per population plane.  The 19 population loads of a cell
// are issued together with its mask load (see bulk_loads), every load and
// store addresses a host-computed plane base plus one shared 32-bit lane
// offset, and the collision works on +-c direction pairs.
//
//  k_init           LBM::Init (lbm.cpp:31-57), f = feq in every cell
//  k_tables         mask bits, bounce-back link words and wall velocities
//                   (lbm.cpp:186-295, 941-962) from the cell types
//  k_collide        LBM::Collision/ConvLBM (lbm.cpp:60-118), in place
//  k_stream         LBM::Streaming (lbm.cpp:182-295) in pull form, optionally
//                   fused with MacroProp (lbm.cpp:657-776) and the next
//                   step's collision -- the hot kernel
//  k_macro          LBM::MacroProp (lbm.cpp:657-776)
//  k_backeq_solid   the back-face inflow write into solid cells
//  k_links          the streaming tables as the kernels use them
//  k_f_out          populations to FP64 for download
//  k_reduce_*       deterministic FP64 sums (mass, momentum, kinetic energy)
#include <string.h>

#include <type_traits>

#include "lbm_device.h"
#include "lbm_launch.h"
#include "lbm_step.h"
#include "lbm_stream.h"

namespace lbm {

// LBM::Collision (lbm.cpp:60-73) in place, with rho, u and F as stored
// (modes 0-2).  In mode 3 (the IBM step) rho, u and F come from the cell's
// own populations and the IBM box (cell_state), and rho, u are emitted: they
// are what the reference holds in fluid[][][] after IBLBM.
template <class S, int FORCE>
__global__ __launch_bounds__(256) void k_collide(S *__restrict__ buf, const uint8_t *__restrict__ mask,
                                                 KArgs<typename Store<S>::C> a) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  const uint8_t mk = mask[idx];
  if (!(mk & M_FLUID)) return;
  const uint32_t ps = a.g.ps;
  C f[Q];
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    f[m] = Store<S>::ld(buf + (uint32_t)m * ps + idx);
  });
  C rho, drho, ux, uy, uz, Fx, Fy, Fz;
  if constexpr (FORCE == 3) {
    cell_state<S, FORCE>(f, mk, idx, a, rho, drho, ux, uy, uz, Fx, Fy, Fz);
    a.rho[idx] = rho;
    a.ux[idx] = ux;
    a.uy[idx] = uy;
    a.uz[idx] = uz;
  } else {
    rho = a.rho[idx];
    drho = rho - C(1);
    ux = a.ux[idx];
    uy = a.uy[idx];
    uz = a.uz[idx];
    load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
  }
  const bool neg = relax<S, FORCE>(f, rho, drho, ux, uy, uz, Fx, Fy, Fz, a.omega);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    Store<S>::st(buf + (uint32_t)m * ps + idx, f[m]);
  });
  if (a.wall_slots && (mk & M_WALL)) {
    S *planes[Q];
    for (int m = 0; m < Q; m++) planes[m] = buf + (size_t)m * ps;
    wall_slot_stores<S>(f, mk, idx, a.links[idx], a, planes);
  }
  if (neg) flag(a.status, a.first_bad, ST_FEQ_NEG, idx);
}

// LBM::MacroProp (lbm.cpp:657-776) on the cells whose mask has a bit of
// `sel` (TYPE_F and TYPE_S cells).  Cells of other types keep their stored
// rho,u: their population slots carry the frozen-link values of their fluid
// neighbours.
template <class S, int FORCE>
__global__ __launch_bounds__(256) void k_macro(const S *__restrict__ buf,
                                               const uint8_t *__restrict__ mask, uint8_t sel,
                                               KArgs<typename Store<S>::C> a) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  if (!(mask[idx] & sel)) return;
  const uint32_t ps = a.g.ps;
  C f[Q];
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    f[m] = Store<S>::ld(buf + (uint32_t)m * ps + idx);
  });
  C rho, drho, jx, jy, jz;
  moments<S>(f, rho, drho, jx, jy, jz);
  C Fx, Fy, Fz;
  load_force<C, (FORCE == 3 ? 0 : FORCE)>(a, idx, Fx, Fy, Fz);
  a.rho[idx] = rho;
  C ux, uy, uz;
  velocity(jx + C(0.5) * Fx, jy + C(0.5) * Fy, jz + C(0.5) * Fz, rho, ux, uy, uz);
  a.ux[idx] = ux;
  a.uy[idx] = uy;
  a.uz[idx] = uz;
  if (rho < C(0) || rho > C(100)) flag(a.status, a.first_bad, ST_RHO_RANGE, idx);
}

// LBM::Init (lbm.cpp:31-57): f = feq(rho,u) in every cell, written to both
// buffers.  A cell that never collides (not TYPE_F or TYPE_S) instead stores
// in slot m the Init value of its fluid neighbour x + c_m in direction m:
// the reference never rewrites that population (frozen link).
template <class S>
__global__ __launch_bounds__(256) void k_init(S *__restrict__ A, S *__restrict__ B,
                                              const uint8_t *__restrict__ mask,
                                              KArgs<typename Store<S>::C> a) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  const uint8_t mk = mask[idx];
  uint32_t x, y, z;
  coords(a.g, idx, x, y, z);
  const uint32_t ps = a.g.ps;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    uint32_t c = idx;
    if (mk & M_OTHER) {
      const uint32_t d = wrap_dst(a.g, x, y, z, CX[m], CY[m], CZ[m]);
      if (mask[d] & M_FLUID) c = d;
    }
    const C rho = a.rho[c], ux = a.ux[c], uy = a.uy[c], uz = a.uz[c];
    const C uu = ux * ux + uy * uy + uz * uz;
    const C cu = C(CX[m]) * ux + C(CY[m]) * uy + C(CZ[m]) * uz;
    const C A_ = eq_A(cu, uu);
    C r;
    if constexpr (Store<S>::shifted) r = C(Store<S>::K * W[m]) * ((rho - C(1)) + rho * A_);
    else r = rho * C(W[m]) * (C(1) + A_);
    const S s = Store<S>::enc(r);
    A[(uint32_t)m * ps + idx] = s;
    B[(uint32_t)m * ps + idx] = s;
  });
}

// Per-cell kernel mask, bounce-back link word and wall velocity from the
// cell types (device order; row-padding cells get mask 0) and the solids'
// current velocities.
//
// With wall slots (line_cells > 0, the cells of one 128-byte line of a
// population plane), solid and row-padding cells get a fill word instead of
// link bits: the slots that no wall cell writes, set only where the cell's
// line also holds a fluid cell.  The stream kernel writes those slots (with
// values nobody reads), so that the lines a row of fluid cells stores are
// written whole rather than with holes at the solids and the row padding.
template <class C>
__global__ __launch_bounds__(256) void k_tables(const int16_t *__restrict__ type,
                                                uint8_t *__restrict__ mask,
                                                uint32_t *__restrict__ links, C *uwx, C *uwy,
                                                C *uwz, Geo g, const C *__restrict__ ux,
                                                const C *__restrict__ uy,
                                                const C *__restrict__ uz, uint32_t line_cells) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell) return;
  uint32_t x, y, z;
  coords(g, idx, x, y, z);
  uint8_t mk = 0;
  uint32_t lk = 0;
  C wx = C(0), wy = C(0), wz = C(0);
  if (line_cells && (x >= g.nx || type[idx] == 1)) {
    bool line_fluid = false;
    const uint32_t j0 = idx - idx % line_cells;
    for (uint32_t j = j0; j < j0 + line_cells && j < g.ncell; j++)
      if (j % g.px < g.nx && type[j] == 0) line_fluid = true;
    if (line_fluid) {
      lk = (1u << Q) - 1;
      // slot m of a solid is written by its neighbour x + c_m when that is
      // a fluid cell (wall_slot_stores), unless the solid's face is folded
      // (its wall links are resolved in the pull address)
      if (x < g.nx && !on_fold_face(g, x, y, z)) {
        for_dirs([&](auto M) {
          constexpr int m = decltype(M)::value;
          if (type[wrap_dst(g, x, y, z, CX[m], CY[m], CZ[m])] == 0) lk &= ~(1u << m);
        });
      }
    }
  }
  if (x < g.nx) {
    const int16_t t = type[idx];
    if (t == 0) mk |= M_FLUID;
    else if (t == 1) mk |= M_SOLID;
    else mk |= M_OTHER;
    if (x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1)
      mk |= M_WRAP;
    if (z == g.nz - 1 && type[idx - (g.nz - 1) * g.plane] == 3) mk |= M_BACKEQ;
    if (t == 0) {
      // bounce-back links: source x - c_m is solid; the link word keeps
      // those not folded into the pull address (fold_mask)
      uint32_t all = 0;
      for_dirs([&](auto M) {
        constexpr int m = decltype(M)::value;
        if (type[wrap_src(g, x, y, z, CX[m], CY[m], CZ[m])] == 1) all |= 1u << m;
      });
      lk = all & ~fold_mask(g, x, y, z);
      // Determine_Wall_Velocity, lbm.cpp:941-962: neighbours x + c_l in l order
      int num = 0;
      for_dirs([&](auto M) {
        constexpr int l = decltype(M)::value;
        const uint32_t n = wrap_dst(g, x, y, z, CX[l], CY[l], CZ[l]);
        if (type[n] == 1 && (ux[n] != C(0) || uy[n] != C(0) || uz[n] != C(0))) {
          wx += ux[n];
          wy += uy[n];
          wz += uz[n];
          num++;
        }
      });
      if (num != 0) {
        wx /= C(num);
        wy /= C(num);
        wz /= C(num);
      }
      if (lk) mk |= M_WALL;
      if (all && (wx != C(0) || wy != C(0) || wz != C(0))) mk |= M_MOVING;
    }
  }
  mask[idx] = mk;
  if (links) links[idx] = lk;
  if (uwx) {
    uwx[idx] = wx;
    uwy[idx] = wy;
    uwz[idx] = wz;
  }
}

// TYPE_S cells hit by the back-face inflow write (M_BACKEQ): at every
// Streaming they take feq2 of their TYPE_IN partner's rho, u as of the last
// MacroProp (lbm.cpp:599-611).  With wall slots a solid's buffer entries are
// bounce-back slots that the face loops may still read (an outlet copying a
// fluid neighbour's pulled value), so only the solid's MacroProp values
// (moments of the stored populations, lbm.cpp:668-675) are written; they
// are what the wall tables and the output read.  Without wall slots the
// populations themselves are written to both buffers.
template <class S, int FORCE>
__global__ __launch_bounds__(256) void k_backeq_solid(S *__restrict__ A, S *__restrict__ B,
                                                      const uint8_t *__restrict__ mask,
                                                      KArgs<typename Store<S>::C> a) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  const uint8_t mk = mask[idx];
  if ((mk & (M_SOLID | M_BACKEQ)) != (M_SOLID | M_BACKEQ)) return;
  const uint32_t s = idx - (a.g.nz - 1) * a.g.plane;
  const C r = a.rho[s], vx = a.ux[s], vy = a.uy[s], vz = a.uz[s];
  const C uu = vx * vx + vy * vy + vz * vz;
  const uint32_t ps = a.g.ps;
  C f[Q];
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    const C cu = C(CX[m]) * vx + C(CY[m]) * vy + C(CZ[m]) * vz;
    const S v = Store<S>::enc(feq2_raw<C, Store<S>::shifted>(r, cu, uu, C(W[m]), C(Store<S>::K)));
    f[m] = Store<S>::ld(&v);  // the stored value
    if (!a.wall_slots) {
      A[(uint32_t)m * ps + idx] = v;
      B[(uint32_t)m * ps + idx] = v;
    }
  });
  if (a.wall_slots) {
    C rho, drho, jx, jy, jz, Fx, Fy, Fz;
    moments<S>(f, rho, drho, jx, jy, jz);
    load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
    a.rho[idx] = rho;
    C ux, uy, uz;
    velocity(jx + C(0.5) * Fx, jy + C(0.5) * Fy, jz + C(0.5) * Fz, rho, ux, uy, uz);
    a.ux[idx] = ux;
    a.uy[idx] = uy;
    a.uz[idx] = uz;
  }
}

// The streaming tables exactly as the stream kernels resolve them.
__global__ __launch_bounds__(256) void k_links(const uint8_t *__restrict__ mask,
                                               const uint32_t *__restrict__ links, Geo g,
                                               int32_t *__restrict__ src_out,
                                               uint8_t *__restrict__ kind_out) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell) return;
  const uint8_t mk = mask[idx];
  if (!(mk & M_FLUID)) {
    for (int m = 0; m < Q; m++) {
      src_out[(size_t)idx * Q + m] = -1;
      kind_out[(size_t)idx * Q + m] = 255;
    }
    return;
  }
  uint32_t x, y, z;
  coords(g, idx, x, y, z);
  const uint32_t lk = ((mk & M_WALL) ? links[idx] : 0u) | fold_mask(g, x, y, z);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    const uint32_t src = (mk & M_WRAP) ? wrap_src(g, x, y, z, CX[m], CY[m], CZ[m])
                                       : idx - pull_off<m>(g);
    uint8_t kind;
    if (lk & (1u << m)) kind = 1;
    else kind = (mask[src] & M_FLUID) ? 0 : 2;
    src_out[(size_t)idx * Q + m] = (int32_t)src;
    kind_out[(size_t)idx * Q + m] = kind;
  });
}

// Populations as FP64 f (unshifted), [idx][m].  TYPE_F cells (and TYPE_S
// cells unless their slots carry bounce-back values) come from the buffer;
// other cells report feq of their stored rho,u (their Init value; their
// slots hold values for their neighbours).
template <class S>
__global__ __launch_bounds__(256) void k_f_out(const S *__restrict__ buf,
                                               const uint8_t *__restrict__ mask,
                                               KArgs<typename Store<S>::C> a,
                                               double *__restrict__ out) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  const uint8_t mk = mask[idx];
  const uint32_t ps = a.g.ps;
  const C rho = a.rho[idx], ux = a.ux[idx], uy = a.uy[idx], uz = a.uz[idx];
  const C uu = ux * ux + uy * uy + uz * uz;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    double v;
    if ((mk & M_FLUID) || ((mk & M_SOLID) && !a.wall_slots)) {
      v = (double)Store<S>::ld(buf + (uint32_t)m * ps + idx);
      if constexpr (Store<S>::shifted) v = v / (double)Store<S>::K + W[m];
    } else {
      const C cu = C(CX[m]) * ux + C(CY[m]) * uy + C(CZ[m]) * uz;
      v = (double)(rho * C(W[m]) * (C(1) + eq_A(cu, uu)));
    }
    out[(size_t)idx * Q + m] = v;
  });
}

// ---- deterministic FP64 reductions over TYPE_F cells ----------------------
// Pass 1: block b sums cells [b*chunk, (b+1)*chunk) in a fixed tree order.
// Pass 2: one block sums the partials in a fixed tree order.
constexpr int RED_THREADS = 256;

__device__ __forceinline__ double block_sum(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

template <class C>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_1(const uint8_t *__restrict__ mask,
                                                          const C *__restrict__ rho,
                                                          const C *__restrict__ ux,
                                                          const C *__restrict__ uy,
                                                          const C *__restrict__ uz, uint32_t n,
                                                          uint32_t chunk, int what,
                                                          double *__restrict__ partial) {
  __shared__ double sh[RED_THREADS];
  const uint32_t b0 = blockIdx.x * chunk;
  const uint32_t b1 = min(n, b0 + chunk);
  double s[3] = {0.0, 0.0, 0.0};
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += RED_THREADS) {
    if (!(mask[i] & M_FLUID)) continue;
    const double r = (double)rho[i];
    if (what == 0) {
      s[0] += r;
    } else if (what == 1) {
      s[0] += r * (double)ux[i];
      s[1] += r * (double)uy[i];
      s[2] += r * (double)uz[i];
    } else {
      const double x = ux[i], y = uy[i], z = uz[i];
      s[0] += 0.5 * r * (x * x + y * y + z * z);
    }
  }
  const int nc = (what == 1) ? 3 : 1;
  for (int c = 0; c < nc; c++) {
    const double t = block_sum(s[c], sh);
    if (threadIdx.x == 0) partial[(size_t)c * gridDim.x + blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_2(const double *__restrict__ partial,
                                                          int nblocks, int ncomp,
                                                          double *__restrict__ out) {
  __shared__ double sh[RED_THREADS];
  for (int c = 0; c < ncomp; c++) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += RED_THREADS) s += partial[(size_t)c * nblocks + i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) out[c] = t;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class S, int FORCE>
static hipError_t collide_dispatch(const LaunchCtx &L, void *buf, bool macro_only) {
  auto a = make_args<S>(L);
  const dim3 grid = grid_for(L.g.ncell, L.block);
  if (macro_only)
    hipLaunchKernelGGL((k_macro<S, FORCE>), grid, dim3(L.block), 0, L.stream, (const S *)buf, L.mask, L.macro_sel, a);
  else
    hipLaunchKernelGGL((k_collide<S, FORCE>), grid, dim3(L.block), 0, L.stream, (S *)buf, L.mask, a);
  return hipGetLastError();
}

template <class S>
static hipError_t collide_s(const LaunchCtx &L, void *buf, bool macro_only) {
  switch (L.force_mode) {
    case 0: return collide_dispatch<S, 0>(L, buf, macro_only);
    case 1: return collide_dispatch<S, 1>(L, buf, macro_only);
    case 2: return collide_dispatch<S, 2>(L, buf, macro_only);
    case 3: return collide_dispatch<S, 3>(L, buf, macro_only);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_collide(const LaunchCtx &L, void *buf) {
  switch (L.storage) {
    case 0: return collide_s<double>(L, buf, false);
    case 1: return collide_s<float>(L, buf, false);
    case 2: return collide_s<__half>(L, buf, false);
    case 3: return collide_s<f16c_t>(L, buf, false);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_macro(const LaunchCtx &L, const void *buf) {
  switch (L.storage) {
    case 0: return collide_s<double>(L, (void *)buf, true);
    case 1: return collide_s<float>(L, (void *)buf, true);
    case 2: return collide_s<__half>(L, (void *)buf, true);
    case 3: return collide_s<f16c_t>(L, (void *)buf, true);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_init(const LaunchCtx &L, void *A, void *B) {
  const dim3 grid = grid_for(L.g.ncell, L.block);
  switch (L.storage) {
    case 0:
      hipLaunchKernelGGL(k_init<double>, grid, dim3(L.block), 0, L.stream, (double *)A, (double *)B, L.mask, make_args<double>(L));
      break;
    case 1:
      hipLaunchKernelGGL(k_init<float>, grid, dim3(L.block), 0, L.stream, (float *)A, (float *)B, L.mask, make_args<float>(L));
      break;
    case 2:
      hipLaunchKernelGGL(k_init<__half>, grid, dim3(L.block), 0, L.stream, (__half *)A, (__half *)B, L.mask, make_args<__half>(L));
      break;
    case 3:
      hipLaunchKernelGGL(k_init<f16c_t>, grid, dim3(L.block), 0, L.stream, (f16c_t *)A, (f16c_t *)B, L.mask, make_args<f16c_t>(L));
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_tables(const LaunchCtx &L, uint8_t *mask, uint32_t *links, void *uwx,
                         void *uwy, void *uwz, const int16_t *type) {
  const dim3 grid = grid_for(L.g.ncell, L.block);
  const uint32_t line_cells = L.wall_slots ? 128u / (L.storage == 0 ? 8u : L.storage == 1 ? 4u : 2u) : 0u;
  if (L.storage == 0)
    hipLaunchKernelGGL(k_tables<double>, grid, dim3(L.block), 0, L.stream, type, mask, links,
                       (double *)uwx, (double *)uwy, (double *)uwz, L.g, (const double *)L.ux,
                       (const double *)L.uy, (const double *)L.uz, line_cells);
  else
    hipLaunchKernelGGL(k_tables<float>, grid, dim3(L.block), 0, L.stream, type, mask, links,
                       (float *)uwx, (float *)uwy, (float *)uwz, L.g, (const float *)L.ux,
                       (const float *)L.uy, (const float *)L.uz, line_cells);
  return hipGetLastError();
}

template <class S>
static void backeq_s(const LaunchCtx &L, void *A, void *B) {
  const dim3 grid = grid_for(L.g.ncell, L.block);
  auto a = make_args<S>(L);
  switch (L.force_mode) {
    case 1: hipLaunchKernelGGL((k_backeq_solid<S, 1>), grid, dim3(L.block), 0, L.stream, (S *)A, (S *)B, L.mask, a); break;
    case 2: hipLaunchKernelGGL((k_backeq_solid<S, 2>), grid, dim3(L.block), 0, L.stream, (S *)A, (S *)B, L.mask, a); break;
    default: hipLaunchKernelGGL((k_backeq_solid<S, 0>), grid, dim3(L.block), 0, L.stream, (S *)A, (S *)B, L.mask, a); break;
  }
}

hipError_t launch_backeq_solid(const LaunchCtx &L, void *A, void *B) {
  switch (L.storage) {
    case 0: backeq_s<double>(L, A, B); break;
    case 1: backeq_s<float>(L, A, B); break;
    case 2: backeq_s<__half>(L, A, B); break;
    case 3: backeq_s<f16c_t>(L, A, B); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_links(const LaunchCtx &L, int32_t *src, uint8_t *kind) {
  hipLaunchKernelGGL(k_links, grid_for(L.g.ncell, L.block), dim3(L.block), 0, L.stream, L.mask, L.links, L.g, src, kind);
  return hipGetLastError();
}

hipError_t launch_f_out(const LaunchCtx &L, const void *buf, double *out) {
  const dim3 grid = grid_for(L.g.ncell, L.block);
  switch (L.storage) {
    case 0:
      hipLaunchKernelGGL(k_f_out<double>, grid, dim3(L.block), 0, L.stream, (const double *)buf, L.mask, make_args<double>(L), out);
      break;
    case 1:
      hipLaunchKernelGGL(k_f_out<float>, grid, dim3(L.block), 0, L.stream, (const float *)buf, L.mask, make_args<float>(L), out);
      break;
    case 2:
      hipLaunchKernelGGL(k_f_out<__half>, grid, dim3(L.block), 0, L.stream, (const __half *)buf, L.mask, make_args<__half>(L), out);
      break;
    case 3:
      hipLaunchKernelGGL(k_f_out<f16c_t>, grid, dim3(L.block), 0, L.stream, (const f16c_t *)buf, L.mask, make_args<f16c_t>(L), out);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_stream(const LaunchCtx &L, const void *fin, void *fout, int mode) {
  if (L.g.fold) return launch_stream_fold(L, fin, fout, mode);
  return stream_body<false>(L, fin, fout, mode);
}

size_t batch_args_size(int storage) {
  switch (storage) {
    case 0: return sizeof(LatArgs<double>);
    case 1: return sizeof(LatArgs<float>);
    case 2: return sizeof(LatArgs<__half>);
    default: return sizeof(LatArgs<f16c_t>);
  }
}

template <class S>
static void fill_args_s(const LaunchCtx &L, const void *in, void *out, void *dst) {
  LatArgs<S> A;
  A.P = make_planes(L, (const S *)in, (S *)out);
  A.a = make_args<S>(L);
  A.mask = L.mask;
  memcpy(dst, &A, sizeof A);
}

void batch_fill_args(const LaunchCtx &L, const void *in, void *out, void *dst) {
  switch (L.storage) {
    case 0: fill_args_s<double>(L, in, out, dst); break;
    case 1: fill_args_s<float>(L, in, out, dst); break;
    case 2: fill_args_s<__half>(L, in, out, dst); break;
    default: fill_args_s<f16c_t>(L, in, out, dst); break;
  }
}

hipError_t launch_stream_batch(const BatchLaunch &B) {
  if (B.fold) return launch_stream_batch_fold(B);
  return batch_body<false>(B);
}

int reduce_blocks(uint32_t n) {
  // fixed partition independent of the device: 4096 cells per thread block
  return (int)((n + 4095) / 4096);
}

hipError_t launch_reduce(const LaunchCtx &L, int what, double *partial, double *out) {
  const uint32_t n = L.g.ncell;
  const int nb = reduce_blocks(n);
  if (L.storage == 0)
    hipLaunchKernelGGL(k_reduce_1<double>, dim3(nb), dim3(RED_THREADS), 0, L.stream, L.mask, (const double *)L.rho,
                       (const double *)L.ux, (const double *)L.uy, (const double *)L.uz, n, 4096u, what, partial);
  else
    hipLaunchKernelGGL(k_reduce_1<float>, dim3(nb), dim3(RED_THREADS), 0, L.stream, L.mask, (const float *)L.rho,
                       (const float *)L.ux, (const float *)L.uy, (const float *)L.uz, n, 4096u, what, partial);
  hipLaunchKernelGGL(k_reduce_2, dim3(1), dim3(RED_THREADS), 0, L.stream, partial, nb, what == 1 ? 3 : 1, out);
  return hipGetLastError();
}

}  // namespace lbm
