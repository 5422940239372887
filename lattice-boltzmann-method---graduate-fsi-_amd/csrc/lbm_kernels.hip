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
#include <type_traits>

#include "lbm_device.h"
#include "lbm_launch.h"

namespace lbm {

template <class C>
struct KArgs {
  Geo g;
  C omega;
  C Fx, Fy, Fz;                  // uniform force (force mode 2)
  const C *Fxa, *Fya, *Fza;      // per-cell force (force mode 1)
  const uint32_t *links;         // bit m: x - c_m is TYPE_S
  const C *uwx, *uwy, *uwz;      // per-cell wall velocity (lbm.cpp:941-962)
  C *rho, *ux, *uy, *uz;         // macroscopic arrays
  int *status;
  unsigned long long *first_bad;
};

// Per-direction plane pointers, computed on the host for each launch so
// that no wave recomputes them: src[m] = fin + m*ps - off_m (bulk pull
// source), in[m] = fin + m*ps, out[m] = fout + m*ps.
template <class S>
struct Planes {
  const S *src[Q];
  const S *in[Q];
  S *out[Q];
};

template <class C, int FORCE>
__device__ __forceinline__ void load_force(const KArgs<C> &a, uint32_t idx, C &Fx, C &Fy, C &Fz) {
  if constexpr (FORCE == 1) {
    Fx = a.Fxa[idx];
    Fy = a.Fya[idx];
    Fz = a.Fza[idx];
  } else if constexpr (FORCE == 2) {
    Fx = a.Fx;
    Fy = a.Fy;
    Fz = a.Fz;
  } else {
    Fx = C(0);
    Fy = C(0);
    Fz = C(0);
  }
}

__device__ __forceinline__ void flag(int *status, unsigned long long *first_bad, int bits,
                                     uint32_t idx) {
  atomicOr(status, bits);
  atomicMin(first_bad, (unsigned long long)idx);
}

__device__ __forceinline__ float vmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double vmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return __builtin_fmax(a, b); }

// u = (j + F/2)/rho: FP64 divides each component (the reference's
// expression); FP32 multiplies by one correctly rounded reciprocal.
template <class C>
__device__ __forceinline__ void velocity(C jx, C jy, C jz, C rho, C &ux, C &uy, C &uz) {
  if constexpr (sizeof(C) == 8) {
    ux = jx / rho;
    uy = jy / rho;
    uz = jz / rho;
  } else {
    const C ir = C(1) / rho;
    ux = jx * ir;
    uy = jy * ir;
    uz = jz * ir;
  }
}

// Moments of one cell from raw values (MacroProp, lbm.cpp:668-671), summed
// pairwise over the +-c direction pairs.
template <class S>
__device__ __forceinline__ void moments(const typename Store<S>::C (&f)[Q],
                                        typename Store<S>::C &rho, typename Store<S>::C &drho,
                                        typename Store<S>::C &jx, typename Store<S>::C &jy,
                                        typename Store<S>::C &jz) {
  using C = typename Store<S>::C;
  // p_k = f_m + f_opp, d_k = f_m - f_opp for the pairs (1,2), (3,4), ... (17,18)
  const C p1 = f[1] + f[2], p2 = f[3] + f[4], p3 = f[5] + f[6], p4 = f[7] + f[8];
  const C p5 = f[9] + f[10], p6 = f[11] + f[12], p7 = f[13] + f[14], p8 = f[15] + f[16];
  const C p9 = f[17] + f[18];
  const C d1 = f[1] - f[2], d2 = f[3] - f[4], d3 = f[5] - f[6], d4 = f[7] - f[8];
  const C d5 = f[9] - f[10], d6 = f[11] - f[12], d7 = f[13] - f[14], d8 = f[15] - f[16];
  const C d9 = f[17] - f[18];
  const C s0 = ((f[0] + (p1 + p2)) + ((p3 + p4) + (p5 + p6))) + ((p7 + p8) + p9);
  // c(7) = (1,1,0), c(9) = (1,0,1), c(11) = (0,1,1), c(13) = (1,-1,0),
  // c(15) = (1,0,-1), c(17) = (0,1,-1)
  jx = (d1 + d4) + ((d5 + d7) + d8);
  jy = (d2 + d4) + ((d6 - d7) + d9);
  jz = (d3 + d5) + ((d6 - d8) - d9);
  if constexpr (Store<S>::shifted) {
    constexpr C iK = C(1) / C(Store<S>::K);
    drho = s0 * iK;
    rho = C(1) + drho;
    jx *= iK;
    jy *= iK;
    jz *= iK;
  } else {
    rho = s0;
    drho = s0 - C(1);
  }
}

// BGK collision with the Guo source (ConvLBM, lbm.cpp:75-118) of one cell,
// in place on raw values.  Returns true when some feq < 0 (lbm.cpp:114).
// Directions come in pairs m, m+1 = opp(m) with c_{m+1} = -c_m, so
// A(-cu) = s - p and A(cu) = s + p share s = 4.5cu^2 - 1.5uu and
// p = cu (3 + 4.5cu^2 - 4.5uu).  f(1-omega) + omega feq is evaluated as
// feq + (1-omega)(f - feq), which rounds the large term once.
template <class S, int FORCE>
__device__ __forceinline__ bool relax(typename Store<S>::C (&f)[Q], typename Store<S>::C rho,
                                      typename Store<S>::C drho, typename Store<S>::C ux,
                                      typename Store<S>::C uy, typename Store<S>::C uz,
                                      typename Store<S>::C Fx, typename Store<S>::C Fy,
                                      typename Store<S>::C Fz, typename Store<S>::C omega) {
  using C = typename Store<S>::C;
  constexpr C K = C(Store<S>::K);
  const C uu = ux * ux + uy * uy + uz * uz;
  const C om1 = C(1) - omega;
  const C c3 = C(3) - C(4.5) * uu;  // 3 - 4.5uu
  const C mh = C(-1.5) * uu;        // -1.5uu
  // raw equilibrium feq = Kw*(drho + rho*A) (shifted) or w*rho*(1 + A)
  // (unshifted) = ra*A + da per weight class; relaxation in the
  // non-equilibrium form f* = feq + (1 - omega)(f - feq)
  C ra[3], da[3];
  {
    constexpr C wc[3] = {C(1.0 / 3), C(1.0 / 18), C(1.0 / 36)};
#pragma unroll
    for (int q = 0; q < 3; q++) {
      ra[q] = (K * wc[q]) * rho;
      da[q] = Store<S>::shifted ? (K * wc[q]) * drho : ra[q];
    }
  }
  C fpre = C(0), Fu = C(0);
  if constexpr (FORCE != 0) {
    fpre = C(1) - omega * C(0.5);
    Fu = Fx * ux + Fy * uy + Fz * uz;
  }
  C amin = mh, amax = mh;  // A of the rest direction is -1.5uu
  // rest direction
  {
    const C fe = ra[0] * mh + da[0];
    f[0] = fe + om1 * (f[0] - fe);
  }
  if constexpr (FORCE != 0) f[0] += (K * C(1.0 / 3)) * fpre * (C(-3) * Fu);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    if constexpr (m % 2 == 1) {
      constexpr int q = (m <= 6) ? 1 : 2;
      const C cu = C(CX[m]) * ux + C(CY[m]) * uy + C(CZ[m]) * uz;
      const C t2 = cu * cu;
      const C s_ = C(4.5) * t2 + mh;
      const C p_ = cu * (C(4.5) * t2 + c3);
      const C Ap = s_ + p_, Am = s_ - p_;
      amin = vmin(amin, vmin(Ap, Am));
      amax = vmax(amax, vmax(Ap, Am));
      const C fep = ra[q] * Ap + da[q], fem = ra[q] * Am + da[q];
      C fp = fep + om1 * (f[m] - fep);
      C fm = fem + om1 * (f[m + 1] - fem);
      if constexpr (FORCE != 0) {
        constexpr C Kw = C(Store<S>::K * W[m]);
        const C Fc = C(CX[m]) * Fx + C(CY[m]) * Fy + C(CZ[m]) * Fz;
        const C kf = Kw * fpre;
        // S(c) = Kw fpre (3Fc + 9 Fc cu - 3Fu); S(-c) has Fc, cu negated
        fp += kf * (Fc * (C(3) + C(9) * cu) - C(3) * Fu);
        fm += kf * (-Fc * (C(3) - C(9) * cu) - C(3) * Fu);
      }
      f[m] = fp;
      f[m + 1] = fm;
    }
  });
  // feq = rho w (1 + A) < 0 for some direction (lbm.cpp:114-117)
  return (rho > C(0) && C(1) + amin < C(0)) || (rho < C(0) && C(1) + amax > C(0));
}

// ---------------------------------------------------------------------------
// Pull of one fluid cell.  The value arriving at cell x in direction m is,
// as in the push loop of lbm.cpp:252-292:
//   x - c_m is TYPE_F      : f*(x - c_m, m)
//   x - c_m is TYPE_S      : f*(x, opp m) + 2 w_m (u_w . c_m) / cs2
//   x - c_m any other type : f(x, m) untouched since Init; that value is kept
//                            in slot m of the source cell, which never
//                            collides, so the same load reads it.
// ---------------------------------------------------------------------------
template <class S>
__device__ __forceinline__ const S *at(const S *base, uint32_t boff) {
  return reinterpret_cast<const S *>(reinterpret_cast<const char *>(base) + boff);
}
template <class S>
__device__ __forceinline__ S *at(S *base, uint32_t boff) {
  return reinterpret_cast<S *>(reinterpret_cast<char *>(base) + boff);
}

// Addressing: a wave-uniform 64-bit base per direction plus one 32-bit
// per-lane byte offset, so every load uses the scalar-base form
// (global_load saddr) with no per-access 64-bit vector arithmetic.
//
// Bulk loads f[m] = src[m][idx] are issued for every thread before its mask
// is known (they overlap the mask load instead of waiting a full memory
// round-trip for it).  They never leave the population allocation: for any
// idx < ncell, m*ps + idx - off_m lies in [0, 19*ps) because |off_m| <=
// plane + px + 1 < ps.  Cells that are not plain bulk fluid then patch the
// directions whose source differs (patch_cell).
template <class S>
__device__ __forceinline__ void bulk_loads(const Planes<S> &P, uint32_t idx,
                                           typename Store<S>::C (&f)[Q]) {
  const uint32_t boff = idx * (uint32_t)sizeof(S);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    f[m] = Store<S>::ld(at(P.src[m], boff));
  });
}

// An empty asm statement that takes the bulk-loaded values as inputs and
// passes the mask byte through: the mask test then depends on every load
// having completed, so neither the loads nor the test can be moved across
// each other (the compiler would otherwise sink the loads into the fluid
// branch, or hoist the test above them, and serialise two memory
// round-trips per wave).  It emits no instruction.
template <class C>
__device__ __forceinline__ uint32_t after_loads(uint32_t mk, const C (&f)[Q]) {
  asm volatile(""
               : "+v"(mk)
               : "v"(f[0]), "v"(f[1]), "v"(f[2]), "v"(f[3]), "v"(f[4]), "v"(f[5]), "v"(f[6]),
                 "v"(f[7]), "v"(f[8]), "v"(f[9]), "v"(f[10]), "v"(f[11]), "v"(f[12]),
                 "v"(f[13]), "v"(f[14]), "v"(f[15]), "v"(f[16]), "v"(f[17]), "v"(f[18]));
  return mk;
}

template <class S>
__device__ __forceinline__ void patch_cell(const Planes<S> &P, uint8_t mk, uint32_t idx,
                                           const KArgs<typename Store<S>::C> &a,
                                           typename Store<S>::C (&f)[Q]) {
  using C = typename Store<S>::C;
  const uint32_t boff = idx * (uint32_t)sizeof(S);
  if (mk & M_WRAP) {
    // periodic wrap of the padded lattice (lbm.cpp:197-199)
    uint32_t x, y, z;
    coords(a.g, idx, x, y, z);
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      const uint32_t src = wrap_src(a.g, x, y, z, CX[m], CY[m], CZ[m]);
      if (src != idx - pull_off<m>(a.g)) f[m] = Store<S>::ld(at(P.in[m], src * (uint32_t)sizeof(S)));
    });
  }
  if (mk & M_WALL) {
    const uint32_t lk = a.links[idx];
    C uwx = C(0), uwy = C(0), uwz = C(0);
    if (mk & M_MOVING) {
      uwx = a.uwx[idx];
      uwy = a.uwy[idx];
      uwz = a.uwz[idx];
    }
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      if (lk & (1u << m)) {
        const C ucw = C(CX[m]) * uwx + C(CY[m]) * uwy + C(CZ[m]) * uwz;
        f[m] = Store<S>::ld(at(P.in[OPP[m]], boff)) + C(Store<S>::K * 6.0 * W[m]) * ucw;
      }
    });
  }
  if (mk & M_BACKEQ) {
    // after the push, the back-face inflow loop overwrites this cell with
    // feq2 of the TYPE_IN cell (i,j,0) (lbm.cpp:599-611)
    const uint32_t s = idx - (a.g.nz - 1) * a.g.plane;
    const C r = a.rho[s], vx = a.ux[s], vy = a.uy[s], vz = a.uz[s];
    const C uu = vx * vx + vy * vy + vz * vz;
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      const C cu = C(CX[m]) * vx + C(CY[m]) * vy + C(CZ[m]) * vz;
      f[m] = feq2_raw<C, Store<S>::shifted>(r, cu, uu, C(W[m]), C(Store<S>::K));
    });
  }
}

template <class S>
__device__ __forceinline__ void pull_cell(const Planes<S> &P, uint8_t mk, uint32_t idx,
                                          const KArgs<typename Store<S>::C> &a,
                                          typename Store<S>::C (&f)[Q]) {
  bulk_loads<S>(P, idx, f);
  if (mk & (M_WALL | M_WRAP | M_BACKEQ)) patch_cell<S>(P, mk, idx, a, f);
}

// Moments, optional macro emission, optional collision, store of one cell.
template <class S, int FORCE, bool COLLIDE, bool EMIT>
__device__ __forceinline__ void finish_cell(typename Store<S>::C (&f)[Q], uint32_t idx,
                                            const KArgs<typename Store<S>::C> &a,
                                            const Planes<S> &P) {
  using C = typename Store<S>::C;
  if constexpr (COLLIDE || EMIT) {
    C rho, drho, jx, jy, jz;
    moments<S>(f, rho, drho, jx, jy, jz);
    C Fx, Fy, Fz;
    load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
    C ux, uy, uz;
    if constexpr (FORCE != 0) velocity(jx + C(0.5) * Fx, jy + C(0.5) * Fy, jz + C(0.5) * Fz, rho, ux, uy, uz);
    else velocity(jx, jy, jz, rho, ux, uy, uz);
    int bad = (rho < C(0) || rho > C(100)) ? ST_RHO_RANGE : 0;
    if constexpr (EMIT) {
      a.rho[idx] = rho;
      a.ux[idx] = ux;
      a.uy[idx] = uy;
      a.uz[idx] = uz;
    }
    if constexpr (COLLIDE) {
      if (relax<S, FORCE>(f, rho, drho, ux, uy, uz, Fx, Fy, Fz, a.omega)) bad |= ST_FEQ_NEG;
    }
    if (bad) flag(a.status, a.first_bad, bad, idx);
  }
  const uint32_t boff = idx * (uint32_t)sizeof(S);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    Store<S>::st(at(P.out[m], boff), f[m]);
  });
}

// One cell per thread.
// COLLIDE: relax and store f*; otherwise store f (post-streaming).
// EMIT   : write rho, u (u = (sum f c + F/2)/rho, lbm.cpp:672-675).
template <class S, int FORCE, bool COLLIDE, bool EMIT>
__global__ __launch_bounds__(256) void k_stream(const Planes<S> P, const uint8_t *__restrict__ mask,
                                                KArgs<typename Store<S>::C> a) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  C f[Q];
  bulk_loads<S>(P, idx, f);  // in flight together with the mask load
  const uint8_t mk = (uint8_t)after_loads(mask[idx], f);
  if (!(mk & M_FLUID)) return;
  if (mk & (M_WALL | M_WRAP | M_BACKEQ)) patch_cell<S>(P, mk, idx, a, f);
  finish_cell<S, FORCE, COLLIDE, EMIT>(f, idx, a, P);
}

// LBM::Collision (lbm.cpp:60-73) in place, with rho, u and F as stored.
template <class S, int FORCE>
__global__ __launch_bounds__(256) void k_collide(S *__restrict__ buf, const uint8_t *__restrict__ mask,
                                                 KArgs<typename Store<S>::C> a) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  if (!(mask[idx] & M_FLUID)) return;
  const uint32_t ps = a.g.ps;
  C f[Q];
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    f[m] = Store<S>::ld(buf + (uint32_t)m * ps + idx);
  });
  const C rho = a.rho[idx];
  const C ux = a.ux[idx], uy = a.uy[idx], uz = a.uz[idx];
  C Fx, Fy, Fz;
  load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
  const bool neg = relax<S, FORCE>(f, rho, rho - C(1), ux, uy, uz, Fx, Fy, Fz, a.omega);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    Store<S>::st(buf + (uint32_t)m * ps + idx, f[m]);
  });
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
  load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
  a.rho[idx] = rho;
  a.ux[idx] = (jx + C(0.5) * Fx) / rho;
  a.uy[idx] = (jy + C(0.5) * Fy) / rho;
  a.uz[idx] = (jz + C(0.5) * Fz) / rho;
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
template <class C>
__global__ __launch_bounds__(256) void k_tables(const int16_t *__restrict__ type,
                                                uint8_t *__restrict__ mask,
                                                uint32_t *__restrict__ links, C *uwx, C *uwy,
                                                C *uwz, Geo g, const C *__restrict__ ux,
                                                const C *__restrict__ uy,
                                                const C *__restrict__ uz) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell) return;
  uint32_t x, y, z;
  coords(g, idx, x, y, z);
  uint8_t mk = 0;
  uint32_t lk = 0;
  C wx = C(0), wy = C(0), wz = C(0);
  if (x < g.nx) {
    const int16_t t = type[idx];
    if (t == 0) mk |= M_FLUID;
    else if (t == 1) mk |= M_SOLID;
    else mk |= M_OTHER;
    if (x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1)
      mk |= M_WRAP;
    if (z == g.nz - 1 && type[idx - (g.nz - 1) * g.plane] == 3) mk |= M_BACKEQ;
    if (t == 0) {
      // bounce-back links: source x - c_m is solid
      for_dirs([&](auto M) {
        constexpr int m = decltype(M)::value;
        if (type[wrap_src(g, x, y, z, CX[m], CY[m], CZ[m])] == 1) lk |= 1u << m;
      });
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
      if (lk && (wx != C(0) || wy != C(0) || wz != C(0))) mk |= M_MOVING;
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

// TYPE_S cells hit by the back-face inflow write (M_BACKEQ): from the first
// Streaming on they hold feq2 of their TYPE_IN partner (lbm.cpp:599-611).
template <class S>
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
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    const C cu = C(CX[m]) * vx + C(CY[m]) * vy + C(CZ[m]) * vz;
    const S v = Store<S>::enc(feq2_raw<C, Store<S>::shifted>(r, cu, uu, C(W[m]), C(Store<S>::K)));
    A[(uint32_t)m * ps + idx] = v;
    B[(uint32_t)m * ps + idx] = v;
  });
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
  const uint32_t lk = (mk & M_WALL) ? links[idx] : 0u;
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

// Populations as FP64 f (unshifted), [idx][m].  TYPE_F and TYPE_S cells come
// from the buffer; other cells report feq of their stored rho,u (their Init
// value; their slots hold their neighbours' frozen values).
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
    if (mk & (M_FLUID | M_SOLID)) {
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

template <class S, int FORCE, bool COLLIDE, bool EMIT>
static void stream_launch(const LaunchCtx &L, const S *in, S *out) {
  auto a = make_args<S>(L);
  const Planes<S> P = make_planes(L, in, out);
  hipLaunchKernelGGL((k_stream<S, FORCE, COLLIDE, EMIT>), grid_for(L.g.ncell, L.block),
                     dim3(L.block), 0, L.stream, P, L.mask, a);
}

template <class S, int FORCE>
static hipError_t stream_dispatch(const LaunchCtx &L, const void *fin, void *fout, int mode) {
  const S *in = (const S *)fin;
  S *out = (S *)fout;
  switch (mode) {
    case STREAM_ONLY: stream_launch<S, FORCE, false, false>(L, in, out); break;
    case STREAM_MACRO: stream_launch<S, FORCE, false, true>(L, in, out); break;
    case STREAM_COLLIDE: stream_launch<S, FORCE, true, false>(L, in, out); break;
    case STREAM_COLLIDE_MACRO: stream_launch<S, FORCE, true, true>(L, in, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class S>
static hipError_t stream_s(const LaunchCtx &L, const void *fin, void *fout, int mode) {
  switch (L.force_mode) {
    case 0: return stream_dispatch<S, 0>(L, fin, fout, mode);
    case 1: return stream_dispatch<S, 1>(L, fin, fout, mode);
    case 2: return stream_dispatch<S, 2>(L, fin, fout, mode);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_stream(const LaunchCtx &L, const void *fin, void *fout, int mode) {
  switch (L.storage) {
    case 0: return stream_s<double>(L, fin, fout, mode);
    case 1: return stream_s<float>(L, fin, fout, mode);
    case 2: return stream_s<__half>(L, fin, fout, mode);
  }
  return hipErrorInvalidValue;
}

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
  }
  return hipErrorInvalidValue;
}

hipError_t launch_collide(const LaunchCtx &L, void *buf) {
  switch (L.storage) {
    case 0: return collide_s<double>(L, buf, false);
    case 1: return collide_s<float>(L, buf, false);
    case 2: return collide_s<__half>(L, buf, false);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_macro(const LaunchCtx &L, const void *buf) {
  switch (L.storage) {
    case 0: return collide_s<double>(L, (void *)buf, true);
    case 1: return collide_s<float>(L, (void *)buf, true);
    case 2: return collide_s<__half>(L, (void *)buf, true);
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
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_tables(const LaunchCtx &L, uint8_t *mask, uint32_t *links, void *uwx,
                         void *uwy, void *uwz, const int16_t *type) {
  const dim3 grid = grid_for(L.g.ncell, L.block);
  if (L.storage == 0)
    hipLaunchKernelGGL(k_tables<double>, grid, dim3(L.block), 0, L.stream, type, mask, links,
                       (double *)uwx, (double *)uwy, (double *)uwz, L.g, (const double *)L.ux,
                       (const double *)L.uy, (const double *)L.uz);
  else
    hipLaunchKernelGGL(k_tables<float>, grid, dim3(L.block), 0, L.stream, type, mask, links,
                       (float *)uwx, (float *)uwy, (float *)uwz, L.g, (const float *)L.ux,
                       (const float *)L.uy, (const float *)L.uz);
  return hipGetLastError();
}

hipError_t launch_backeq_solid(const LaunchCtx &L, void *A, void *B) {
  const dim3 grid = grid_for(L.g.ncell, L.block);
  switch (L.storage) {
    case 0:
      hipLaunchKernelGGL(k_backeq_solid<double>, grid, dim3(L.block), 0, L.stream, (double *)A, (double *)B, L.mask, make_args<double>(L));
      break;
    case 1:
      hipLaunchKernelGGL(k_backeq_solid<float>, grid, dim3(L.block), 0, L.stream, (float *)A, (float *)B, L.mask, make_args<float>(L));
      break;
    case 2:
      hipLaunchKernelGGL(k_backeq_solid<__half>, grid, dim3(L.block), 0, L.stream, (__half *)A, (__half *)B, L.mask, make_args<__half>(L));
      break;
    default:
      return hipErrorInvalidValue;
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
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
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
