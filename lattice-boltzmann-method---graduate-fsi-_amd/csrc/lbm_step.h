// lbm_step.h -- device helpers of the D3Q19 step shared by the kernel
// translation units: kernel arguments, moments, the BGK/Guo relaxation, the
// pull of one cell and its store.
#pragma once

#include "lbm_device.h"
#include "lbm_launch.h"

namespace lbm {


// FORCE: 0 none, 1 per-cell field, 2 uniform, 3 immersed-boundary step
// (u = sum f c / rho without the half force, F = 0, box cells overridden).
template <class C>
struct KArgs {
  Geo g;
  BoxArgs box;
  C omega;
  C Fx, Fy, Fz;                  // uniform force (force mode 2)
  const C *Fxa, *Fya, *Fza;      // per-cell force (force mode 1)
  const uint32_t *links;         // bit m: x - c_m is TYPE_S
  const C *uwx, *uwy, *uwz;      // per-cell wall velocity (lbm.cpp:941-962)
  C *rho, *ux, *uy, *uz;         // macroscopic arrays
  int *status;
  unsigned long long *first_bad;
  // Bounce-back through the solid's slots: after its collision a wall cell x
  // writes f*(x, opp m) + 2 w_m (u_w . c_m)/cs2 into slot m of the solid
  // x - c_m, which the next pull then reads like any streamed value.
  int wall_slots;
  // the stream kernel reads the link words: some wall link is not folded
  // into the pull address (g.fold), so some wall cell writes a solid's slot
  int slot_links;
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

// Posts a cell's status bits and index from the lanes that call it (a
// divergent branch): at most one atomic pair per wavefront -- the OR of the
// calling lanes' bits (status bits are below 256) and the smallest index,
// which is the first calling lane's unless a later lane holds a smaller one
// (then every calling lane posts its own).  Both words only grow (bits) or
// shrink (index) during a launch, so a post that the current value already
// covers is skipped: a launch with many bad cells does not serialise on
// the two words.
__device__ __forceinline__ void flag(int *status, unsigned long long *first_bad, int bits,
                                     uint32_t idx) {
  int all = 0;
  for (int b = 0; b < 8; b++)
    if (__ballot(bits & (1 << b))) all |= 1 << b;
  const uint32_t first = __builtin_amdgcn_readfirstlane(idx);
  const bool lead = (int)__lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1;
  const unsigned long long fb = __hip_atomic_load(first_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (__ballot(idx < first)) {
    if (idx < fb) atomicMin(first_bad, (unsigned long long)idx);
  } else if (lead && first < fb) {
    atomicMin(first_bad, (unsigned long long)first);
  }
  if (lead) {
    const int st = __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((st & all) != all) atomicOr(status, all);
  }
}

__device__ __forceinline__ float fma_c(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_c(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float vmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double vmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return __builtin_fmax(a, b); }

// u = (j + F/2)/rho.  FP64 divides each component (the reference's
// expression).  FP32 takes the quotient from the hardware reciprocal (1 ulp)
// with one correction step (Markstein 1990: q = j r, q += (j - q rho) r, the
// residual by fma), which rounds once at u's own scale except next to a tie.
//
// The plain product j fl(1/rho) is biased instead: a float rho near 1 lies
// on the grid 1 + m 2^-23 above 1 and 1 - m 2^-24 below, aligned with the
// grid of 1/rho, so fl(1/rho) errs in a pattern set by m (below 1 with m
// odd, 1/rho = 1 + m 2^-24 + O(m^2 2^-48) sits just above a tie and rounds
// up by 2^-24; above 1 the O(drho^2) term is dropped) rather than at random.
// The collision (omega near 2) feeds such an error of u into the momentum
// every step: the 256^3 Taylor-Green velocity at tau 0.503 drifted by 4.7e-5
// relative L2 in 5000 steps against FP64 storage, 1.3e-6 with u correctly
// rounded.  The divisor is the same rounded rho that the equilibrium takes,
// so the equilibrium's momentum rho u stays j to rounding.
template <class C>
__device__ __forceinline__ void velocity(C jx, C jy, C jz, C rho, C &ux, C &uy, C &uz) {
  if constexpr (sizeof(C) == 8) {
    ux = jx / rho;
    uy = jy / rho;
    uz = jz / rho;
  } else {
    const C ir = __builtin_amdgcn_rcpf(rho);
    const C qx = jx * ir, qy = jy * ir, qz = jz * ir;
    ux = __builtin_fmaf(__builtin_fmaf(-qx, rho, jx), ir, qx);
    uy = __builtin_fmaf(__builtin_fmaf(-qy, rho, jy), ir, qy);
    uz = __builtin_fmaf(__builtin_fmaf(-qz, rho, jz), ir, qz);
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
  //
  // K w_q is carried as hi + lo (hi = K w_q rounded to C, lo = the rest), and
  // rho K w_q = fma(hi, rho, lo rho) rounds once.  With hi alone every
  // product would carry hi's own rounding, the same sign in every cell and
  // step: the equilibrium's momentum would be (1 + 36 fl(1/36) - 1) rho u in
  // FP32, a gain of omega * 7.5e-9 per step that accumulates (1e-4 of the
  // velocity after 5000 steps of a Taylor-Green vortex).  In FP64 lo is 0.
  C ra[3], da[3];
  {
    constexpr double wd[3] = {1.0 / 3, 1.0 / 18, 1.0 / 36};
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const double kw = Store<S>::K * wd[q];
      const C hi = C(kw), lo = C(kw - double(hi));
      ra[q] = fma_c(hi, rho, lo * rho);
      da[q] = Store<S>::shifted ? fma_c(hi, drho, lo * drho) : ra[q];
    }
  }
  C fpre = C(0), Fu = C(0);
  if constexpr (FORCE != 0) {
    fpre = C(1) - omega * C(0.5);
    Fu = Fx * ux + Fy * uy + Fz * uz;
  }
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
  // feq = rho w (1 + A) < 0 for some direction (lbm.cpp:114-117).  For
  // |u|^2 <= 0.09 every 1 + A >= 0.23: over D3Q19 the smallest A is at
  // c.u = -sqrt(2)|u| (A grows with c.u while 13.5 a^2 + 9a + 3 - 4.5|u|^2
  // has no root, i.e. |u|^2 < 1/3), 1 + A = 1 - 3 sqrt(2) U + 7.5 U^2 -
  // 4.5 sqrt(2) U^3 >= 0.23 at U = 0.3, far above rounding; feq then has the
  // sign of rho.  Faster cells (and NaN) take the direction-by-direction test.
  if (uu <= C(0.09)) return rho < C(0);
  C amin = mh, amax = mh;  // A of the rest direction is -1.5uu
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    if constexpr (m % 2 == 1) {
      const C cu = C(CX[m]) * ux + C(CY[m]) * uy + C(CZ[m]) * uz;
      const C t2 = cu * cu;
      const C s_ = C(4.5) * t2 + mh;
      const C p_ = cu * (C(4.5) * t2 + c3);
      const C Ap = s_ + p_, Am = s_ - p_;
      amin = vmin(amin, vmin(Ap, Am));
      amax = vmax(amax, vmax(Ap, Am));
    }
  });
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

// Bulk loads with the periodic wrap of the padded lattice (lbm.cpp:197-199)
// folded into the addresses: a lane on the shell adds, per direction, the
// offset that takes its unwrapped source back into the lattice (+-nx cells
// in x, +-ny rows in y, +-nz planes in z).  Off the shell every correction is
// zero, and no lane needs a second, dependent round of loads.
// The lane's 32-bit byte offset must not go negative: it is taken from a
// base lowered by the largest negative correction (bias), so that
// offset = idx*size + bias + correction >= 0 for every lane.
template <class S, int N>
__device__ __forceinline__ void wrapped_loads(const S *const (&src)[N], const Geo &g, uint32_t idx,
                                              typename Store<S>::C (&f)[N]) {
  uint32_t x, y, z;
  coords(g, idx, x, y, z);
  const uint32_t e = (uint32_t)sizeof(S);
  const uint32_t ex = g.nx * e, ey = g.ny * g.px * e, ez = g.nz * g.plane * e;
  const uint32_t bias_elems = g.nx + g.ny * g.px + g.nz * g.plane;
  const uint32_t bias = bias_elems * e;
  // byte corrections: a source at x - 1 (direction c_x = +1) on the x = 0
  // face wraps by +nx cells, a source at x + 1 on the x = nx - 1 face by -nx
  const uint32_t xp = x == 0 ? ex : 0u, xm = x == g.nx - 1 ? ex : 0u;
  const uint32_t yp = y == 0 ? ey : 0u, ym = y == g.ny - 1 ? ey : 0u;
  const uint32_t zp = z == 0 ? ez : 0u, zm = z == g.nz - 1 ? ez : 0u;
  const uint32_t boff = idx * e + bias;
  for_dirs_impl(std::make_integer_sequence<int, N>{}, [&](auto M) {
    constexpr int m = decltype(M)::value;
    uint32_t o = boff;
    if constexpr (CX[m] > 0) o += xp;
    if constexpr (CX[m] < 0) o -= xm;
    if constexpr (CY[m] > 0) o += yp;
    if constexpr (CY[m] < 0) o -= ym;
    if constexpr (CZ[m] > 0) o += zp;
    if constexpr (CZ[m] < 0) o -= zm;
    f[m] = Store<S>::ld(at(src[m] - bias_elems, o));
  });
}
// the 19 population planes (N < 19: the first N directions, e.g. the seven
// of D3Q7, which are D3Q19's rest and axis directions)
template <class S>
__device__ __forceinline__ void bulk_loads_wrapped(const Planes<S> &P, const Geo &g, uint32_t idx,
                                                   typename Store<S>::C (&f)[Q]) {
  wrapped_loads<S, Q>(P.src, g, idx, f);
}

// Bulk loads with the periodic wrap and the shell fold (g.fold) both in the
// addresses.  A folded direction m reads the cell's own plane opp m,
//   in[opp m] + idx = src[m] + idx + (opp m - m) ps + off_m,
// instead of its source's slot, so the solids of a folded face need no
// bounce-back stores and the pull no link word.  A lane on a keep face
// (bits 8-13) reads its own cell in every plane, in[m] + idx = src[m] + idx
// + off_m: an inlet or outlet cell there stores its frozen-link values back
// unchanged.  The lane offset is taken from a base lowered by ps + plane +
// px + 1 elements, the most negative correction (it exceeds every wrap
// correction); the host enables the fold only where every resulting offset
// fits 32 bits.  Every address stays inside the population allocation: a
// folded or kept one is the cell's own entry, the others are those of
// wrapped_loads.  Returns whether the lane is on a keep face.
template <class S>
__device__ __forceinline__ bool folded_loads(const Planes<S> &P, const Geo &g, uint32_t idx,
                                             typename Store<S>::C (&f)[Q]) {
  uint32_t x, y, z;
  coords(g, idx, x, y, z);
  const uint32_t e = (uint32_t)sizeof(S);
  const uint32_t ex = g.nx * e, ey = g.ny * g.px * e, ez = g.nz * g.plane * e;
  const uint32_t bias_elems = g.ps + g.plane + g.px + 1;
  const uint32_t xp = x == 0 ? ex : 0u, xm = x == g.nx - 1 ? ex : 0u;
  const uint32_t yp = y == 0 ? ey : 0u, ym = y == g.ny - 1 ? ey : 0u;
  const uint32_t zp = z == 0 ? ez : 0u, zm = z == g.nz - 1 ? ez : 0u;
  const FoldLane fl = fold_lane(g, x, y, z);
  const bool kp = on_fold_face<8>(g, x, y, z);
  const uint32_t boff = (idx + bias_elems) * e;
  const uint32_t pse = g.ps * e, pxe = g.px * e, ple = g.plane * e;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    uint32_t o = boff;
    if constexpr (CX[m] > 0) o += xp;
    if constexpr (CX[m] < 0) o -= xm;
    if constexpr (CY[m] > 0) o += yp;
    if constexpr (CY[m] < 0) o -= ym;
    if constexpr (CZ[m] > 0) o += zp;
    if constexpr (CZ[m] < 0) o -= zm;
    if constexpr (m > 0) {
      // off_m and (opp m - m) ps + off_m in bytes, modulo 2^32
      const uint32_t off = (uint32_t)CX[m] * e + (uint32_t)CY[m] * pxe + (uint32_t)CZ[m] * ple;
      const uint32_t d = (OPP[m] > m ? pse : 0u - pse) + off;
      o = folds<m>(fl) ? boff + d : o;
      o = kp ? boff + off : o;
    }
    f[m] = Store<S>::ld(at(P.src[m] - bias_elems, o));
  });
  return kp;
}

// The moving-wall term of the folded directions, 6 K w_m (u_w . c_m)
// (lbm.cpp:268-275), for a fluid cell with a wall velocity.
template <class S>
__device__ __forceinline__ void fold_moving(const KArgs<typename Store<S>::C> &a, uint32_t idx,
                                            typename Store<S>::C (&f)[Q]) {
  using C = typename Store<S>::C;
  uint32_t x, y, z;
  coords(a.g, idx, x, y, z);
  const FoldLane fl = fold_lane(a.g, x, y, z);
  const C uwx = a.uwx[idx], uwy = a.uwy[idx], uwz = a.uwz[idx];
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    if constexpr (m > 0) {
      if (folds<m>(fl)) {
        const C ucw = C(CX[m]) * uwx + C(CY[m]) * uwy + C(CZ[m]) * uwz;
        f[m] = f[m] + C(Store<S>::K * 6.0 * W[m]) * ucw;
      }
    }
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
// as after_loads, for a second value loaded alongside the mask (the
// bounce-back link word)
__device__ __forceinline__ void after_loads2(uint32_t &mk, uint32_t &lk) {
  asm volatile("" : "+v"(mk), "+v"(lk));
}

template <class S, bool WRAPPED = false>
__device__ __forceinline__ void patch_cell(const Planes<S> &P, uint8_t mk, uint32_t idx,
                                           const KArgs<typename Store<S>::C> &a,
                                           typename Store<S>::C (&f)[Q]) {
  using C = typename Store<S>::C;
  const uint32_t boff = idx * (uint32_t)sizeof(S);
  if (!WRAPPED && (mk & M_WRAP)) {
    // periodic wrap of the padded lattice (lbm.cpp:197-199)
    uint32_t x, y, z;
    coords(a.g, idx, x, y, z);
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      const uint32_t src = wrap_src(a.g, x, y, z, CX[m], CY[m], CZ[m]);
      if (src != idx - pull_off<m>(a.g)) f[m] = Store<S>::ld(at(P.in[m], src * (uint32_t)sizeof(S)));
    });
  }
  if ((mk & M_WALL) && !a.wall_slots) {
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

// mask bits whose cells need patch_cell after the bulk loads
template <class C>
__device__ __forceinline__ uint8_t patch_bits(const KArgs<C> &a) {
  return a.wall_slots ? (uint8_t)(M_WRAP | M_BACKEQ) : (uint8_t)(M_WALL | M_WRAP | M_BACKEQ);
}

// The bounce-back values a wall cell hands to its solid neighbours' slots
// (wall_slots): slot m of x - c_m receives f*(x, opp m) + 6 K w_m (u_w . c_m),
// the value the reference's push writes into f(x, m) (lbm.cpp:268-275).
template <class S>
__device__ __forceinline__ void wall_slot_stores(const typename Store<S>::C (&f)[Q], uint8_t mk,
                                                 uint32_t idx, uint32_t lk,
                                                 const KArgs<typename Store<S>::C> &a,
                                                 S *const (&out)[Q]) {
  using C = typename Store<S>::C;
  C uwx = C(0), uwy = C(0), uwz = C(0);
  if (mk & M_MOVING) {
    uwx = a.uwx[idx];
    uwy = a.uwy[idx];
    uwz = a.uwz[idx];
  }
  uint32_t x = 0, y = 0, z = 0;
  if (mk & M_WRAP) coords(a.g, idx, x, y, z);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    if (lk & (1u << m)) {
      const uint32_t s = (mk & M_WRAP) ? wrap_src(a.g, x, y, z, CX[m], CY[m], CZ[m]) : idx - pull_off<m>(a.g);
      const C ucw = C(CX[m]) * uwx + C(CY[m]) * uwy + C(CZ[m]) * uwz;
      Store<S>::st(at(out[m], s * (uint32_t)sizeof(S)), f[OPP[m]] + C(Store<S>::K * 6.0 * W[m]) * ucw);
    }
  });
}

template <class S>
__device__ __forceinline__ void pull_cell(const Planes<S> &P, uint8_t mk, uint32_t idx,
                                          const KArgs<typename Store<S>::C> &a,
                                          typename Store<S>::C (&f)[Q]) {
  // the periodic wrap and the shell fold in the addresses, as stream_cell
  if (a.g.fold) folded_loads<S>(P, a.g, idx, f);
  else bulk_loads_wrapped<S>(P, a.g, idx, f);
  if (a.g.fold && (mk & M_MOVING)) fold_moving<S>(a, idx, f);
  if (mk & (patch_bits(a) & ~M_WRAP)) patch_cell<S, true>(P, mk, idx, a, f);
}

// Box index of a cell inside the IBM box (mask bit M_BOX).
__device__ __forceinline__ uint32_t box_index(const Geo &g, const BoxArgs &b, uint32_t idx) {
  uint32_t x, y, z;
  coords(g, idx, x, y, z);
  return (uint32_t)((int)x - b.x0) + (uint32_t)b.nx * ((uint32_t)((int)y - b.y0) + (uint32_t)b.ny * (uint32_t)((int)z - b.z0));
}

// rho, u and F a cell collides with.  Modes 0-2: MacroProp's velocity
// u = (sum f c + F/2)/rho (lbm.cpp:672-675).  Mode 3, the immersed-boundary
// step: the unforced velocity sum f c / rho and F = 0 of lbm.cpp:1262-1285
// (MacroProp is not called in IBM runs, main.cpp:197-211), replaced inside
// the box by the IBM result (lbm.cpp:1347-1352).
template <class S, int FORCE>
__device__ __forceinline__ void cell_state(const typename Store<S>::C (&f)[Q], uint8_t mk,
                                           uint32_t idx, const KArgs<typename Store<S>::C> &a,
                                           typename Store<S>::C &rho, typename Store<S>::C &drho,
                                           typename Store<S>::C &ux, typename Store<S>::C &uy,
                                           typename Store<S>::C &uz, typename Store<S>::C &Fx,
                                           typename Store<S>::C &Fy, typename Store<S>::C &Fz) {
  using C = typename Store<S>::C;
  C jx, jy, jz;
  moments<S>(f, rho, drho, jx, jy, jz);
  if constexpr (FORCE == 3) {
    if (mk & M_BOX) {
      const uint32_t b = box_index(a.g, a.box, idx);
      rho = (C)a.box.rho[b];
      drho = rho - C(1);
      ux = (C)a.box.ux[b];
      uy = (C)a.box.uy[b];
      uz = (C)a.box.uz[b];
      Fx = (C)a.box.Fx[b];
      Fy = (C)a.box.Fy[b];
      Fz = (C)a.box.Fz[b];
    } else {
      velocity(jx, jy, jz, rho, ux, uy, uz);
      Fx = Fy = Fz = C(0);
    }
  } else {
    load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
    if constexpr (FORCE != 0) velocity(jx + C(0.5) * Fx, jy + C(0.5) * Fy, jz + C(0.5) * Fz, rho, ux, uy, uz);
    else velocity(jx, jy, jz, rho, ux, uy, uz);
  }
}

// Moments, optional macro emission and optional collision of one fluid cell.
template <class S, int FORCE, bool COLLIDE, bool EMIT>
__device__ __forceinline__ void update_cell(typename Store<S>::C (&f)[Q], uint8_t mk, uint32_t idx,
                                            const KArgs<typename Store<S>::C> &a) {
  using C = typename Store<S>::C;
  if constexpr (COLLIDE || EMIT) {
    C rho, drho, ux, uy, uz, Fx, Fy, Fz;
    cell_state<S, FORCE>(f, mk, idx, a, rho, drho, ux, uy, uz, Fx, Fy, Fz);
    int bad = (rho < C(0) || rho > C(100)) ? ST_RHO_RANGE : 0;
    if constexpr (FORCE == 3) bad = 0;  // no MacroProp (and no density check) in IBM runs
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
}

// The loads of stream_cell: the 19 populations (FOLD: folded_loads, else
// the wrapped bulk loads), the mask and the link word, completed together.
template <class S, bool FOLD>
__device__ __forceinline__ uint32_t issue_loads(const Planes<S> &P, const uint8_t *__restrict__ mask,
                                                const KArgs<typename Store<S>::C> &a, uint32_t idx,
                                                typename Store<S>::C (&f)[Q], uint32_t &lk, bool &keep) {
  if constexpr (FOLD) keep = folded_loads<S>(P, a.g, idx, f);
  else bulk_loads_wrapped<S>(P, a.g, idx, f);
  uint32_t mk = mask[idx];
  lk = 0;
  if (a.slot_links && a.wall_slots) {
    lk = a.links[idx];
    after_loads2(mk, lk);
  }
  return after_loads(mk, f);
}

// The whole hot-kernel body for one cell (see k_stream).
//
// Solid and row-padding cells whose 128-byte lines hold fluid cells join the
// fluid cells' stores with zeros in the slots of their fill word (all slots
// of a padding cell when there are no solids): every such line then leaves
// the wave whole, from one store instruction, instead of with holes at the
// solids and the padding.  Nothing reads these slots: with wall slots a
// solid's populations are only bounce-back carriers.  Cells of the other
// types keep their frozen-link values.
template <class S, int FORCE, bool COLLIDE, bool EMIT, bool FOLD>
__device__ __forceinline__ void stream_cell(const Planes<S> &P, const uint8_t *__restrict__ mask,
                                            const KArgs<typename Store<S>::C> &a, uint32_t idx) {
  using C = typename Store<S>::C;
  constexpr uint32_t ALL = (1u << Q) - 1;
  C f[Q];
  // the populations (periodic wrap and, with FOLD, the shell fold in the
  // addresses), the mask and, where some wall link is not folded, the
  // bounce-back link word are all in flight together
  bool keep = false;  // on a keep face: rewrite the loaded own values
  uint32_t lk;
  uint32_t mk = issue_loads<S, FOLD>(P, mask, a, idx, f, lk, keep);
  // with no link word, solid and row-padding lanes write every slot (nothing
  // reads them) in waves that hold a fluid cell, so those lines leave whole
  const bool wave_fluid = __builtin_amdgcn_ballot_w64(mk & M_FLUID) != 0;
  uint32_t st = ALL;  // slots this lane stores
  if (mk & M_FLUID) {
    if (FOLD && (mk & M_MOVING)) fold_moving<S>(a, idx, f);
    if (mk & (patch_bits(a) & ~M_WRAP)) patch_cell<S, true>(P, (uint8_t)mk, idx, a, f);
    update_cell<S, FORCE, COLLIDE, EMIT>(f, (uint8_t)mk, idx, a);
  } else if (FOLD && keep && (mk & M_OTHER) && wave_fluid) {
    // an inlet/outlet cell of a keep face: its own (frozen-link) values back
  } else {
    st = !a.wall_slots ? 0u
         : a.slot_links ? lk
                        : ((mk == 0 || (mk & M_SOLID)) && wave_fluid ? ALL : 0u);
    if (!st) return;
    for (int m = 0; m < Q; m++) f[m] = C(0);
  }
  // encoded once, stored by one of two paths
  typename Store<S>::E e[Q];
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    e[m] = Store<S>::encw(f[m]);
  });
  const uint32_t boff = idx * (uint32_t)sizeof(S);
  if (__builtin_amdgcn_ballot_w64(st != ALL) == 0) {
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      Store<S>::st_e(at(P.out[m], boff), e[m]);
    });
  } else {
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      if (st & (1u << m)) Store<S>::st_e(at(P.out[m], boff), e[m]);
    });
  }
  if constexpr (COLLIDE) {
    if (a.wall_slots && (mk & (M_FLUID | M_WALL)) == (M_FLUID | M_WALL))
      wall_slot_stores<S>(f, (uint8_t)mk, idx, lk, a, P.out);
  }
}

}  // namespace lbm
