// lbm_thermal.hip -- the thermal branch of the reference (#define ENERGY) on
// gfx950: D3Q19 energy populations g on the same lattice as f.
//
// Per fluid cell and step, as the reference's ConvLBM / Streaming /
// MacroProp do with ENERGY defined:
//   pull     g like f (x - c_m fluid: g*(x - c_m, m); other types: frozen
//            slot); walls by the solid's type_Energy:
//              TYPE_T  anti-bounce-back g(x,m) = -g*(x,opp m) + 2 geq_wall
//                      with the wall temperature averaged over the cell's
//                      solid neighbours (lbm.cpp:276-280, 636-645, 965-981)
//              TYPE_Q  g(x,m) = w rho T_w (1 + 2nd-order(u_w)) with T_w from
//                      the zero-net-flux balance of Determine_Twall_Neumann
//                      (lbm.cpp:281-287, 983-1028), over the pulled values
//   moments  rho, u (with the force), T = sum g, q = sum g c (lbm.cpp:777-798)
//   collide  f: BGK + Guo as the plain step; the strain rate from the
//            non-equilibrium f gives Phi = 2 nu rho S:S; g: BGK with tau_g
//            and the source F = w Phi/cp (1 + c.u (tau_g - 1/2)/(cs2 tau_g))
//            plus dt^2/2 dF/dt (lbm.cpp:121-177)
// The previous collision's F_vis (needed for dF/dt) is rebuilt from that
// collision's Phi and u, which are kept per cell (4 values instead of 19;
// the same expression on the same inputs gives the same F_vis).
//
// Storage: FP64 holds g; FP32 holds g - w T0 (T0 near the mean temperature),
// the 16-bit storages (g - w T0) K in the units of their f
// so that small temperature differences keep their digits.
//
// The face loops' g (lbm_faces.hip) reach the fluid through the back-face
// inflow write only: a cell at k = Nz+1 above a TYPE_IN cell at k = 0 takes
// g = rho w T (1 + 3cu + 4.5cu^2 - 1.5uu) of that inlet after every Streaming
// (lbm.cpp:613-619) -- in the pull of a fluid cell (backeq_g), in both
// buffers of a solid (k_th_backeq_solid, whose T then feeds the wall
// temperature tables).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_step.h"

using namespace lbm;

namespace {


enum ThMode {
  TH_FUSED = 0,        // Streaming + MacroProp + next Collision
  TH_FUSED_EMIT = 1,   // as TH_FUSED, and write rho, u, T, q
  TH_STREAM = 2,       // Streaming alone
  TH_COLLIDE = 3,      // Collision in place (rho, u, T, F as stored)
  TH_MACRO = 4,        // MacroProp of the cells selected by `sel`
  TH_STREAM_EMIT = 5,  // Streaming + MacroProp
};

template <class C>
struct ThArgs {
  C tau, tau_g, nu, cp, j_dot_n, g_beta, T_ref, T0;
  C *T, *qx, *qy, *qz;
  C *phi, *upx, *upy, *upz;
  C *rho_w;
  const C *Tw;
  const uint64_t *tl;
  C *Fy;  // Boussinesq: per-cell F_y, rewritten after each MacroProp
};

// g storage: FP64 raw = g; FP32 raw = g - w T0; 16-bit raw = (g - w T0) K,
// the units of f (K a power of two: exact both ways)
template <class S>
__device__ __forceinline__ typename Store<S>::C g_val(typename Store<S>::C raw, double w,
                                                      typename Store<S>::C T0) {
  using C = typename Store<S>::C;
  if constexpr (Store<S>::K != C(1)) return raw * C(1.0 / Store<S>::K) + C(w) * T0;
  else if constexpr (Store<S>::shifted) return raw + C(w) * T0;
  else return raw;
}
template <class S>
__device__ __forceinline__ typename Store<S>::C g_raw(typename Store<S>::C val, double w,
                                                      typename Store<S>::C T0) {
  using C = typename Store<S>::C;
  if constexpr (Store<S>::K != C(1)) return (val - C(w) * T0) * C(Store<S>::K);
  else if constexpr (Store<S>::shifted) return val - C(w) * T0;
  else return val;
}

// F16C holds |g - w T0| < 0.5, at the rest weight 1/3 |rho T - T0| < 1.5: a
// cell with |T - T0| past 1.4 flags the temperature range rather than let
// its g saturate.
template <class S>
__device__ __forceinline__ bool beyond_storage(typename Store<S>::C T, typename Store<S>::C T0) {
  if constexpr (std::is_same<S, f16c_t>::value) return !(fabs(T - T0) < 1.4f);
  else return false;
}

// 1 + cubic Hermite terms of Calc_feq / Calc_geq (lbm.cpp:628-653), cs2 = 1/3
template <class C>
__device__ __forceinline__ C one_plus_A(C cu, C uu) {
  return C(1) + eq_A(cu, uu);
}

// The back-face inflow write's g (lbm.cpp:613-619): rho w T (1 + 3cu +
// 4.5cu^2 - 1.5uu) of the TYPE_IN cell s (its rho, u, T of the last
// MacroProp), as raw storage values.
template <class S>
__device__ __forceinline__ void backeq_g(uint32_t s, const KArgs<typename Store<S>::C> &a,
                                         const ThArgs<typename Store<S>::C> &t,
                                         typename Store<S>::C (&g)[Q]) {
  using C = typename Store<S>::C;
  const C r = a.rho[s], vx = a.ux[s], vy = a.uy[s], vz = a.uz[s], T = t.T[s];
  const C uu = vx * vx + vy * vy + vz * vz;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    const C cu = C(CX[m]) * vx + C(CY[m]) * vy + C(CZ[m]) * vz;
    const C B = C(3) * cu + C(4.5) * cu * cu - C(1.5) * uu;
    g[m] = g_raw<S>(r * C(W[m]) * T * (C(1) + B), W[m], t.T0);
  });
}

// Pull of g for one fluid cell: periodic wrap and the thermal walls.  f must
// already be pulled (the Neumann balance does not read it, but the order
// matches the reference's: every push first, then the walls).
template <class S>
__device__ __forceinline__ void pull_g(const Planes<S> &Pg, uint8_t mk, uint32_t idx,
                                       const KArgs<typename Store<S>::C> &a,
                                       const ThArgs<typename Store<S>::C> &t,
                                       typename Store<S>::C (&g)[Q]) {
  using C = typename Store<S>::C;
  if (mk & M_WRAP) {
    uint32_t x, y, z;
    coords(a.g, idx, x, y, z);
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      const uint32_t src = wrap_src(a.g, x, y, z, CX[m], CY[m], CZ[m]);
      if (src != idx - pull_off<m>(a.g)) g[m] = Store<S>::ld(at(Pg.in[m], src * (uint32_t)sizeof(S)));
    });
  }
  if (!(mk & M_WALL)) return;
  const uint64_t tl = t.tl[idx];
  if (!tl) return;
  const uint32_t boff = idx * (uint32_t)sizeof(S);
  const C rw = t.rho_w[idx], Tw = t.Tw[idx];
  C ux = C(0), uy = C(0), uz = C(0);
  if (mk & M_MOVING) {
    ux = a.uwx[idx];
    uy = a.uwy[idx];
    uz = a.uwz[idx];
  }
  const C uu = ux * ux + uy * uy + uz * uz;
  // Dirichlet: the wall equilibrium of the solid direction l = opp m
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    constexpr int l = OPP[m];
    if (tl & (1ull << m)) {
      const C cu = ux * C(CX[l]) + uy * C(CY[l]) + uz * C(CZ[l]);
      const C geq = rw * C(W[l]) * Tw * one_plus_A(cu, uu);
      const C gs = g_val<S>(Store<S>::ld(at(Pg.in[l], boff)), W[l], t.T0);
      g[m] = g_raw<S>(-gs + C(2) * geq, W[m], t.T0);
    }
  });
  if (!(tl >> 32)) return;
  // Neumann: pos = the directions m whose source x - c_m is solid
  const uint32_t lk = a.links[idx];
  int npos = 0, nx = 0, ny = 0, nz = 0;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    if (lk & (1u << m)) {
      npos++;
      nx += CX[m];
      ny += CY[m];
      nz += CZ[m];
    }
  });
  const C n0 = C(nx) / C(npos), n1 = C(ny) / C(npos), n2 = C(nz) / C(npos);
  C sum_notpos = C(0), sum_pos = C(0);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    const C cn = C(CX[m]) * n0 + C(CY[m]) * n1 + C(CZ[m]) * n2;
    if (!(lk & (1u << m))) {
      sum_notpos += g_val<S>(g[m], W[m], t.T0) * cn;
    }
  });
  // the pos terms in the reference's order: solid neighbours x + c_l for
  // l = 0..18, each contributing p = opp l
  for_dirs([&](auto M) {
    constexpr int l = decltype(M)::value;
    constexpr int p = OPP[l];
    if (lk & (1u << p)) {
      const C cn = C(CX[p]) * n0 + C(CY[p]) * n1 + C(CZ[p]) * n2;
      const C cu = ux * C(CX[p]) + uy * C(CY[p]) + uz * C(CZ[p]);
      sum_pos += C(W[p]) * rw * one_plus_A(cu, uu) * cn;
    }
  });
  const C TwN = (t.j_dot_n - sum_notpos) / sum_pos;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    constexpr int l = OPP[m];
    if (tl & (1ull << (32 + m))) {
      const C cu = ux * C(CX[l]) + uy * C(CY[l]) + uz * C(CZ[l]);
      const C v = C(W[l]) * rw * TwN * (C(1) + C(3) * cu + C(4.5) * cu * cu - C(1.5) * uu);
      g[m] = g_raw<S>(v, W[m], t.T0);
    }
  });
}

// One cell of the thermal step.  PULL: f, g pulled from the other buffer's
// neighbours (Streaming); else read in place.  MACRO: MacroProp semantics
// (moments of the cell's populations, Boussinesq force update).  COLLIDE:
// relax f and g (with rho, u, T from the moments when MACRO, from the
// macroscopic arrays otherwise, as LBM::Collision reads them).  EMIT: write
// rho, u, T, q.
template <class S, int FORCE, bool BQ, bool PULL, bool MACRO, bool COLLIDE, bool EMIT>
__global__ __launch_bounds__(256) void k_thermal(const Planes<S> Pf, const Planes<S> Pg,
                                                 const uint8_t *__restrict__ mask, uint8_t sel,
                                                 KArgs<typename Store<S>::C> a,
                                                 ThArgs<typename Store<S>::C> t) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  const uint8_t mk = mask[idx];
  if (!(mk & sel)) return;
  const uint32_t boff = idx * (uint32_t)sizeof(S);
  C f[Q], g[Q];
  if constexpr (PULL) {
    bulk_loads<S>(Pf, idx, f);
    bulk_loads<S>(Pg, idx, g);
    if (mk & (M_WALL | M_WRAP | M_BACKEQ)) patch_cell<S>(Pf, mk, idx, a, f);
    if (mk & (M_WALL | M_WRAP)) pull_g<S>(Pg, mk, idx, a, t, g);
    if (mk & M_BACKEQ) backeq_g<S>(idx - (a.g.nz - 1) * a.g.plane, a, t, g);
  } else {
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      f[m] = Store<S>::ld(at(Pf.in[m], boff));
      g[m] = Store<S>::ld(at(Pg.in[m], boff));
    });
  }
  C rho, drho, ux, uy, uz, T, Fx, Fy, Fz;
  int bad = 0;
  if constexpr (MACRO) {
    C jx, jy, jz;
    moments<S>(f, rho, drho, jx, jy, jz);
    if constexpr (FORCE == 1) {
      Fx = a.Fxa[idx];
      Fy = BQ ? t.Fy[idx] : a.Fya[idx];
      Fz = a.Fza[idx];
    } else {
      load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
    }
    // u = (sum f c + F/2)/rho with the force before this MacroProp's update
    velocity(jx + C(0.5) * Fx, jy + C(0.5) * Fy, jz + C(0.5) * Fz, rho, ux, uy, uz);
    C Ts = C(0), qx = C(0), qy = C(0), qz = C(0);
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      const C v = g_val<S>(g[m], W[m], t.T0);
      Ts += v;
      qx += v * C(CX[m]);
      qy += v * C(CY[m]);
      qz += v * C(CZ[m]);
    });
    T = Ts;
    if (rho < C(0) || rho > C(100)) bad |= ST_RHO_RANGE;
    if (T < C(0) || T > C(100) || beyond_storage<S>(T, t.T0)) bad |= ST_T_RANGE;
    if constexpr (BQ) {
      Fy = t.g_beta * (T - t.T_ref);  // lbm.cpp:803-806
      t.Fy[idx] = Fy;
    }
    if constexpr (EMIT) {
      a.rho[idx] = rho;
      a.ux[idx] = ux;
      a.uy[idx] = uy;
      a.uz[idx] = uz;
      t.T[idx] = T;
      t.qx[idx] = qx;
      t.qy[idx] = qy;
      t.qz[idx] = qz;
    }
  } else if constexpr (COLLIDE) {
    rho = a.rho[idx];
    drho = rho - C(1);
    ux = a.ux[idx];
    uy = a.uy[idx];
    uz = a.uz[idx];
    T = t.T[idx];
    if constexpr (FORCE == 1) {
      Fx = a.Fxa[idx];
      Fy = BQ ? t.Fy[idx] : a.Fya[idx];
      Fz = a.Fza[idx];
    } else {
      load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
    }
  }
  if constexpr (COLLIDE) {
    // ---- f: BGK + Guo (lbm.cpp:93-112), and the non-equilibrium
    // second moment for the strain rate (lbm.cpp:121-146)
    constexpr C K = C(Store<S>::K);
    const C omega = a.omega;
    const C uu = ux * ux + uy * uy + uz * uz;
    const C Fu = Fx * ux + Fy * uy + Fz * uz;
    const C fpre = C(1) - omega * C(0.5);
    C Pxx = C(0), Pxy = C(0), Pxz = C(0), Pyy = C(0), Pyz = C(0), Pzz = C(0);
    bool neg = false;
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      const C cu = ux * C(CX[m]) + uy * C(CY[m]) + uz * C(CZ[m]);
      const C A_ = eq_A(cu, uu);
      C fe;
      if constexpr (Store<S>::shifted) fe = C(K * W[m]) * (drho + rho * A_);
      else fe = rho * C(W[m]) * (C(1) + A_);
      if (rho * (C(1) + A_) < C(0)) neg = true;
      const C ne = f[m] - fe;
      if constexpr (CX[m] != 0) Pxx += C(CX[m] * CX[m]) * ne;
      if constexpr (CX[m] * CY[m] != 0) Pxy += C(CX[m] * CY[m]) * ne;
      if constexpr (CX[m] * CZ[m] != 0) Pxz += C(CX[m] * CZ[m]) * ne;
      if constexpr (CY[m] != 0) Pyy += C(CY[m] * CY[m]) * ne;
      if constexpr (CY[m] * CZ[m] != 0) Pyz += C(CY[m] * CZ[m]) * ne;
      if constexpr (CZ[m] != 0) Pzz += C(CZ[m] * CZ[m]) * ne;
      C fs = fe + (C(1) - omega) * ne;
      if constexpr (FORCE != 0) {
        const C Fc = Fx * C(CX[m]) + Fy * C(CY[m]) + Fz * C(CZ[m]);
        fs += C(K * W[m]) * fpre * (C(3) * Fc + C(9) * Fc * cu - C(3) * Fu);
      }
      f[m] = fs;
    });
    if (neg) bad |= ST_FEQ_NEG;
    if (mk & M_WALL) t.rho_w[idx] = rho;
    // S = -P/(2 rho tau cs2), Phi = 2 nu rho S:S (lbm.cpp:149-164)
    const C sd = C(2) * rho * t.tau * C(1.0 / 3.0) * K;
    const C Sxx = -Pxx / sd, Sxy = -Pxy / sd, Sxz = -Pxz / sd;
    const C Syy = -Pyy / sd, Syz = -Pyz / sd, Szz = -Pzz / sd;
    C Phi = Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Sxy * Sxy + Syy * Syy + Syz * Syz + Sxz * Sxz +
            Syz * Syz + Szz * Szz;
    Phi = C(2) * t.nu * rho * Phi;
    // ---- g: BGK with tau_g and the viscous-heating source (lbm.cpp:166-176)
    const C phi_p = t.phi[idx], px = t.upx[idx], py = t.upy[idx], pz = t.upz[idx];
    const C cs_num = t.tau_g - C(0.5), cs_den = C(1.0 / 3.0) * t.tau_g;
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      const C cu = ux * C(CX[m]) + uy * C(CY[m]) + uz * C(CZ[m]);
      const C cup = px * C(CX[m]) + py * C(CY[m]) + pz * C(CZ[m]);
      const C Fb = C(W[m]) * phi_p / t.cp * (C(1) + cup * cs_num / cs_den);
      const C Fn = C(W[m]) * Phi / t.cp * (C(1) + cu * cs_num / cs_den);
      const C geq = rho * C(W[m]) * T * one_plus_A(cu, uu);
      const C gv = g_val<S>(g[m], W[m], t.T0);
      g[m] = g_raw<S>(gv - (gv - geq) / t.tau_g + Fn + C(0.5) * (Fn - Fb), W[m], t.T0);
    });
    t.phi[idx] = Phi;
    t.upx[idx] = ux;
    t.upy[idx] = uy;
    t.upz[idx] = uz;
  }
  if (bad) flag(a.status, a.first_bad, bad, idx);
  if constexpr (PULL || COLLIDE) {
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      Store<S>::st(at(Pf.out[m], boff), f[m]);
      Store<S>::st(at(Pg.out[m], boff), g[m]);
    });
  }
}

// Link bits and wall temperature of every fluid cell (device order).
// tl bit m: x - c_m is a TYPE_S cell with type_Energy TYPE_T (Dirichlet),
// bit 32+m: TYPE_Q (Neumann).  Tw = Determine_Wall_Temperature(x)
// (lbm.cpp:965-981): the mean of Tsrc over the solid neighbours x + c_l in
// l order.
template <class C>
__global__ __launch_bounds__(256) void k_th_tables(const int16_t *__restrict__ type,
                                                   const int16_t *__restrict__ typeE,
                                                   const uint8_t *__restrict__ mask, Geo g,
                                                   const C *__restrict__ Tsrc,
                                                   uint64_t *__restrict__ tl, C *__restrict__ Tw) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell) return;
  uint64_t bits = 0;
  C tw = C(0);
  if (mask[idx] & M_FLUID) {
    uint32_t x, y, z;
    coords(g, idx, x, y, z);
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      const uint32_t s = wrap_src(g, x, y, z, CX[m], CY[m], CZ[m]);
      if (type[s] == 1) {
        if (typeE[s] == 6) bits |= 1ull << m;
        else if (typeE[s] == 7) bits |= 1ull << (32 + m);
      }
    });
    int num = 0;
    for_dirs([&](auto M) {
      constexpr int l = decltype(M)::value;
      const uint32_t n = wrap_dst(g, x, y, z, CX[l], CY[l], CZ[l]);
      if (type[n] == 1) {
        tw += Tsrc[n];
        num++;
      }
    });
    if (num) tw /= C(num);
  }
  tl[idx] = bits;
  Tw[idx] = tw;
}

// LBM::Init with ENERGY (lbm.cpp:31-57): f = feq and g = Calc_geq with the
// cell's own type and zero wall velocity, in both buffers.  Solids take
// rho_s w T_w with T_w the mean set-up T of their own solid neighbours
// (lbm.cpp:637-641, l = 0 included).  Cells that never collide (not TYPE_F
// or TYPE_S) store in slot m the Init values of their fluid neighbour x + c_m
// (frozen link, as for f).
template <class S>
__global__ __launch_bounds__(256) void k_th_init(S *__restrict__ A, S *__restrict__ B,
                                                 const uint8_t *__restrict__ mask,
                                                 const int16_t *__restrict__ type,
                                                 KArgs<typename Store<S>::C> a,
                                                 ThArgs<typename Store<S>::C> t) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  const uint8_t mk = mask[idx];
  uint32_t x, y, z;
  coords(a.g, idx, x, y, z);
  const uint32_t ps = a.g.ps;
  C Ts = C(0);
  if (mk & M_SOLID) {
    int num = 0;
    for_dirs([&](auto M) {
      constexpr int l = decltype(M)::value;
      const uint32_t n = wrap_dst(a.g, x, y, z, CX[l], CY[l], CZ[l]);
      if (type[n] == 1) {
        Ts += t.T[n];
        num++;
      }
    });
    Ts /= C(num);
  }
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
    const S sf = Store<S>::enc(r);
    A[(uint32_t)m * ps + idx] = sf;
    B[(uint32_t)m * ps + idx] = sf;
    C gv;
    if (mk & M_SOLID) gv = rho * C(W[m]) * Ts;
    else gv = rho * C(W[m]) * t.T[c] * (C(1) + A_);
    const S sg = Store<S>::enc(g_raw<S>(gv, W[m], t.T0));
    A[(uint32_t)(Q + m) * ps + idx] = sg;
    B[(uint32_t)(Q + m) * ps + idx] = sg;
  });
  t.phi[idx] = C(0);
  t.upx[idx] = C(0);
  t.upy[idx] = C(0);
  t.upz[idx] = C(0);
  t.rho_w[idx] = a.rho[idx];
}

// The pulled f and g of the fluid cells an outlet copies from, into their
// face slots (FP64, direction-major: f rows, then g rows): the values the
// next pulling launch computes for them, from the same buffer and tables,
// before its back-face inflow write.  Under that write (M_BACKEQ) the links
// from cells of the other types keep the slot's value: the reference's
// frozen link holds the previous step's write there.
template <class S>
__global__ __launch_bounds__(256) void k_th_face_src(const Planes<S> Pf, const Planes<S> Pg,
                                                     const uint8_t *__restrict__ mask,
                                                     const int16_t *__restrict__ type,
                                                     KArgs<typename Store<S>::C> a,
                                                     ThArgs<typename Store<S>::C> t,
                                                     const int32_t *__restrict__ list, int n,
                                                     double *__restrict__ of, int ns) {
  using C = typename Store<S>::C;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t idx = (uint32_t)list[2 * k];
  const size_t slot = (size_t)list[2 * k + 1];
  C f[Q], g[Q];
  bulk_loads<S>(Pf, idx, f);
  bulk_loads<S>(Pg, idx, g);
  const uint8_t mk0 = mask[idx];
  const uint8_t mk = mk0 & (uint8_t)~M_BACKEQ;
  if (mk & (M_WALL | M_WRAP)) {
    patch_cell<S>(Pf, mk, idx, a, f);
    pull_g<S>(Pg, mk, idx, a, t, g);
  }
  uint32_t x = 0, y = 0, z = 0;
  if (mk0 & M_BACKEQ) coords(a.g, idx, x, y, z);
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    if (mk0 & M_BACKEQ) {
      const int16_t tq = type[wrap_src(a.g, x, y, z, CX[m], CY[m], CZ[m])];
      if (tq != LBM_TYPE_F && tq != LBM_TYPE_S) return;
    }
    double fv = (double)f[m], gv = (double)g[m];
    if constexpr (Store<S>::shifted) {
      fv = fv / (double)Store<S>::K + W[m];
      gv = gv / (double)Store<S>::K + W[m] * (double)t.T0;
    }
    of[(size_t)m * ns + slot] = fv;
    of[(size_t)(Q + m) * ns + slot] = gv;
  });
}

// The back-face inflow write's g into the solids at k = Nz+1 above a
// TYPE_IN cell (lbm.cpp:613-619), both buffers (f: k_backeq_solid)
template <class S>
__global__ __launch_bounds__(256) void k_th_backeq_solid(S *__restrict__ A, S *__restrict__ B,
                                                         const uint8_t *__restrict__ mask,
                                                         KArgs<typename Store<S>::C> a,
                                                         ThArgs<typename Store<S>::C> t) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  if ((mask[idx] & (M_SOLID | M_BACKEQ)) != (M_SOLID | M_BACKEQ)) return;
  C g[Q];
  backeq_g<S>(idx - (a.g.nz - 1) * a.g.plane, a, t, g);
  const uint32_t ps = a.g.ps;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    const S v = Store<S>::enc(g[m]);
    A[(uint32_t)(Q + m) * ps + idx] = v;
    B[(uint32_t)(Q + m) * ps + idx] = v;
  });
}

// g as FP64 values, [idx][m]
template <class S>
__global__ __launch_bounds__(256) void k_g_out(const S *__restrict__ buf, Geo g,
                                               typename Store<S>::C T0, double *__restrict__ out) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell) return;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    const C raw = Store<S>::ld(buf + (uint32_t)(Q + m) * g.ps + idx);
    out[(size_t)idx * Q + m] = (double)g_val<S>(raw, W[m], T0);
  });
}

inline dim3 grid_of(const lbm_ctx *c) { return dim3((c->g.ncell + c->block - 1) / c->block); }

template <class S>
KArgs<typename Store<S>::C> kargs(const LaunchCtx &L) {
  using C = typename Store<S>::C;
  KArgs<C> a{};
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
ThArgs<typename Store<S>::C> thargs(const lbm_ctx *c) {
  using C = typename Store<S>::C;
  ThArgs<C> t{};
  t.tau = (C)c->d.tau;
  t.tau_g = (C)c->td.tau_g;
  t.nu = (C)c->td.nu;
  t.cp = (C)c->td.cp;
  t.j_dot_n = (C)c->td.j_dot_n;
  t.g_beta = (C)c->td.g_beta;
  t.T_ref = (C)c->td.T_ref;
  t.T0 = (C)c->T0;
  t.T = (C *)c->T;
  t.qx = (C *)c->qv[0];
  t.qy = (C *)c->qv[1];
  t.qz = (C *)c->qv[2];
  t.phi = (C *)c->phi;
  t.upx = (C *)c->up[0];
  t.upy = (C *)c->up[1];
  t.upz = (C *)c->up[2];
  t.rho_w = (C *)c->rho_w;
  t.Tw = (const C *)c->Tw;
  t.tl = c->tlinks;
  t.Fy = (C *)c->F[1];
  return t;
}

template <class S>
Planes<S> planes(const lbm_ctx *c, const void *in, void *out, int first) {
  Planes<S> P;
  const size_t ps = c->g.ps;
  const S *bi = (const S *)in + (size_t)first * ps;
  S *bo = out ? (S *)out + (size_t)first * ps : nullptr;
  for (int m = 0; m < Q; m++) {
    const ptrdiff_t off = (ptrdiff_t)CX[m] + (ptrdiff_t)CY[m] * c->g.px + (ptrdiff_t)CZ[m] * c->g.plane;
    P.in[m] = bi + m * ps;
    P.src[m] = bi + (ptrdiff_t)(m * ps) - off;
    P.out[m] = bo ? bo + m * ps : nullptr;
  }
  return P;
}

template <class S, int FORCE, bool BQ, bool PULL, bool MACRO, bool COLLIDE, bool EMIT>
void launch_k(lbm_ctx *c, const void *in, void *out, uint8_t sel) {
  const LaunchCtx L = launch_ctx(c);
  const Planes<S> Pf = planes<S>(c, in, out, 0), Pg = planes<S>(c, in, out, Q);
  hipLaunchKernelGGL((k_thermal<S, FORCE, BQ, PULL, MACRO, COLLIDE, EMIT>), grid_of(c), dim3(c->block), 0,
                     c->stream, Pf, Pg, c->mask, sel, kargs<S>(L), thargs<S>(c));
}

template <class S, int FORCE, bool BQ>
void launch_mode(lbm_ctx *c, int mode, uint8_t sel) {
  switch (mode) {
    case TH_FUSED: launch_k<S, FORCE, BQ, true, true, true, false>(c, c->cur, c->oth, M_FLUID); break;
    case TH_FUSED_EMIT: launch_k<S, FORCE, BQ, true, true, true, true>(c, c->cur, c->oth, M_FLUID); break;
    case TH_STREAM: launch_k<S, FORCE, BQ, true, false, false, false>(c, c->cur, c->oth, M_FLUID); break;
    case TH_STREAM_EMIT: launch_k<S, FORCE, BQ, true, true, false, true>(c, c->cur, c->oth, M_FLUID); break;
    case TH_COLLIDE: launch_k<S, FORCE, BQ, false, false, true, false>(c, c->cur, c->cur, M_FLUID); break;
    case TH_MACRO: launch_k<S, FORCE, BQ, false, true, false, true>(c, c->cur, nullptr, sel); break;
  }
}

template <class S>
void launch_s(lbm_ctx *c, int mode, uint8_t sel) {
  if (c->td.boussinesq) {
    launch_mode<S, 1, true>(c, mode, sel);
    return;
  }
  switch (c->force_mode) {
    case 0: launch_mode<S, 0, false>(c, mode, sel); break;
    case 1: launch_mode<S, 1, false>(c, mode, sel); break;
    default: launch_mode<S, 2, false>(c, mode, sel); break;
  }
}

// f(Tag<S>{}) with S the context's storage type
template <class S>
struct Tag {
  using type = S;
};
template <class F>
void with_storage(const lbm_ctx *c, F &&f) {
  switch (c->storage) {
    case LBM_STORE_F64: f(Tag<double>{}); break;
    case LBM_STORE_F16S: f(Tag<__half>{}); break;
    case LBM_STORE_F16C: f(Tag<f16c_t>{}); break;
    default: f(Tag<float>{}); break;
  }
}

int upload_cells_c(lbm_ctx *c, void *dst, const std::vector<double> &h) {
  const size_t n = c->g.ncell;
  if (c->csize == 8) {
    HIPCHK(hipMemcpy(dst, h.data(), n * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(h.begin(), h.end());
    HIPCHK(hipMemcpy(dst, f.data(), n * 4, hipMemcpyHostToDevice));
  }
  return LBM_OK;
}

}  // namespace

void thermal_free(lbm_ctx *c) {
  void **ptrs[] = {&c->T, &c->qv[0], &c->qv[1], &c->qv[2], &c->phi, &c->up[0], &c->up[1],
                   &c->up[2], &c->rho_w, &c->Tw, (void **)&c->tlinks, (void **)&c->typeE};
  for (void **p : ptrs) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  c->thermal = 0;
}

int thermal_tables(lbm_ctx *c) {
  const int16_t *typeE = c->typeE;
  hipError_t e = hipSuccess;
  {
    if (c->storage == LBM_STORE_F64)
      hipLaunchKernelGGL(k_th_tables<double>, grid_of(c), dim3(c->block), 0, c->stream, c->type, typeE, c->mask, c->g,
                         (const double *)c->T, c->tlinks, (double *)c->Tw);
    else
      hipLaunchKernelGGL(k_th_tables<float>, grid_of(c), dim3(c->block), 0, c->stream, c->type, typeE, c->mask, c->g,
                         (const float *)c->T, c->tlinks, (float *)c->Tw);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? LBM_OK : LBM_ERR_HIP;
}

int thermal_init(lbm_ctx *c) {
  int e = thermal_tables(c);
  if (e != LBM_OK) return e;
  const LaunchCtx L = launch_ctx(c);
  with_storage(c, [&](auto tag) {
    using S = typename decltype(tag)::type;
    hipLaunchKernelGGL(k_th_init<S>, grid_of(c), dim3(c->block), 0, c->stream, (S *)c->cur, (S *)c->oth, c->mask,
                       c->type, kargs<S>(L), thargs<S>(c));
  });
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int thermal_launch(lbm_ctx *c, int mode, uint8_t sel) {
  with_storage(c, [&](auto tag) { launch_s<typename decltype(tag)::type>(c, mode, sel); });
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int thermal_download_g(lbm_ctx *c, double *g_dev) {
  with_storage(c, [&](auto tag) {
    using S = typename decltype(tag)::type;
    using C = typename Store<S>::C;
    hipLaunchKernelGGL(k_g_out<S>, grid_of(c), dim3(c->block), 0, c->stream, (const S *)c->cur, c->g, (C)c->T0,
                       g_dev);
  });
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

template <class S>
void face_src_s(lbm_ctx *c) {
  const LaunchCtx L = launch_ctx(c);
  const Planes<S> Pf = planes<S>(c, c->cur, nullptr, 0), Pg = planes<S>(c, c->cur, nullptr, Q);
  hipLaunchKernelGGL(k_th_face_src<S>, dim3((c->fnsrc + 255) / 256), dim3(256), 0, c->stream, Pf, Pg, c->mask,
                     c->type, kargs<S>(L), thargs<S>(c), c->fsrc, c->fnsrc, c->fof, c->fnslot);
}

int thermal_face_sources(lbm_ctx *c) {
  if (c->thermal != 1 || !c->fnsrc) return LBM_OK;
  with_storage(c, [&](auto tag) { face_src_s<typename decltype(tag)::type>(c); });
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int thermal_backeq_solid(lbm_ctx *c) {
  if (c->thermal != 1) return LBM_OK;
  const LaunchCtx L = launch_ctx(c);
  with_storage(c, [&](auto tag) {
    using S = typename decltype(tag)::type;
    hipLaunchKernelGGL(k_th_backeq_solid<S>, grid_of(c), dim3(c->block), 0, c->stream, (S *)c->cur, (S *)c->oth,
                       c->mask, kargs<S>(L), thargs<S>(c));
  });
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

extern "C" {

int lbm_upload_thermal(lbm_ctx *c, const lbm_thermal_desc *d, const double *T,
                       const int16_t *type_energy) {
  if (!c || !d) return LBM_ERR_ARG;
  if (!c->uploaded || c->state != ST_NONE) return LBM_ERR_STATE;
  if (c->ibm) return LBM_ERR_UNSUPPORTED;
  if (d->model != LBM_THERMAL_ENERGY && d->model != LBM_THERMAL_D3Q7) return LBM_ERR_ARG;
  if (d->model == LBM_THERMAL_D3Q7) {
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipStreamSynchronize(c->stream));
    return q7_upload(c, d, T, type_energy);
  }
  if (!(d->tau_g > 0.5) || !(d->cp > 0) || !std::isfinite(d->nu)) return LBM_ERR_ARG;
  if ((uint64_t)c->g.ps * 38ull >= (1ull << 32)) return LBM_ERR_UNSUPPORTED;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t n = c->g.ncell;
  const uint32_t nx = c->g.nx, ny = c->g.ny, nz = c->g.nz, px = c->g.px;
  // host copies in device order
  std::vector<double> hT(n, 0.0);
  c->h_typeE.assign(n, (int16_t)LBM_TYPE_F);
  for (uint32_t i = 0; i < nx; i++)
    for (uint32_t j = 0; j < ny; j++)
      for (uint32_t k = 0; k < nz; k++) {
        const size_t r = ((size_t)i * ny + j) * nz + k, dd = ((size_t)k * ny + j) * px + i;
        if (T) hT[dd] = T[r];
        if (type_energy) c->h_typeE[dd] = type_energy[r];
      }
  // every solid that borders a fluid cell needs a thermal condition: a
  // solid of type_Energy TYPE_F would freeze the fluid cell's g link
  for (uint32_t z = 0; z < nz; z++)
    for (uint32_t y = 0; y < ny; y++)
      for (uint32_t x = 0; x < nx; x++) {
        const size_t dd = ((size_t)z * ny + y) * px + x;
        if (c->h_type[dd] != LBM_TYPE_S || c->h_typeE[dd] == LBM_TYPE_T || c->h_typeE[dd] == LBM_TYPE_Q) continue;
        for (int m = 0; m < 19; m++) {
          const uint32_t xs = (x + nx + CX[m]) % nx, ys = (y + ny + CY[m]) % ny, zs = (z + nz + CZ[m]) % nz;
          if (c->h_type[((size_t)zs * ny + ys) * px + xs] == LBM_TYPE_F) return LBM_ERR_UNSUPPORTED;
        }
      }
  thermal_free(c);
  c->td = *d;
  // storage shift: the mean set-up T of the fluid cells unless given
  double T0 = d->T0;
  if (!std::isfinite(T0)) {
    double s = 0;
    size_t k = 0;
    for (size_t i = 0; i < n; i++)
      if (c->h_type[i] == LBM_TYPE_F) {
        s += hT[i];
        k++;
      }
    T0 = k ? s / (double)k : 0.0;
  }
  c->T0 = c->storage == LBM_STORE_F64 ? 0.0 : T0;
  // 38-plane population buffers
  const size_t popbytes = c->esize * (size_t)c->g.ps * 38;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, popbytes) != hipSuccess) return LBM_ERR_NOMEM;
  if (hipMalloc(&b, popbytes) != hipSuccess) {
    (void)hipFree(a);
    return LBM_ERR_NOMEM;
  }
  (void)hipFree(c->cur);
  (void)hipFree(c->oth);
  c->cur = a;
  c->oth = b;
  void **arr[] = {&c->T, &c->qv[0], &c->qv[1], &c->qv[2], &c->phi, &c->up[0], &c->up[1], &c->up[2], &c->rho_w, &c->Tw};
  for (void **p : arr)
    if (hipMalloc(p, n * c->csize) != hipSuccess) return LBM_ERR_NOMEM;
  if (hipMalloc((void **)&c->tlinks, n * 8) != hipSuccess) return LBM_ERR_NOMEM;
  if (hipMalloc((void **)&c->typeE, n * 2) != hipSuccess) return LBM_ERR_NOMEM;
  HIPCHK(hipMemcpy(c->typeE, c->h_typeE.data(), n * 2, hipMemcpyHostToDevice));
  int e = upload_cells_c(c, c->T, hT);
  if (e != LBM_OK) return e;
  for (int q = 0; q < 3; q++) HIPCHK(hipMemsetAsync(c->qv[q], 0, n * c->csize, c->stream));
  // the Boussinesq force is rewritten per cell by every MacroProp: keep it
  // as a per-cell field
  if (d->boussinesq && c->force_mode != 1) {
    for (int q = 0; q < 3; q++) {
      std::vector<double> h(n, c->force_mode == 2 ? c->Fu[q] : 0.0);
      if (hipMalloc(&c->F[q], n * c->csize) != hipSuccess) return LBM_ERR_NOMEM;
      if ((e = upload_cells_c(c, c->F[q], h)) != LBM_OK) return e;
    }
    c->force_mode = 1;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->thermal = 1;
  return LBM_OK;
}

int lbm_download_thermal(lbm_ctx *c, double *T, double *qx, double *qy, double *qz) {
  if (!c) return LBM_ERR_ARG;
  if (!c->thermal) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  double *outs[4] = {T, qx, qy, qz};
  void *src[4] = {c->T, c->qv[0], c->qv[1], c->qv[2]};
  const size_t n = c->g.ncell;
  const uint32_t nx = c->g.nx, ny = c->g.ny, nz = c->g.nz, px = c->g.px;
  std::vector<double> h(n);
  std::vector<float> hf(c->csize == 4 ? n : 0);
  for (int q = 0; q < 4; q++) {
    if (!outs[q]) continue;
    if (c->csize == 8) {
      HIPCHK(hipMemcpy(h.data(), src[q], n * 8, hipMemcpyDeviceToHost));
    } else {
      HIPCHK(hipMemcpy(hf.data(), src[q], n * 4, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; i++) h[i] = hf[i];
    }
    for (uint32_t i = 0; i < nx; i++)
      for (uint32_t j = 0; j < ny; j++)
        for (uint32_t k = 0; k < nz; k++)
          outs[q][((size_t)i * ny + j) * nz + k] = h[((size_t)k * ny + j) * px + i];
  }
  return LBM_OK;
}

int lbm_download_g(lbm_ctx *c, double *g) {
  if (!c || !g) return LBM_ERR_ARG;
  if (!c->thermal || c->state == ST_NONE) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = c->g.ncell;
  const int ng = c->thermal == 2 ? 7 : 19;
  double *d = nullptr;
  HIPCHK(hipMalloc((void **)&d, n * ng * sizeof(double)));
  int e = c->thermal == 2 ? q7_download_g(c, d) : thermal_download_g(c, d);
  if (e == LBM_OK && c->thermal == 1 && faces_g_out(c, d) != hipSuccess) e = LBM_ERR_HIP;
  std::vector<double> h;
  if (e == LBM_OK) {
    h.resize(n * ng);
    if (hipMemcpyAsync(h.data(), d, n * ng * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      e = LBM_ERR_HIP;
  }
  (void)hipFree(d);
  if (e != LBM_OK) return e;
  const uint32_t nx = c->g.nx, ny = c->g.ny, nz = c->g.nz, px = c->g.px;
  for (uint32_t i = 0; i < nx; i++)
    for (uint32_t j = 0; j < ny; j++)
      for (uint32_t k = 0; k < nz; k++)
        memcpy(g + (((size_t)i * ny + j) * nz + k) * ng, h.data() + (((size_t)k * ny + j) * px + i) * ng,
               ng * sizeof(double));
  return LBM_OK;
}

}  // extern "C"
