// lbm_q7.hip -- temperature as a D3Q7 advection-diffusion distribution
// coupled to the D3Q19 flow (LBM_THERMAL_D3Q7).  The reference has no D3Q7:
// its thermal branch is the same-lattice D3Q19 g of lbm_thermal.hip.
//
// Model (Krueger et al. 2017, The Lattice Boltzmann Method, ch. 8):
//   g_i, i = 0..6 over D3Q19's rest and axis directions, w_0 = 1/4,
//   w_1..6 = 1/8, cs2 = 1/4;  T = sum g_i;  geq_i = w_i T (1 + c_i.u/cs2);
//   BGK  g* = geq + (1 - omega_g)(g - geq),  omega_g = 1/tau_g,
//   diffusivity alpha = cs2 (tau_g - 1/2).
// Coupling, per fluid cell and step: T from the pulled g; the Boussinesq
// force F_y += g_beta (T - T_ref) at that time level enters the velocity
// u = (sum f c + F/2)/rho and the Guo source; g relaxes with that u.
// Walls (half-way, like f): a D3Q7 link from fluid cell x to solid s = x - c_m
// carries, by the solid's type_Energy,
//   TYPE_T  g(x,m) = -g*(x, opp m) + 2 w_m T_s   (Dirichlet at T_s)
//   TYPE_Q  g(x,m) =  g*(x, opp m) + j           (heat flux j into the fluid
//                                                  through each wall face)
//   other   g(x,m) =  g*(x, opp m)               (adiabatic)
// written by x into the solid's slot m at its collision (the wall slots of
// the flow populations).  Other cell types (IN, OUT, ...) freeze the link as
// they do for f: x keeps its own Init value in that direction.
//
// T is a relative temperature here (it may be negative about T_ref): its
// guard flags blow-up only, |T| > 100 or not finite, where the ENERGY
// branch flags T outside [0, 100] as the reference does (lbm.cpp:794-797).
//
// Buffers: 26 planes, f in 0..18, g in 19..25.  FP64 stores g; the other
// storages store g - w_i T0 (T0 near the mean temperature), the 16-bit ones
// in the units of their f (scaled by Store<S>::K, a power of two: the scaling
// is exact).  The kernels take g into temperature units after the loads and
// back at the stores.  T and q = sum g c are written for fluid cells only; a
// solid's T is its Dirichlet temperature.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <cmath>
#include <type_traits>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_step.h"

using namespace lbm;

namespace {

constexpr int Q7 = 7;
constexpr uint32_t ALL19 = (1u << Q) - 1;
// links bit 20 + m (fluid cells): the D3Q7 link m ends on a TYPE_T solid;
// neumann_bit(m): it ends on a TYPE_Q solid (bits 27..31, then 20, which no
// Dirichlet link uses)
constexpr int DIRICHLET_BIT = 20;
__host__ __device__ constexpr uint32_t neumann_bit(int m) { return m < 6 ? 1u << (26 + m) : 1u << 20; }
constexpr double W7[Q7] = {0.25, 0.125, 0.125, 0.125, 0.125, 0.125, 0.125};

template <int N, class F>
__device__ __forceinline__ void for_n(F &&f) {
  for_dirs_impl(std::make_integer_sequence<int, N>{}, f);
}

template <class S>
struct G7 {
  const S *src[Q7];
  const S *in[Q7];
  S *out[Q7];
};

template <class C>
struct Q7Args {
  C omega_g, g_beta, T_ref, T0;
  C jn;  // heat flux through a TYPE_Q wall face (lbm_thermal_desc.j_dot_n)
  C *T, *qx, *qy, *qz;
};

// F16C holds |g - w T0| < 0.5, at the rest weight 1/4 |T - T0| < 2: a cell
// past 1.9 flags the temperature range rather than let its g saturate.
template <class S>
__device__ __forceinline__ bool beyond_storage(typename Store<S>::C dT) {
  if constexpr (std::is_same<S, f16c_t>::value) return !(fabs(dT) < 1.9f);
  else return false;
}

// Loaded g (storage units) into temperature units, in place.
template <class S>
__device__ __forceinline__ void g_units(typename Store<S>::C (&g)[Q7]) {
  using C = typename Store<S>::C;
  if constexpr (Store<S>::K != C(1)) {
    constexpr C inv = C(1) / Store<S>::K;
#pragma unroll
    for (int m = 0; m < Q7; m++) g[m] = g[m] * inv;
  }
}

// Store of one g value given in temperature units.
template <class S>
__device__ __forceinline__ void g_st(S *p, typename Store<S>::C v) {
  Store<S>::st(p, v * Store<S>::K);
}

template <class C>
__device__ __forceinline__ uint32_t after_loads7(uint32_t mk, const C (&g)[Q7]) {
  asm volatile("" : "+v"(mk) : "v"(g[0]), "v"(g[1]), "v"(g[2]), "v"(g[3]), "v"(g[4]), "v"(g[5]), "v"(g[6]));
  return mk;
}

// T - T0 and q = sum g c from stored (shifted) values
template <class C>
__device__ __forceinline__ C g_sum(const C (&g)[Q7]) {
  return g[0] + ((g[1] + g[2]) + ((g[3] + g[4]) + (g[5] + g[6])));
}

// The force the flow collides with: the set force (modes 0-2) plus the
// buoyancy of this cell's temperature.
template <class C, int FORCE, bool BQ>
__device__ __forceinline__ void q7_force(const KArgs<C> &a, const Q7Args<C> &t, uint32_t idx, C T, C &Fx,
                                         C &Fy, C &Fz) {
  load_force<C, FORCE>(a, idx, Fx, Fy, Fz);
  if constexpr (BQ) Fy += t.g_beta * (T - t.T_ref);
}

// BGK of the seven g in place (shifted values: g - w T0), at temperature T
template <class C>
__device__ __forceinline__ void g_relax(C (&g)[Q7], C T, C dT, C ux, C uy, C uz, C omega_g) {
  const C om1 = C(1) - omega_g;
  {
    const C ge = C(W7[0]) * dT;
    g[0] = ge + om1 * (g[0] - ge);
  }
  const C u[3] = {ux, uy, uz};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    // directions 1+2d (+e_d) and 2+2d (-e_d): w (dT +- T c.u / cs2)
    const C a = C(W7[1]) * dT, b = C(W7[1] * 4.0) * T * u[d];
    const C gp = a + b, gm = a - b;
    g[1 + 2 * d] = gp + om1 * (g[1 + 2 * d] - gp);
    g[2 + 2 * d] = gm + om1 * (g[2 + 2 * d] - gm);
  }
}

// The D3Q7 wall-slot values of a wall cell after its collision.
template <class S>
__device__ __forceinline__ void g_wall_slots(const typename Store<S>::C (&g)[Q7], uint8_t mk, uint32_t idx,
                                             uint32_t lk, const KArgs<typename Store<S>::C> &a,
                                             const Q7Args<typename Store<S>::C> &t, S *const (&out)[Q7]) {
  using C = typename Store<S>::C;
  uint32_t x = 0, y = 0, z = 0;
  if (mk & M_WRAP) coords(a.g, idx, x, y, z);
  for_n<Q7>([&](auto M) {
    constexpr int m = decltype(M)::value;
    if constexpr (m > 0) {
      if (lk & (1u << m)) {
        const uint32_t s = (mk & M_WRAP) ? wrap_src(a.g, x, y, z, CX[m], CY[m], CZ[m]) : idx - pull_off<m>(a.g);
        C v = g[OPP[m]];
        if (lk & (1u << (DIRICHLET_BIT + m))) v = C(2 * W7[m]) * (t.T[s] - t.T0) - v;
        else if (lk & neumann_bit(m)) v = v + t.jn;
        g_st<S>(at(out[m], s * (uint32_t)sizeof(S)), v);
      }
    }
  });
}

// Streaming of f and g, MacroProp (rho, u, T, q), and optionally the next
// collision of both; the structure of stream_cell (lbm_step.h).
template <class S, int FORCE, bool BQ, bool COLLIDE, bool EMIT>
__global__ __launch_bounds__(256) void k_q7(const Planes<S> P, const G7<S> G, const uint8_t *__restrict__ mask,
                                            KArgs<typename Store<S>::C> a, Q7Args<typename Store<S>::C> t) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  C f[Q], g[Q7];
  bulk_loads_wrapped<S>(P, a.g, idx, f);
  wrapped_loads<S, Q7>(G.src, a.g, idx, g);
  uint32_t mk = mask[idx];
  uint32_t lk = 0;
  if (a.links) {
    lk = a.links[idx];
    after_loads2(mk, lk);
  }
  mk = after_loads(mk, f);
  mk = after_loads7(mk, g);
  g_units<S>(g);
  uint32_t st = ALL19;
  if (mk & M_FLUID) {
    if (mk & M_BACKEQ) patch_cell<S, true>(P, (uint8_t)mk, idx, a, f);
    const C dT = g_sum(g);
    const C T = t.T0 + dT;
    C rho, drho, Fx, Fy, Fz, ux, uy, uz;
    int bad = 0;
    if constexpr (FORCE == 3) {
      // the IBM step (no MacroProp): the box's rho, u, F, unforced moments
      // elsewhere; the buoyancy of this cell's T added on top
      cell_state<S, 3>(f, (uint8_t)mk, idx, a, rho, drho, ux, uy, uz, Fx, Fy, Fz);
      if constexpr (BQ) {
        const C b = t.g_beta * (T - t.T_ref);
        Fy = Fy + b;
        uy = uy + C(0.5) * b / rho;
      }
    } else {
      C jx, jy, jz;
      moments<S>(f, rho, drho, jx, jy, jz);
      q7_force<C, FORCE, BQ>(a, t, idx, T, Fx, Fy, Fz);
      if constexpr (FORCE != 0 || BQ) velocity(jx + C(0.5) * Fx, jy + C(0.5) * Fy, jz + C(0.5) * Fz, rho, ux, uy, uz);
      else velocity(jx, jy, jz, rho, ux, uy, uz);
      bad = (rho < C(0) || rho > C(100)) ? ST_RHO_RANGE : 0;
      if (!(fabs(T) <= C(100)) || beyond_storage<S>(dT)) bad |= ST_T_RANGE;
    }
    if constexpr (EMIT) {
      a.rho[idx] = rho;
      a.ux[idx] = ux;
      a.uy[idx] = uy;
      a.uz[idx] = uz;
      t.T[idx] = T;
      t.qx[idx] = g[1] - g[2];
      t.qy[idx] = g[3] - g[4];
      t.qz[idx] = g[5] - g[6];
    }
    if constexpr (COLLIDE) {
      if (relax<S, (FORCE != 0 || BQ) ? 2 : 0>(f, rho, drho, ux, uy, uz, Fx, Fy, Fz, a.omega)) bad |= ST_FEQ_NEG;
      g_relax(g, T, dT, ux, uy, uz, t.omega_g);
    }
    if (bad) flag(a.status, a.first_bad, bad, idx);
  } else {
    // solids and row padding fill the lines they share with fluid cells
    // (stream_cell); other types keep their frozen-link values
    st = a.links ? (lk & ALL19) : (mk == 0 ? ALL19 : 0u);
    if (!st) return;
    for (int m = 0; m < Q; m++) f[m] = C(0);
    for (int m = 0; m < Q7; m++) g[m] = C(0);
  }
  const uint32_t boff = idx * (uint32_t)sizeof(S);
  if (__builtin_amdgcn_ballot_w64(st != ALL19) == 0) {
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      Store<S>::st(at(P.out[m], boff), f[m]);
    });
    for_n<Q7>([&](auto M) {
      constexpr int m = decltype(M)::value;
      g_st<S>(at(G.out[m], boff), g[m]);
    });
  } else {
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      if (st & (1u << m)) Store<S>::st(at(P.out[m], boff), f[m]);
    });
    for_n<Q7>([&](auto M) {
      constexpr int m = decltype(M)::value;
      if (st & (1u << m)) g_st<S>(at(G.out[m], boff), g[m]);
    });
  }
  if constexpr (COLLIDE) {
    if ((mk & (M_FLUID | M_WALL)) == (M_FLUID | M_WALL)) {
      wall_slot_stores<S>(f, (uint8_t)mk, idx, lk, a, P.out);
      g_wall_slots<S>(g, (uint8_t)mk, idx, lk, a, t, G.out);
    }
  }
}

// Collision in place with the stored rho, u and T (lbm_collide), wall slots
// included.
template <class S, int FORCE, bool BQ>
__global__ __launch_bounds__(256) void k_q7_collide(S *__restrict__ buf, const uint8_t *__restrict__ mask,
                                                    KArgs<typename Store<S>::C> a,
                                                    Q7Args<typename Store<S>::C> t) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  const uint8_t mk = mask[idx];
  if (!(mk & M_FLUID)) return;
  const uint32_t ps = a.g.ps;
  C f[Q], g[Q7];
  for (int m = 0; m < Q; m++) f[m] = Store<S>::ld(buf + (uint32_t)m * ps + idx);
  for (int m = 0; m < Q7; m++) g[m] = Store<S>::ld(buf + (uint32_t)(Q + m) * ps + idx);
  g_units<S>(g);
  C rho, drho, ux, uy, uz, T, Fx, Fy, Fz;
  if constexpr (FORCE == 3) {
    // the IBM step: as k_collide's mode 3, with T from the cell's own g
    // and its buoyancy added; rho, u, T, q are emitted
    cell_state<S, 3>(f, mk, idx, a, rho, drho, ux, uy, uz, Fx, Fy, Fz);
    T = t.T0 + g_sum(g);
    if constexpr (BQ) {
      const C b = t.g_beta * (T - t.T_ref);
      Fy = Fy + b;
      uy = uy + C(0.5) * b / rho;
    }
    a.rho[idx] = rho;
    a.ux[idx] = ux;
    a.uy[idx] = uy;
    a.uz[idx] = uz;
    t.T[idx] = T;
    t.qx[idx] = g[1] - g[2];
    t.qy[idx] = g[3] - g[4];
    t.qz[idx] = g[5] - g[6];
  } else {
    rho = a.rho[idx];
    drho = rho - C(1);
    ux = a.ux[idx];
    uy = a.uy[idx];
    uz = a.uz[idx];
    T = t.T[idx];
    q7_force<C, FORCE, BQ>(a, t, idx, T, Fx, Fy, Fz);
  }
  const bool neg = relax<S, (FORCE != 0 || BQ) ? 2 : 0>(f, rho, drho, ux, uy, uz, Fx, Fy, Fz, a.omega);
  g_relax(g, T, T - t.T0, ux, uy, uz, t.omega_g);
  for (int m = 0; m < Q; m++) Store<S>::st(buf + (uint32_t)m * ps + idx, f[m]);
  for (int m = 0; m < Q7; m++) g_st<S>(buf + (uint32_t)(Q + m) * ps + idx, g[m]);
  if (mk & M_WALL) {
    const uint32_t lk = a.links[idx];
    S *fp[Q], *gp[Q7];
    for (int m = 0; m < Q; m++) fp[m] = buf + (size_t)m * ps;
    for (int m = 0; m < Q7; m++) gp[m] = buf + (size_t)(Q + m) * ps;
    wall_slot_stores<S>(f, mk, idx, lk, a, fp);
    g_wall_slots<S>(g, mk, idx, lk, a, t, gp);
  }
  if (neg) flag(a.status, a.first_bad, ST_FEQ_NEG, idx);
}

// MacroProp of the fluid cells from the stored f and g (lbm_macro): T first,
// then the force, then u.
template <class S, int FORCE, bool BQ>
__global__ __launch_bounds__(256) void k_q7_macro(const S *__restrict__ buf, const uint8_t *__restrict__ mask,
                                                  KArgs<typename Store<S>::C> a,
                                                  Q7Args<typename Store<S>::C> t) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  if (!(mask[idx] & M_FLUID)) return;
  const uint32_t ps = a.g.ps;
  C f[Q], g[Q7];
  for (int m = 0; m < Q; m++) f[m] = Store<S>::ld(buf + (uint32_t)m * ps + idx);
  for (int m = 0; m < Q7; m++) g[m] = Store<S>::ld(buf + (uint32_t)(Q + m) * ps + idx);
  g_units<S>(g);
  const C T = t.T0 + g_sum(g);
  C rho, drho, jx, jy, jz, Fx, Fy, Fz, ux, uy, uz;
  moments<S>(f, rho, drho, jx, jy, jz);
  q7_force<C, FORCE, BQ>(a, t, idx, T, Fx, Fy, Fz);
  velocity(jx + C(0.5) * Fx, jy + C(0.5) * Fy, jz + C(0.5) * Fz, rho, ux, uy, uz);
  a.rho[idx] = rho;
  a.ux[idx] = ux;
  a.uy[idx] = uy;
  a.uz[idx] = uz;
  t.T[idx] = T;
  t.qx[idx] = g[1] - g[2];
  t.qy[idx] = g[3] - g[4];
  t.qz[idx] = g[5] - g[6];
  int bad = (rho < C(0) || rho > C(100)) ? ST_RHO_RANGE : 0;
  if (!(fabs(T) <= C(100)) || beyond_storage<S>(T - t.T0)) bad |= ST_T_RANGE;
  if (bad) flag(a.status, a.first_bad, bad, idx);
}

// g = g* = geq(T, u) in every cell, both buffers.  A cell that never
// collides and is not solid stores in slot m the Init value of its fluid
// neighbour x + c_m (the frozen link, as k_init does for f).
template <class S>
__global__ __launch_bounds__(256) void k_q7_init(S *__restrict__ A, S *__restrict__ B,
                                                 const uint8_t *__restrict__ mask, KArgs<typename Store<S>::C> a,
                                                 Q7Args<typename Store<S>::C> t) {
  using C = typename Store<S>::C;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.g.ncell) return;
  const uint8_t mk = mask[idx];
  uint32_t x, y, z;
  coords(a.g, idx, x, y, z);
  const uint32_t ps = a.g.ps;
  for_n<Q7>([&](auto M) {
    constexpr int m = decltype(M)::value;
    uint32_t c = idx;
    if (mk & M_OTHER) {
      const uint32_t d = wrap_dst(a.g, x, y, z, CX[m], CY[m], CZ[m]);
      if (mask[d] & M_FLUID) c = d;
    }
    const C T = t.T[c];
    const C cu = C(CX[m]) * a.ux[c] + C(CY[m]) * a.uy[c] + C(CZ[m]) * a.uz[c];
    const C v = C(W7[m]) * (T - t.T0) + C(W7[m] * 4.0) * T * cu;
    const S e = Store<S>::enc(v * Store<S>::K);
    A[(uint32_t)(Q + m) * ps + idx] = e;
    B[(uint32_t)(Q + m) * ps + idx] = e;
  });
}

// Dirichlet bits of the fluid cells' link words (after k_tables).
__global__ __launch_bounds__(256) void k_q7_tables(const int16_t *__restrict__ type,
                                                   const int16_t *__restrict__ typeE, const uint8_t *__restrict__ mask,
                                                   uint32_t *__restrict__ links, Geo g) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell) return;
  if ((mask[idx] & (M_FLUID | M_WALL)) != (M_FLUID | M_WALL)) return;
  uint32_t x, y, z;
  coords(g, idx, x, y, z);
  uint32_t lk = links[idx] & ((1u << DIRICHLET_BIT) - 1);
  for (int m = 1; m < Q7; m++) {
    const uint32_t s = wrap_src(g, x, y, z, CX[m], CY[m], CZ[m]);
    if (type[s] == LBM_TYPE_S && typeE[s] == LBM_TYPE_T) lk |= 1u << (DIRICHLET_BIT + m);
    if (type[s] == LBM_TYPE_S && typeE[s] == LBM_TYPE_Q) lk |= neumann_bit(m);
  }
  links[idx] = lk;
}

// g as FP64 [idx][7], unshifted
template <class S>
__global__ __launch_bounds__(256) void k_g7_out(const S *__restrict__ buf, Geo g, double T0, double *__restrict__ out) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ncell) return;
  using C = typename Store<S>::C;
  C v[Q7];
  for (int m = 0; m < Q7; m++) v[m] = Store<S>::ld(buf + (uint32_t)(Q + m) * g.ps + idx);
  g_units<S>(v);
  for (int m = 0; m < Q7; m++) out[(size_t)idx * Q7 + m] = (double)v[m] + W7[m] * T0;
}

// ---- host -----------------------------------------------------------------

template <class S>
Planes<S> planes_of(const lbm_ctx *c, const void *in, void *out) {
  Planes<S> P;
  const size_t ps = c->g.ps;
  for (int m = 0; m < Q; m++) {
    const ptrdiff_t off = (ptrdiff_t)CX[m] + (ptrdiff_t)CY[m] * c->g.px + (ptrdiff_t)CZ[m] * c->g.plane;
    P.in[m] = (const S *)in + m * ps;
    P.src[m] = (const S *)in + (ptrdiff_t)(m * ps) - off;
    P.out[m] = (S *)out + m * ps;
  }
  return P;
}

template <class S>
G7<S> g7_of(const lbm_ctx *c, const void *in, void *out) {
  G7<S> G;
  const size_t ps = c->g.ps;
  for (int m = 0; m < Q7; m++) {
    const ptrdiff_t off = (ptrdiff_t)CX[m] + (ptrdiff_t)CY[m] * c->g.px + (ptrdiff_t)CZ[m] * c->g.plane;
    G.in[m] = (const S *)in + (Q + m) * ps;
    G.src[m] = (const S *)in + (ptrdiff_t)((Q + m) * ps) - off;
    G.out[m] = (S *)out + (Q + m) * ps;
  }
  return G;
}

template <class C>
KArgs<C> kargs7(const LaunchCtx &L) {
  KArgs<C> a;
  memset(&a, 0, sizeof a);
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
  a.wall_slots = 1;
  a.slot_links = L.links != nullptr;
  return a;
}

template <class C>
Q7Args<C> q7args(const lbm_ctx *c) {
  Q7Args<C> t;
  t.omega_g = (C)(1.0 / c->td.tau_g);
  t.g_beta = (C)(c->td.boussinesq ? c->td.g_beta : 0.0);
  t.T_ref = (C)c->td.T_ref;
  t.T0 = (C)c->T0;
  t.jn = (C)c->td.j_dot_n;
  t.T = (C *)c->T;
  t.qx = (C *)c->qv[0];
  t.qy = (C *)c->qv[1];
  t.qz = (C *)c->qv[2];
  return t;
}

inline dim3 grid_of(const lbm_ctx *c) { return dim3((c->g.ncell + c->block - 1) / c->block); }

template <class S, int FORCE, bool BQ>
void launch_fb(lbm_ctx *c, int mode) {
  using C = typename Store<S>::C;
  const LaunchCtx L = launch_ctx(c);
  const KArgs<C> a = kargs7<C>(L);
  const Q7Args<C> t = q7args<C>(c);
  const dim3 grid = grid_of(c), block(c->block);
  switch (mode) {
    case Q7_FUSED:
    case Q7_FUSED_EMIT:
    case Q7_STREAM:
    case Q7_STREAM_EMIT: {
      const Planes<S> P = planes_of<S>(c, c->cur, c->oth);
      const G7<S> G = g7_of<S>(c, c->cur, c->oth);
      if (mode == Q7_FUSED)
        hipLaunchKernelGGL((k_q7<S, FORCE, BQ, true, false>), grid, block, 0, c->stream, P, G, c->mask, a, t);
      else if (mode == Q7_FUSED_EMIT)
        hipLaunchKernelGGL((k_q7<S, FORCE, BQ, true, true>), grid, block, 0, c->stream, P, G, c->mask, a, t);
      else if (mode == Q7_STREAM)
        hipLaunchKernelGGL((k_q7<S, FORCE, BQ, false, false>), grid, block, 0, c->stream, P, G, c->mask, a, t);
      else
        hipLaunchKernelGGL((k_q7<S, FORCE, BQ, false, true>), grid, block, 0, c->stream, P, G, c->mask, a, t);
      break;
    }
    case Q7_COLLIDE:
      hipLaunchKernelGGL((k_q7_collide<S, FORCE, BQ>), grid, block, 0, c->stream, (S *)c->cur, c->mask, a, t);
      break;
    case Q7_MACRO:
      hipLaunchKernelGGL((k_q7_macro<S, FORCE, BQ>), grid, block, 0, c->stream, (const S *)c->cur, c->mask, a, t);
      break;
  }
}

template <class S, bool BQ>
void launch_b(lbm_ctx *c, int mode, int fm) {
  switch (fm) {
    case 1: launch_fb<S, 1, BQ>(c, mode); break;
    case 2: launch_fb<S, 2, BQ>(c, mode); break;
    case 3: launch_fb<S, 3, BQ>(c, mode); break;
    default: launch_fb<S, 0, BQ>(c, mode); break;
  }
}

template <class S>
void launch_s7(lbm_ctx *c, int mode, int fm) {
  if (c->td.boussinesq) launch_b<S, true>(c, mode, fm);
  else launch_b<S, false>(c, mode, fm);
}

template <class S>
void init_s7(lbm_ctx *c) {
  using C = typename Store<S>::C;
  const LaunchCtx L = launch_ctx(c);
  hipLaunchKernelGGL(k_q7_init<S>, grid_of(c), dim3(c->block), 0, c->stream, (S *)c->cur, (S *)c->oth, c->mask,
                     kargs7<C>(L), q7args<C>(c));
}

template <class S>
void g_out_s7(lbm_ctx *c, double *g_dev) {
  hipLaunchKernelGGL(k_g7_out<S>, grid_of(c), dim3(c->block), 0, c->stream, (const S *)c->cur, c->g, c->T0, g_dev);
}

int upload_field(lbm_ctx *c, void *dst, const std::vector<double> &h) {
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

int q7_launch(lbm_ctx *c, int mode, int force_mode) {
  const int fm = force_mode < 0 ? c->force_mode : force_mode;
  switch (c->storage) {
    case LBM_STORE_F64: launch_s7<double>(c, mode, fm); break;
    case LBM_STORE_F16S: launch_s7<__half>(c, mode, fm); break;
    case LBM_STORE_F16C: launch_s7<f16c_t>(c, mode, fm); break;
    default: launch_s7<float>(c, mode, fm); break;
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int q7_tables(lbm_ctx *c) {
  if (!c->links) return LBM_OK;
  hipLaunchKernelGGL(k_q7_tables, grid_of(c), dim3(c->block), 0, c->stream, c->type, c->typeE, c->mask, c->links,
                     c->g);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int q7_init(lbm_ctx *c) {
  if (int e = q7_tables(c)) return e;
  switch (c->storage) {
    case LBM_STORE_F64: init_s7<double>(c); break;
    case LBM_STORE_F16S: init_s7<__half>(c); break;
    case LBM_STORE_F16C: init_s7<f16c_t>(c); break;
    default: init_s7<float>(c); break;
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

int q7_download_g(lbm_ctx *c, double *g_dev) {
  switch (c->storage) {
    case LBM_STORE_F64: g_out_s7<double>(c, g_dev); break;
    case LBM_STORE_F16S: g_out_s7<__half>(c, g_dev); break;
    case LBM_STORE_F16C: g_out_s7<f16c_t>(c, g_dev); break;
    default: g_out_s7<float>(c, g_dev); break;
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// lbm_upload_thermal for LBM_THERMAL_D3Q7 (arguments checked by the caller)
int q7_upload(lbm_ctx *c, const lbm_thermal_desc *d, const double *T, const int16_t *type_energy) {
  if (!(d->tau_g > 0.5) || !std::isfinite(d->j_dot_n)) return LBM_ERR_ARG;
  if ((uint64_t)c->g.ps * 26ull >= (1ull << 32)) return LBM_ERR_UNSUPPORTED;
  const size_t n = c->g.ncell;
  const uint32_t nx = c->g.nx, ny = c->g.ny, nz = c->g.nz, px = c->g.px;
  std::vector<double> hT(n, 0.0);
  c->h_typeE.assign(n, (int16_t)LBM_TYPE_F);
  for (uint32_t i = 0; i < nx; i++)
    for (uint32_t j = 0; j < ny; j++)
      for (uint32_t k = 0; k < nz; k++) {
        const size_t r = ((size_t)i * ny + j) * nz + k, dd = ((size_t)k * ny + j) * px + i;
        if (T) hT[dd] = T[r];
        if (type_energy) c->h_typeE[dd] = type_energy[r];
      }
  thermal_free(c);
  c->td = *d;
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
  const size_t popbytes = c->esize * (size_t)c->g.ps * (Q + Q7);
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
  void **arr[] = {&c->T, &c->qv[0], &c->qv[1], &c->qv[2]};
  for (void **p : arr)
    if (hipMalloc(p, n * c->csize) != hipSuccess) return LBM_ERR_NOMEM;
  if (hipMalloc((void **)&c->typeE, n * 2) != hipSuccess) return LBM_ERR_NOMEM;
  HIPCHK(hipMemcpy(c->typeE, c->h_typeE.data(), n * 2, hipMemcpyHostToDevice));
  if (int e = upload_field(c, c->T, hT)) return e;
  for (int q = 0; q < 3; q++) HIPCHK(hipMemset(c->qv[q], 0, n * c->csize));
  c->thermal = 2;
  return LBM_OK;
}
