// lbm_faces.hip -- the face loops at the end of LBM::Streaming (lbm.cpp:
// 299-625) and MacroProp of the cells that are neither fluid nor solid
// (lbm.cpp:667-775), on gfx950.
//
// After the push the reference runs six loops over the faces of the padded
// lattice, one after the other: i = 0, i = Nx+1, j = Ny+1, j = 0, k = Nz+1,
// k = 0.  On each face a TYPE_OUT cell copies three populations of its inner
// neighbour, with the D2Q9 indices of the reference's code kept under D3Q19
// ({1,5,8}, {3,6,7}, {4,7,8}, {2,5,6}, {4,7,8}, {4,7,8}), and a TYPE_IN cell
// takes the second-order equilibrium of its own rho, u (those of the
// previous MacroProp) -- on the k = 0 face written into cell (i,j,Nz+1)
// instead (lbm.cpp:599-611).  MacroProp then takes rho, u of every cell from
// its populations.
//
// None of it reaches the fluid: fluid cells never pull from these cells
// (their links stay frozen).  It sets the populations and the rho, u that
// the reference holds for its inlet and outlet cells.  Here each cell of the
// other types (IN, OUT, P, L, T, Q, wherever it lies), and each solid that an
// outlet copies from, has a slot: its 19 populations and its rho, u in FP64
// arrays, direction-major (`of[l * nslot + k]`, `om[q * nslot + k]`).  Their
// plane entries keep carrying the frozen values of their fluid neighbours.
// So does a fluid cell at k = Nz+1 under an inlet at k = 0 that an outlet
// copies from: the k = 0 loop overwrites all its 19 populations, so its
// frozen links (from cells of the other types) hold that write, which its
// slot keeps, not their Init values.
//
// Per step, after the streaming launch that pulled from buffer `src`:
//   launch 1  faces i = 0 and i = Nx+1 (they write disjoint cells and read
//             none that the other writes);
//   launch 2  faces j = Ny+1 and j = 0 (likewise);
//   launch 3  per (i,j) column the k = Nz+1 loop, then the k = 0 loop, then
//             MacroProp of the column's two end cells; MacroProp of the
//             other slotted cells (an inlet that only its own i or j loop
//             writes takes its MacroProp in launch 1 or 2, right after its
//             equilibrium, so launch 3 is skipped when nothing else is due).
// A loop thus reads the values the earlier loops of this step wrote and the
// previous step's values of cells that later loops write, as the
// reference's order gives.  A fluid neighbour's population is its pulled
// value, read again from `src` where the streaming read it (plane l at
// x - c_l: a fluid cell's f*, a solid's bounce-back slot or a frozen value).
//
// The launches walk host-built lists of the inlet/outlet cells (16 or 32
// bytes per entry, read coalesced) in slot order, so the slotted arrays are
// read and written in runs.  MacroProp: every step for the inlets (their
// rho, u feed the next step's equilibrium and the k = 0 write), for the
// other cells on steps that emit rho, u (their values reach nothing else;
// their density check runs then).
//
// In IBM runs (main.cpp:197-211) MacroProp is not called: the IBM step takes
// rho = sum f, u = sum f c / rho of the TYPE_IN/TYPE_OUT cells with F = 0 and
// no range check (lbm.cpp:1262-1285, 1173-1193), which mode 3 does.
//
// ENERGY: the loops move g alike (outlet copies of g[a, b, c], inlet
// g = rho w T (1 + 3cu + 4.5cu^2 - 1.5uu) with the inlet's own T,
// lbm.cpp:309-313, 340-346, 613-619), and MacroProp adds T = sum g,
// q = sum g c, the temperature check and the Boussinesq F_y (lbm.cpp:
// 777-806).  The slots hold g after f and T, q after rho, u.  Every cell an
// outlet copies from has a slot: without wall slots a fluid cell's pulled
// values exist nowhere after the fused launch, so thermal_face_sources
// (lbm_thermal.hip) writes them into its slot, from the same buffer and
// tables, right before each pulling launch.  MacroProp of every slotted
// cell runs on every step that has one (F_y of an outlet enters its u).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_step.h"

using namespace lbm;

namespace {

constexpr int FACE_THREADS = 256;

// source codes of a face-loop entry
constexpr int32_t SRC_IN = -1;     // TYPE_IN: its own equilibrium
constexpr int32_t SRC_INERT = -2;  // another type: no loop action, MacroProp only
constexpr int32_t SRC_NONE = -3;   // not a cell of the other types
// MacroProp kinds of a listed cell
constexpr int32_t K_IN = 1, K_OUT = 2, K_INERT = 3;

// one x or y face cell: d (slot ds) takes from s (code ss, below), or an
// inlet's equilibrium (s = SRC_IN; then ss = 1 when no later loop writes the
// cell, so its MacroProp can follow in the same thread)
struct XYEntry {
  int32_t d, ds, s, ss;
};
// one (i,j) column: its k = Nz+1 and k = 0 end cells
struct ZEntry {
  int32_t top, ts, tsrc, tss;
  int32_t bot, bs, bsrc, bss;
};
// a cell's MacroProp (the other slotted cells)
struct MEntry {
  int32_t idx, slot, kind, pad;
};
// source slot code ss: >= 0 a slotted cell of another type or solid (its
// own value); -1 a fluid cell (its pulled value); <= -2 a fluid cell under
// the k = 0 inflow write with slot -ss-2 (pulled, or that write on a frozen
// link)

template <class S>
struct FaceArgs {
  using C = typename Store<S>::C;
  Geo g;
  const S *src;          // the source buffer of the last streaming launch
  const int16_t *type;   // cell types, padded order
  double *of;            // populations, direction-major (ENERGY: f rows, then g rows)
  double *om;            // rho, ux, uy, uz (ENERGY: then T, qx, qy, qz), component-major
  int32_t ns;            // slots
  C *rho, *ux, *uy, *uz;       // the context's macro arrays
  const C *Fx, *Fy, *Fz;       // force mode 1
  double Fu[3];                // force mode 2
  int mode;                    // force mode 0/1/2; 3: the IBM step's moments
  int *status;
  unsigned long long *first_bad;
  int th;                      // ENERGY
  C *T, *qx, *qy, *qz;         // ENERGY: the context's T, q
  C *Fyb;                      // ENERGY with the Boussinesq force: F_y, rewritten by MacroProp
  double g_beta, T_ref;
  const C *uwx, *uwy, *uwz;    // wall velocities (the folded links' moving-wall term)
};

// Population L of source cell s (slot code ss) as the face loops see it.
template <class S, int L>
__device__ __forceinline__ double pop(const FaceArgs<S> &a, uint32_t s, int32_t ss) {
  if (ss >= 0) return a.of[(size_t)L * a.ns + ss];
  uint32_t x, y, z;
  coords(a.g, s, x, y, z);
  if ((fold_mask(a.g, x, y, z) >> L) & 1u) {
    // a wall link folded into the pull: f*(s, opp L) from s's own plane
    // opp L plus the moving-wall term, the value the pull took
    double v = (double)Store<S>::ld(a.src + (size_t)OPP[L] * a.g.ps + s);
    if constexpr (Store<S>::shifted) v = v / (double)Store<S>::K + W[L];
    const double ucw = CX[L] * (double)a.uwx[s] + CY[L] * (double)a.uwy[s] + CZ[L] * (double)a.uwz[s];
    return v + 6.0 * W[L] * ucw;
  }
  const uint32_t q = wrap_src(a.g, x, y, z, CX[L], CY[L], CZ[L]);
  if (ss <= -2) {
    const int16_t tq = a.type[q];
    if (tq != LBM_TYPE_F && tq != LBM_TYPE_S) return a.of[(size_t)L * a.ns + (-ss - 2)];
  }
  double v = (double)Store<S>::ld(a.src + (size_t)L * a.g.ps + q);
  if constexpr (Store<S>::shifted) v = v / (double)Store<S>::K + W[L];
  return v;
}

// TYPE_OUT: f[D] of slot ds = f[D] of its inner neighbour s (ENERGY: g[D]
// too, from the source's slot, which every source has there)
template <class S, int D0, int D1, int D2>
__device__ __forceinline__ void out_copy(const FaceArgs<S> &a, int32_t ds, uint32_t s, int32_t ss) {
  const double v0 = pop<S, D0>(a, s, ss), v1 = pop<S, D1>(a, s, ss), v2 = pop<S, D2>(a, s, ss);
  a.of[(size_t)D0 * a.ns + ds] = v0;
  a.of[(size_t)D1 * a.ns + ds] = v1;
  a.of[(size_t)D2 * a.ns + ds] = v2;
  if (a.th && ss >= 0) {  // under ENERGY every outlet source has a slot
    const size_t g0 = (size_t)Q * a.ns;
    const double w0 = a.of[g0 + (size_t)D0 * a.ns + ss], w1 = a.of[g0 + (size_t)D1 * a.ns + ss],
                 w2 = a.of[g0 + (size_t)D2 * a.ns + ss];
    a.of[g0 + (size_t)D0 * a.ns + ds] = w0;
    a.of[g0 + (size_t)D1 * a.ns + ds] = w1;
    a.of[g0 + (size_t)D2 * a.ns + ds] = w2;
  }
}

// TYPE_IN: the second-order equilibrium of slot k's rho, u into slot kd
// (lbm.cpp:325-337), the reference's expression with cs2 = 1/3; ENERGY: g
// with slot k's T too (lbm.cpp:340-346)
template <class S>
__device__ __forceinline__ void in_feq2(const FaceArgs<S> &a, int32_t k, int32_t kd) {
#pragma clang fp contract(off)
  const double rho = a.om[k], ux = a.om[a.ns + k], uy = a.om[2 * (size_t)a.ns + k],
               uz = a.om[3 * (size_t)a.ns + k];
  const double T = a.th ? a.om[4 * (size_t)a.ns + k] : 0.0;
  const double cs2 = 1.0 / 3.0;
  const double u_dot_u = ux * ux + uy * uy + uz * uz;
  for_dirs([&](auto M) {
    constexpr int l = decltype(M)::value;
    const double c_dot_u = ux * CX[l] + uy * CY[l] + uz * CZ[l];
    const double br = 1 + c_dot_u / cs2 + (c_dot_u * c_dot_u) / (2 * cs2 * cs2) - u_dot_u / (2 * cs2);
    a.of[(size_t)l * a.ns + kd] = rho * W[l] * br;
    if (a.th) a.of[(size_t)(Q + l) * a.ns + kd] = rho * W[l] * T * br;
  });
}

// MacroProp of one cell of the other types (lbm.cpp:668-675, 772-775): the
// sums in the reference's l order.  IBM moments: inlets and outlets only.
template <class S>
__device__ __forceinline__ void other_macro(const FaceArgs<S> &a, uint32_t idx, int32_t k, int32_t kind) {
#pragma clang fp contract(off)
  using C = typename Store<S>::C;
  if (a.mode == 3 && kind == K_INERT) return;
  double rho = 0, jx = 0, jy = 0, jz = 0;
  for_dirs([&](auto M) {
    constexpr int l = decltype(M)::value;
    const double f = a.of[(size_t)l * a.ns + k];
    rho += f;
    jx += f * CX[l];
    jy += f * CY[l];
    jz += f * CZ[l];
  });
  double Fx = 0, Fy = 0, Fz = 0;
  if (a.mode == 1) {
    Fx = (double)a.Fx[idx];
    Fy = (double)a.Fy[idx];
    Fz = (double)a.Fz[idx];
  } else if (a.mode == 2) {
    Fx = a.Fu[0];
    Fy = a.Fu[1];
    Fz = a.Fu[2];
  }
  const double ux = (jx + 0.5 * Fx) / rho, uy = (jy + 0.5 * Fy) / rho, uz = (jz + 0.5 * Fz) / rho;
  a.om[k] = rho;
  a.om[a.ns + k] = ux;
  a.om[2 * (size_t)a.ns + k] = uy;
  a.om[3 * (size_t)a.ns + k] = uz;
  a.rho[idx] = (C)rho;
  a.ux[idx] = (C)ux;
  a.uy[idx] = (C)uy;
  a.uz[idx] = (C)uz;
  int bad = (a.mode != 3 && (rho < 0 || rho > 100)) ? ST_RHO_RANGE : 0;
  if (a.th) {
    // lbm.cpp:777-806
    double T = 0, qx = 0, qy = 0, qz = 0;
    for_dirs([&](auto M) {
      constexpr int l = decltype(M)::value;
      const double g = a.of[(size_t)(Q + l) * a.ns + k];
      T += g;
      qx += g * CX[l];
      qy += g * CY[l];
      qz += g * CZ[l];
    });
    a.om[4 * (size_t)a.ns + k] = T;
    a.om[5 * (size_t)a.ns + k] = qx;
    a.om[6 * (size_t)a.ns + k] = qy;
    a.om[7 * (size_t)a.ns + k] = qz;
    a.T[idx] = (C)T;
    a.qx[idx] = (C)qx;
    a.qy[idx] = (C)qy;
    a.qz[idx] = (C)qz;
    if (T < 0 || T > 100) bad |= ST_T_RANGE;
    if (a.Fyb) a.Fyb[idx] = (C)(a.g_beta * (T - a.T_ref));
  }
  if (bad) flag(a.status, a.first_bad, bad, idx);
}

// launch 1 (AXIS 0: entries [0, n0) on i = 0, then i = Nx+1) and launch 2
// (AXIS 1: j = Ny+1, then j = 0); macro != 0: MacroProp of the inlets that
// no later loop writes
template <class S, int AXIS>
__global__ __launch_bounds__(FACE_THREADS) void k_faces_xy(FaceArgs<S> a, const XYEntry *__restrict__ e, int n0,
                                                           int n, int macro) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const XYEntry v = e[t];
  if (v.s == SRC_IN) {
    in_feq2(a, v.ds, v.ds);
    if (macro && v.ss == 1) other_macro(a, (uint32_t)v.d, v.ds, K_IN);
  } else if (AXIS == 0) {
    if (t < n0) out_copy<S, 1, 5, 8>(a, v.ds, (uint32_t)v.s, v.ss);  // i = 0, lbm.cpp:303-306
    else out_copy<S, 3, 6, 7>(a, v.ds, (uint32_t)v.s, v.ss);         // i = Nx+1, lbm.cpp:357-360
  } else {
    if (t < n0) out_copy<S, 4, 7, 8>(a, v.ds, (uint32_t)v.s, v.ss);  // j = Ny+1, lbm.cpp:412-415
    else out_copy<S, 2, 5, 6>(a, v.ds, (uint32_t)v.s, v.ss);         // j = 0, lbm.cpp:467-470
  }
}

__device__ __forceinline__ bool macro_due(int macro, int32_t src) {
  if (src == SRC_NONE || macro == 0) return false;
  return macro == 2 || src == SRC_IN;
}
__device__ __forceinline__ int32_t kind_of(int32_t src) {
  return src == SRC_IN ? K_IN : src >= 0 ? K_OUT : K_INERT;
}

// launch 3: threads [0, nz) the columns, then [nz, nz + nm) the listed
// cells' MacroProp.  loops = 0: MacroProp alone.  macro: 0 none, 1 the
// inlets, 2 every cell (the listed cells come inlets first).
template <class S>
__global__ __launch_bounds__(FACE_THREADS) void k_faces_z(FaceArgs<S> a, const ZEntry *__restrict__ ze, int nz,
                                                          const MEntry *__restrict__ me, int nm, int loops,
                                                          int macro) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nz) {
    if (t - nz < nm) {
      const MEntry v = me[t - nz];
      other_macro(a, (uint32_t)v.idx, v.slot, v.kind);
    }
    return;
  }
  const ZEntry v = ze[t];
  if (loops) {
    if (v.tsrc >= 0) out_copy<S, 4, 7, 8>(a, v.ts, (uint32_t)v.tsrc, v.tss);  // k = Nz+1, lbm.cpp:518-570
    else if (v.tsrc == SRC_IN) in_feq2(a, v.ts, v.ts);
    if (v.bsrc >= 0) out_copy<S, 4, 7, 8>(a, v.bs, (uint32_t)v.bsrc, v.bss);  // k = 0, lbm.cpp:572-625
    else if (v.bsrc == SRC_IN && v.ts >= 0) in_feq2(a, v.bs, v.ts);
  }
  if (macro_due(macro, v.tsrc)) other_macro(a, (uint32_t)v.top, v.ts, kind_of(v.tsrc));
  if (macro_due(macro, v.bsrc)) other_macro(a, (uint32_t)v.bot, v.bs, kind_of(v.bsrc));
}

// Init (lbm.cpp:31-57): f = the cubic equilibrium of the set-up rho, u;
// ENERGY: g = Calc_geq (lbm.cpp:636-653), rho w T (1 + ...) with the cell's
// set-up T, for a solid rho w T_w with T_w the mean set-up T of its solid
// neighbours (Determine_Wall_Temperature, lbm.cpp:965-981, l = 0 included)
template <class S>
__global__ __launch_bounds__(FACE_THREADS) void k_faces_init(FaceArgs<S> a, const int32_t *__restrict__ cells) {
  using C = typename Store<S>::C;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.ns) return;
  const uint32_t idx = (uint32_t)cells[k];
  const double rho = (double)a.rho[idx], ux = (double)a.ux[idx], uy = (double)a.uy[idx], uz = (double)a.uz[idx];
  const double uu = ux * ux + uy * uy + uz * uz;
  double T = 0.0;
  const bool solid = a.type[idx] == LBM_TYPE_S;
  if (a.th) {
    if (solid) {
      uint32_t x, y, z;
      coords(a.g, idx, x, y, z);
      C s = C(0);
      int num = 0;
      for_dirs([&](auto M) {
        constexpr int l = decltype(M)::value;
        const uint32_t n = wrap_dst(a.g, x, y, z, CX[l], CY[l], CZ[l]);
        if (a.type[n] == LBM_TYPE_S) {
          s += a.T[n];
          num++;
        }
      });
      T = (double)(s / C(num));
    } else {
      T = (double)a.T[idx];
    }
  }
  for_dirs([&](auto M) {
    constexpr int l = decltype(M)::value;
    const double cu = ux * CX[l] + uy * CY[l] + uz * CZ[l];
    a.of[(size_t)l * a.ns + k] = rho * W[l] * (1.0 + eq_A(cu, uu));
    if (a.th) a.of[(size_t)(Q + l) * a.ns + k] = solid ? rho * W[l] * T : rho * W[l] * T * (1.0 + eq_A(cu, uu));
  });
  a.om[k] = rho;
  a.om[a.ns + k] = ux;
  a.om[2 * (size_t)a.ns + k] = uy;
  a.om[3 * (size_t)a.ns + k] = uz;
  if (a.th) {
    a.om[4 * (size_t)a.ns + k] = (double)a.T[idx];
    a.om[5 * (size_t)a.ns + k] = 0.0;
    a.om[6 * (size_t)a.ns + k] = 0.0;
    a.om[7 * (size_t)a.ns + k] = 0.0;
  }
}

// the populations of the slotted cells but the fluid ones into a download
// of f (ncell x 19)
__global__ __launch_bounds__(FACE_THREADS) void k_faces_f_out(const double *__restrict__ of,
                                                              const int32_t *__restrict__ cells,
                                                              const int16_t *__restrict__ type, int n,
                                                              double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * Q) return;
  const int k = t / Q, l = t - k * Q;
  if (type[cells[k]] == LBM_TYPE_F) return;
  out[(size_t)cells[k] * Q + l] = of[(size_t)l * n + k];
}

// ENERGY: the g of the slotted cells but the fluid ones into a download of g
__global__ __launch_bounds__(FACE_THREADS) void k_faces_g_out(const double *__restrict__ of,
                                                              const int32_t *__restrict__ cells,
                                                              const int16_t *__restrict__ type, int n,
                                                              double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * Q) return;
  const int k = t / Q, l = t - k * Q;
  if (type[cells[k]] == LBM_TYPE_F) return;
  out[(size_t)cells[k] * Q + l] = of[(size_t)(Q + l) * n + k];
}

template <class S>
FaceArgs<S> face_args(const lbm_ctx *c, int mode) {
  using C = typename Store<S>::C;
  FaceArgs<S> a;
  a.g = c->g;
  a.src = (const S *)c->oth;
  a.type = c->type;
  a.of = c->fof;
  a.om = c->fom;
  a.ns = c->fnslot;
  a.rho = (C *)c->mac[0];
  a.ux = (C *)c->mac[1];
  a.uy = (C *)c->mac[2];
  a.uz = (C *)c->mac[3];
  a.Fx = (const C *)c->F[0];
  a.Fy = (const C *)c->F[1];
  a.Fz = (const C *)c->F[2];
  for (int q = 0; q < 3; q++) a.Fu[q] = c->Fu[q];
  a.mode = mode;
  a.status = c->status;
  a.first_bad = c->first_bad;
  a.th = c->thermal == 1;
  a.T = a.th ? (C *)c->T : nullptr;
  a.qx = a.th ? (C *)c->qv[0] : nullptr;
  a.qy = a.th ? (C *)c->qv[1] : nullptr;
  a.qz = a.th ? (C *)c->qv[2] : nullptr;
  a.Fyb = a.th && c->td.boussinesq ? (C *)c->F[1] : nullptr;
  a.g_beta = c->td.g_beta;
  a.T_ref = c->td.T_ref;
  a.uwx = (const C *)c->uw[0];
  a.uwy = (const C *)c->uw[1];
  a.uwz = (const C *)c->uw[2];
  return a;
}

unsigned blocks(uint64_t n) { return (unsigned)((n + FACE_THREADS - 1) / FACE_THREADS); }

// The MacroProp list holds the inlets whose MacroProp the x/y launches do
// (fnm_xy of them) first, then the other inlets, then the other cells: a
// loop pass starts after the first group, MacroProp alone takes them all.
template <class S>
hipError_t faces_pass_s(lbm_ctx *c, int mode, int loops, int macro) {
  const FaceArgs<S> a = face_args<S>(c, mode);
  if (loops && c->fnx)
    hipLaunchKernelGGL((k_faces_xy<S, 0>), dim3(blocks(c->fnx)), dim3(FACE_THREADS), 0, c->stream, a,
                       (const XYEntry *)c->fxe, c->fnx0, c->fnx, macro);
  if (loops && c->fny)
    hipLaunchKernelGGL((k_faces_xy<S, 1>), dim3(blocks(c->fny)), dim3(FACE_THREADS), 0, c->stream, a,
                       (const XYEntry *)c->fye, c->fny0, c->fny, macro);
  const int first = loops ? c->fnm_xy : 0;
  const int nm = macro == 2 ? c->fnm - first : macro == 1 ? c->fnm_in - first : 0;
  // columns with loop work; on passes with every cell's MacroProp also the
  // columns whose ends only take MacroProp
  const int nz = (loops && macro != 2) ? c->fnz : c->fnz_macro;
  if (nz + nm > 0)
    hipLaunchKernelGGL(k_faces_z<S>, dim3(blocks((uint64_t)nz + nm)), dim3(FACE_THREADS), 0, c->stream, a,
                       (const ZEntry *)c->fze, nz, (const MEntry *)c->fme + first, nm, loops, macro);
  return hipGetLastError();
}

hipError_t dispatch_pass(lbm_ctx *c, int mode, int loops, int macro) {
  switch (c->storage) {
    case LBM_STORE_F64: return faces_pass_s<double>(c, mode, loops, macro);
    case LBM_STORE_F32: return faces_pass_s<float>(c, mode, loops, macro);
    case LBM_STORE_F16S: return faces_pass_s<__half>(c, mode, loops, macro);
    default: return faces_pass_s<f16c_t>(c, mode, loops, macro);
  }
}

template <class S>
hipError_t faces_init_s(lbm_ctx *c) {
  const FaceArgs<S> a = face_args<S>(c, 0);
  hipLaunchKernelGGL(k_faces_init<S>, dim3(blocks((uint64_t)c->fnslot)), dim3(FACE_THREADS), 0, c->stream, a,
                     c->fcells);
  return hipGetLastError();
}
hipError_t dispatch_init(lbm_ctx *c) {
  switch (c->storage) {
    case LBM_STORE_F64: return faces_init_s<double>(c);
    case LBM_STORE_F32: return faces_init_s<float>(c);
    case LBM_STORE_F16S: return faces_init_s<__half>(c);
    default: return faces_init_s<f16c_t>(c);
  }
}

template <class T>
bool upload_list(lbm_ctx *c, void **dst, const std::vector<T> &v) {
  if (v.empty()) return true;
  if (hipMalloc(dst, v.size() * sizeof(T)) != hipSuccess) return false;
  return hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream) == hipSuccess;
}

}  // namespace

void faces_free(lbm_ctx *c) {
  void *ptrs[] = {c->fslot, c->fof, c->fom, c->fcells, c->fxe, c->fye, c->fze, c->fme, c->fsrc};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->fslot = nullptr;
  c->fof = nullptr;
  c->fom = nullptr;
  c->fcells = nullptr;
  c->fsrc = nullptr;
  c->fxe = c->fye = c->fze = c->fme = nullptr;
  c->fnslot = c->fnx = c->fnx0 = c->fny = c->fny0 = c->fnz = c->fnz_macro = c->fnm = c->fnm_in = c->fnm_xy = 0;
  c->fnsrc = 0;
}

// Slots for the cells of the other types, for the solids that an outlet
// copies from and for the fluid cells under the k = 0 inflow write that an
// outlet copies from (ENERGY: for every cell an outlet copies from); their
// Init populations; the work lists.  Call after the set-up rho, u (and T)
// are in the macro arrays and before anything overwrites the solids'
// entries there.
int faces_setup(lbm_ctx *c) {
  faces_free(c);
  const bool th = c->thermal == 1;
  const lbm::Geo &g = c->g;
  const std::vector<int16_t> &t = c->h_type;
  auto at = [&](uint32_t x, uint32_t y, uint32_t z) { return (int32_t)((z * g.ny + y) * g.px + x); };
  auto other = [&](int32_t d) { return t[d] != LBM_TYPE_F && t[d] != LBM_TYPE_S; };
  auto io = [&](int32_t d) { return t[d] == LBM_TYPE_IN || t[d] == LBM_TYPE_OUT; };
  bool any = false;
  for (uint32_t z = 0; z < g.nz && !any; z++)
    for (uint32_t y = 0; y < g.ny && !any; y++)
      for (uint32_t x = 0; x < g.nx; x++)
        if (other(at(x, y, z))) {
          any = true;
          break;
        }
  if (!any) return LBM_OK;
  std::vector<int32_t> slot(g.ncell, -1), cells;
  auto give = [&](int32_t d) {
    if (slot[d] < 0) {
      slot[d] = (int32_t)cells.size();
      cells.push_back(d);
    }
  };
  // slots in the order the launches visit the cells: the face lists, then
  // the other cells of the other types
  std::vector<std::pair<int32_t, int32_t>> xl, yl;  // (d, s) per face cell
  int nx0 = 0, ny0 = 0;
  for (int side = 0; side < 2; side++) {
    for (uint32_t z = 0; z < g.nz; z++)
      for (uint32_t y = 0; y < g.ny; y++) {
        const int32_t d = at(side ? g.nx - 1 : 0, y, z), s = at(side ? g.nx - 2 : 1, y, z);
        if (io(d)) xl.push_back({d, s});
      }
    if (side == 0) nx0 = (int)xl.size();
  }
  for (int side = 0; side < 2; side++) {  // j = Ny+1 first
    for (uint32_t z = 0; z < g.nz; z++)
      for (uint32_t x = 0; x < g.nx; x++) {
        const int32_t d = at(x, side ? 0 : g.ny - 1, z), s = at(x, side ? 1 : g.ny - 2, z);
        if (io(d)) yl.push_back({d, s});
      }
    if (side == 0) ny0 = (int)yl.size();
  }
  for (auto &e : xl) give(e.first);
  for (auto &e : yl) give(e.first);
  for (uint32_t y = 0; y < g.ny; y++)
    for (uint32_t x = 0; x < g.nx; x++) {
      if (other(at(x, y, g.nz - 1))) give(at(x, y, g.nz - 1));
      if (other(at(x, y, 0))) give(at(x, y, 0));
    }
  for (uint32_t z = 1; z + 1 < g.nz; z++)
    for (uint32_t y = 0; y < g.ny; y++)
      for (uint32_t x = 0; x < g.nx; x++)
        if (other(at(x, y, z))) give(at(x, y, z));
  // outlet sources that need a slot: solids, and fluid cells at k = Nz+1
  // under an inlet at k = 0
  const int32_t top = (int32_t)((g.nz - 1) * g.plane);
  auto backeq_fluid = [&](int32_t s) { return t[s] == LBM_TYPE_F && s >= top && t[s - top] == LBM_TYPE_IN; };
  auto need_src = [&](int32_t d, int32_t s) {
    if (t[d] == LBM_TYPE_OUT && (th || t[s] == LBM_TYPE_S || backeq_fluid(s))) give(s);
  };
  for (auto &e : xl) need_src(e.first, e.second);
  for (auto &e : yl) need_src(e.first, e.second);
  for (uint32_t y = 0; y < g.ny; y++)
    for (uint32_t x = 0; x < g.nx; x++) {
      need_src(at(x, y, g.nz - 1), at(x, y, g.nz - 2));
      need_src(at(x, y, 0), at(x, y, 1));
    }
  auto ss_of = [&](int32_t s) {
    if (slot[s] < 0) return (int32_t)-1;
    return t[s] == LBM_TYPE_F && !th ? -slot[s] - 2 : slot[s];
  };
  auto src_code = [&](int32_t d, int32_t s) {
    if (!other(d)) return SRC_NONE;
    if (t[d] == LBM_TYPE_IN) return SRC_IN;
    if (t[d] == LBM_TYPE_OUT) return s;
    return SRC_INERT;
  };
  // an inlet on an i face off the j and k faces (or on a j face off the i
  // and k faces) is written by its own loop only: its MacroProp can follow
  // its equilibrium in the same thread
  std::vector<char> xy_final(g.ncell, 0);
  auto final_in = [&](int32_t d, bool on_x) {
    const uint32_t x = (uint32_t)d % g.px, y = ((uint32_t)d / g.px) % g.ny, z = (uint32_t)d / g.plane;
    if (t[d] != LBM_TYPE_IN || z == 0 || z == g.nz - 1) return false;
    return on_x ? (y != 0 && y != g.ny - 1) : (x != 0 && x != g.nx - 1);
  };
  std::vector<XYEntry> xe, ye;
  for (auto &e : xl) {
    const bool fin = final_in(e.first, true);
    if (fin) xy_final[e.first] = 1;
    const int32_t s = src_code(e.first, e.second);
    xe.push_back({e.first, slot[e.first], s, s == SRC_IN ? (int32_t)fin : ss_of(e.second)});
  }
  for (auto &e : yl) {
    const bool fin = final_in(e.first, false);
    if (fin) xy_final[e.first] = 1;
    const int32_t s = src_code(e.first, e.second);
    ye.push_back({e.first, slot[e.first], s, s == SRC_IN ? (int32_t)fin : ss_of(e.second)});
  }
  // columns: those with a loop action or a MacroProp first (a loop pass
  // takes fnz of them), then those with an other-type end and no loop
  // action, which only MacroProp alone visits
  std::vector<ZEntry> ze, ze_macro;
  for (uint32_t y = 0; y < g.ny; y++)
    for (uint32_t x = 0; x < g.nx; x++) {
      const int32_t tp = at(x, y, g.nz - 1), bt = at(x, y, 0);
      const int32_t tpi = at(x, y, g.nz - 2), bti = at(x, y, 1);
      if (slot[tp] < 0 && slot[bt] < 0) continue;
      const ZEntry z{tp, slot[tp], src_code(tp, tpi), ss_of(tpi), bt, slot[bt], src_code(bt, bti), ss_of(bti)};
      if (z.tsrc == SRC_NONE && z.bsrc == SRC_NONE) continue;  // slotted sources only: nothing to do
      const bool act = z.tsrc >= 0 || z.tsrc == SRC_IN || z.bsrc >= 0 || z.bsrc == SRC_IN;
      (act ? ze : ze_macro).push_back(z);
    }
  const int nz_loop = (int)ze.size();
  ze.insert(ze.end(), ze_macro.begin(), ze_macro.end());
  // MacroProp list: the other-type cells off the k faces; the inlets done by
  // the x/y launches, then the other inlets, then the rest
  std::vector<MEntry> me;
  for (int pass = 0; pass < 3; pass++)
    for (size_t k = 0; k < cells.size(); k++) {
      const int32_t d = cells[k];
      const uint32_t z = (uint32_t)d / g.plane;
      if (!other(d) || z == 0 || z == g.nz - 1) continue;
      const bool inl = t[d] == LBM_TYPE_IN;
      const int group = inl ? (xy_final[d] ? 0 : 1) : 2;
      if (group != pass) continue;
      me.push_back({d, (int32_t)k, inl ? K_IN : t[d] == LBM_TYPE_OUT ? K_OUT : K_INERT, 0});
    }
  int nm_in = 0, nm_xy = 0;
  for (auto &v : me) {
    nm_in += v.kind == K_IN;
    nm_xy += xy_final[v.idx];
  }
  // ENERGY: the fluid cells with a slot (outlet sources), whose pulled
  // values thermal_face_sources writes
  std::vector<int32_t> fsrc;
  if (th)
    for (size_t k = 0; k < cells.size(); k++)
      if (t[cells[k]] == LBM_TYPE_F) {
        fsrc.push_back(cells[k]);
        fsrc.push_back((int32_t)k);
      }
  const size_t ns = cells.size(), nrow = th ? 2 * Q : Q, nmac = th ? 8 : 4;
  if (ns > (size_t)INT32_MAX / nrow) return LBM_ERR_UNSUPPORTED;
  bool ok = hipMalloc((void **)&c->fslot, g.ncell * sizeof(int32_t)) == hipSuccess &&
            hipMalloc((void **)&c->fof, ns * nrow * sizeof(double)) == hipSuccess &&
            hipMalloc((void **)&c->fom, ns * nmac * sizeof(double)) == hipSuccess &&
            upload_list(c, (void **)&c->fsrc, fsrc) &&
            upload_list(c, (void **)&c->fcells, cells) && upload_list(c, &c->fxe, xe) &&
            upload_list(c, &c->fye, ye) && upload_list(c, &c->fze, ze) && upload_list(c, &c->fme, me) &&
            hipMemcpyAsync(c->fslot, slot.data(), g.ncell * sizeof(int32_t), hipMemcpyHostToDevice, c->stream) ==
                hipSuccess;
  if (!ok) {
    (void)hipStreamSynchronize(c->stream);
    faces_free(c);
    return LBM_ERR_NOMEM;
  }
  c->fnslot = (int)ns;
  c->fnx = (int)xe.size();
  c->fnx0 = nx0;
  c->fny = (int)ye.size();
  c->fny0 = ny0;
  c->fnz = nz_loop;
  c->fnz_macro = (int)ze.size();
  c->fnm = (int)me.size();
  c->fnm_in = nm_in;
  c->fnm_xy = nm_xy;
  c->fnsrc = (int)(fsrc.size() / 2);
  HIPCHK(dispatch_init(c));
  // the host vectors go out of scope: the copies must have landed
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

// The face loops after a streaming launch (the buffers already swapped:
// `oth` is the launch's source) and MacroProp of the slotted cells: macro 0
// none, 1 the inlets, 2 every cell.  ibm: the moments of an IBM run instead
// of MacroProp.
int faces_step(lbm_ctx *c, bool ibm, int macro) {
  if (!c->fnslot) return LBM_OK;
  if (c->thermal == 1 && macro == 1) macro = 2;  // ENERGY: every slotted cell's MacroProp, every step
  HIPCHK(dispatch_pass(c, ibm ? 3 : c->force_mode, 1, macro));
  return LBM_OK;
}

int faces_macro(lbm_ctx *c, bool ibm) {
  if (!c->fnslot) return LBM_OK;
  HIPCHK(dispatch_pass(c, ibm ? 3 : c->force_mode, 0, 2));
  return LBM_OK;
}

hipError_t faces_f_out(lbm_ctx *c, double *out) {
  if (!c->fnslot) return hipSuccess;
  hipLaunchKernelGGL(k_faces_f_out, dim3(blocks((uint64_t)c->fnslot * Q)), dim3(FACE_THREADS), 0, c->stream,
                     c->fof, c->fcells, c->type, c->fnslot, out);
  return hipGetLastError();
}

hipError_t faces_g_out(lbm_ctx *c, double *out) {
  if (!c->fnslot || c->thermal != 1) return hipSuccess;
  hipLaunchKernelGGL(k_faces_g_out, dim3(blocks((uint64_t)c->fnslot * Q)), dim3(FACE_THREADS), 0, c->stream,
                     c->fof, c->fcells, c->type, c->fnslot, out);
  return hipGetLastError();
}
