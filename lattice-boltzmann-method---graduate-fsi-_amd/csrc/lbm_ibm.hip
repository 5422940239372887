// lbm_ibm.hip -- the implicit diffuse-interface immersed-boundary step
// (IBLBM, lbm.cpp:1135-1417, branch Implicit_Diffuse_Interface_Scheme) on
// gfx950, generalised from the reference's 2D slab (box z = {1}) to a 3D box.
//
// Per step:
//   k_ibm_move        trapezoidal marker motion, u_des = V_CM (lbm.cpp:1154-1161)
//   k_ibm_box_moments unforced rho, u = sum f c / rho, F = 0 of the box cells
//                     (lbm.cpp:1262-1285), pulled from f* on the fused path;
//                     per-workgroup sums of the interior density
//   k_ibm_interp      one wavefront per marker, one lane per point of its
//                     4x4x4 stencil: lane terms u phi phi phi, summed by
//                     one lane per component in the reference's i, j, k
//                     order (lbm.cpp:1288-1309, 1363-1387); then the marker
//                     force of the next pass, F_b = 2 rho (u_des - u_b)/dt
//                     with rho at the box point farthest along the marker
//                     normal, and the marker's running drag/lift
//                     (lbm.cpp:1314-1329)
//   m = 1..m_max:
//     k_ibm_spread    Fm = sum F_b ds phi phi phi per box cell, F -= Fm,
//                     u += Fm/(2 rho) (lbm.cpp:1332-1360).  Atomic-free: the
//                     (marker, stencil point) pairs are binned to their cells
//                     once per marker position, sorted by marker; a workgroup
//                     takes a chunk of cells, computes their pairs' terms in
//                     parallel into LDS, and each cell sums its own terms
//                     serially, in the reference's marker order
//     k_ibm_interp
//   k_ibm_finish      FP64 drag, lift (tree sums of the per-marker sums), the
//                     mean density of the interior box cells and the
//                     added-mass term (lbm.cpp:1105-1132, 1390-1398)
// The collision then takes rho, u, F from the box arrays (force mode 3).
// The stencil weights phi(i - x) of every marker (4 per axis) are computed
// once per marker position (k_ibm_weights).
//
// The marker arithmetic is FP64 with a*b+c contraction off, so each product
// and sum over the stencil rounds as the reference's C++ does.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_step.h"

#pragma clang fp contract(off)

using namespace lbm;

// the filament's device state and step (lbm_filament.hip)
void filament_state(lbm_filament *f, double **x, double **xp, int *n, double *ds, int *dev);
// w, out: a coupled step's output (total force with the node weights w, the
// free end's position), or null
int filament_step_dev(lbm_filament *f, const double *F_dev, double dt, int32_t ncalls, hipStream_t s,
                      const double *w, double *out);

namespace {

constexpr int CHUNK_E = 1024;  // (marker, point) pairs per spreading workgroup
constexpr int CHUNK_C = 256;   // cells per spreading workgroup
constexpr int FIN_THREADS = 1024;
// markers (one wavefront each) per workgroup of the marker-space kernels:
// one 1024-thread workgroup per CU for a 4096-marker body dispatches
// faster than four of 256 (0.2008 against 0.2044 ms per IBM channel step)
constexpr int MK_WG = 16;

struct Box {
  int x0, x1, y0, y1, z0, z1;  // inclusive, reference indices (i,j,k)
  int nx, ny, nz;
  __host__ __device__ int ncell() const { return nx * ny * nz; }
};

// kernel(r), lbm.cpp:1095-1102
__device__ __forceinline__ double peskin(double r) {
  double d;
  const double a = fabs(r);
  if (a >= 0 && a <= 1) d = (3 - 2 * a + sqrt(1 + 4 * a - 4 * (r * r))) / 8;
  else if (a >= 1 && a <= 2) d = (5 - 2 * a - sqrt(-7 + 12 * a - 4 * (r * r))) / 8;
  else d = 0;
  return d;
}

__device__ __forceinline__ uint32_t lattice_idx(const Geo &g, int x, int y, int z) {
  return ((uint32_t)z * g.ny + (uint32_t)y) * g.px + (uint32_t)x;
}

// ---- set-up kernels -----------------------------------------------------------
__global__ void k_ibm_mark(uint8_t *__restrict__ mask, Geo g, Box B, uint8_t *__restrict__ live,
                           int set) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.ncell()) return;
  const int x = B.x0 + b % B.nx, y = B.y0 + (b / B.nx) % B.ny, z = B.z0 + b / (B.nx * B.ny);
  const uint32_t idx = lattice_idx(g, x, y, z);
  if (set) {
    mask[idx] |= M_BOX;
    live[b] = (mask[idx] & M_FLUID) ? 1 : 0;
  } else {
    mask[idx] &= (uint8_t)~M_BOX;
  }
}

// NearestLatticeCoordinate, lbm.cpp:1074-1092: the box point with the
// largest positive projection on the marker normal, scanned x, y, z (z
// fastest) with ">=": among equal maxima the last one in scan order.
__global__ __launch_bounds__(256) void k_ibm_nearest(Box B, const double *__restrict__ mx,
                                                     const double *__restrict__ my,
                                                     const double *__restrict__ mz,
                                                     const double *__restrict__ nx,
                                                     const double *__restrict__ ny,
                                                     const double *__restrict__ nz,
                                                     int32_t *__restrict__ near) {
  __shared__ double sd[256];
  __shared__ int ss[256];
  const int n = blockIdx.x;
  const double px = mx[n], py = my[n], pz = mz[n], qx = nx[n], qy = ny[n], qz = nz[n];
  double best = 0.0;
  int bs = -1;
  const int total = B.ncell();
  for (int s = threadIdx.x; s < total; s += blockDim.x) {
    // scan order s = ((x-x0)*ny + (y-y0))*nz + (z-z0)
    const int zz = s % B.nz, yy = (s / B.nz) % B.ny, xx = s / (B.nz * B.ny);
    const double v0 = (B.x0 + xx) - px, v1 = (B.y0 + yy) - py, v2 = (B.z0 + zz) - pz;
    const double d = v0 * qx + v1 * qy + v2 * qz;
    if (d > 0 && (bs < 0 || d > best || (d == best && s > bs))) {
      best = d;
      bs = s;
    }
  }
  sd[threadIdx.x] = best;
  ss[threadIdx.x] = bs;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double d2 = sd[threadIdx.x + w];
      const int s2 = ss[threadIdx.x + w];
      const int s1 = ss[threadIdx.x];
      if (s2 >= 0 && (s1 < 0 || d2 > sd[threadIdx.x] || (d2 == sd[threadIdx.x] && s2 > s1))) {
        sd[threadIdx.x] = d2;
        ss[threadIdx.x] = s2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int s = ss[0];
    if (s < 0) {
      near[n] = -1;
    } else {
      const int zz = s % B.nz, yy = (s / B.nz) % B.ny, xx = s / (B.nz * B.ny);
      near[n] = xx + B.nx * (yy + B.ny * zz);  // box index, x fastest
    }
  }
}

// The explicit branch's density point (lbm.cpp:1216-1222): NearestLatticeCoordinate
// over the marker's floor/ceil neighbourhood (z: floor(z)..ceil(z), the 2D
// call's 1..1 for markers at z = 1), starting from cell (0, 0, 1), which
// stays when no point lies ahead along the normal.  Box index, or -1 when
// the point is outside the box (the force is then NaN).
__global__ void k_ibm_near_local(Box B, int nm, const double *__restrict__ mx, const double *__restrict__ my,
                                 const double *__restrict__ mz, const double *__restrict__ nx,
                                 const double *__restrict__ ny, const double *__restrict__ nz,
                                 int32_t *__restrict__ near) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  const double px = mx[n], py = my[n], pz = mz[n], qx = nx[n], qy = ny[n], qz = nz[n];
  int i = 0, j = 0, k = 1;
  double dot = 0;
  for (int x = (int)floor(px); x <= (int)ceil(px); x++)
    for (int y = (int)floor(py); y <= (int)ceil(py); y++)
      for (int z = (int)floor(pz); z <= (int)ceil(pz); z++) {
        const double d = (x - px) * qx + (y - py) * qy + (z - pz) * qz;
        if (d >= dot && d > 0) {
          i = x;
          j = y;
          k = z;
          dot = d;
        }
      }
  const bool in = i >= B.x0 && i <= B.x1 && j >= B.y0 && j <= B.y1 && k >= B.z0 && k <= B.z1;
  near[n] = in ? (i - B.x0) + B.nx * ((j - B.y0) + B.ny * (k - B.z0)) : -1;
}

// Calc_AverageDensity's selection, lbm.cpp:1110-1124: box cells (any type)
// with (p - x_n).n_n <= 0 for every marker.
__global__ void k_ibm_inside(Box B, int nm, const double *__restrict__ mx,
                             const double *__restrict__ my, const double *__restrict__ mz,
                             const double *__restrict__ nx, const double *__restrict__ ny,
                             const double *__restrict__ nz, uint8_t *__restrict__ inside) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.ncell()) return;
  const double x = B.x0 + b % B.nx, y = B.y0 + (b / B.nx) % B.ny, z = B.z0 + b / (B.nx * B.ny);
  uint8_t in = 1;
  for (int n = 0; n < nm; n++) {
    const double v0 = x - mx[n], v1 = y - my[n], v2 = z - mz[n];
    if (v0 * nx[n] + v1 * ny[n] + v2 * nz[n] > 0) {
      in = 0;
      break;
    }
  }
  inside[b] = in;
}

// ---- per-step kernels ----------------------------------------------------------
// marker update, lbm.cpp:1154-1161 (u_des is the kernel argument vcm)
__global__ void k_ibm_move(int nm, double *__restrict__ mx, double *__restrict__ my,
                           double *__restrict__ mz, double v0, double v1, double v2, double b0,
                           double b1, double b2) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  const double dt = 1.0;
  mx[n] += (v0 + b0) * dt / 2;
  my[n] += (v1 + b1) * dt / 2;
  mz[n] += (v2 + b2) * dt / 2;
}

struct BoxArrays {
  double *rho, *ux, *uy, *uz, *Fx, *Fy, *Fz;
  const uint8_t *live;
};

// Per-marker stencil: base point (floor(x) - 1, ...) and the 4 weights per
// axis phi(i - x), i = base .. base + 3 (kernel(r), lbm.cpp:1095-1102, with
// r = (i dx - x)/dx, dx = 1).
struct Stencils {
  int32_t *base;  // 3 per marker
  double *w;      // 12 per marker: wx[4], wy[4], wz[4]
};

__global__ void k_ibm_weights(int nm, const double *__restrict__ mx, const double *__restrict__ my,
                              const double *__restrict__ mz, Stencils st) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  const double p[3] = {mx[n], my[n], mz[n]};
  for (int a = 0; a < 3; a++) {
    const int b = (int)floor(p[a]) - 1;
    st.base[3 * n + a] = b;
    for (int q = 0; q < 4; q++) st.w[12 * n + 4 * a + q] = peskin(((b + q) * 1.0 - p[a]) / 1.0);
  }
}

// ---- binning of (marker, stencil point) pairs to box cells ---------------------
// pidx[64 n + s] = box index of stencil point s of marker n when it lies in
// the box on a fluid cell, else -1; cnt[b] = pairs on box cell b.
__global__ void k_ibm_bin_count(Box B, int nm, Stencils st, const uint8_t *__restrict__ live,
                                int32_t *__restrict__ cnt, int32_t *__restrict__ pidx) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nm * 64) return;
  const int n = t >> 6, s = t & 63;
  const int i = st.base[3 * n + 0] + (s >> 4), j = st.base[3 * n + 1] + ((s >> 2) & 3),
            k = st.base[3 * n + 2] + (s & 3);
  int b = -1;
  if (i >= B.x0 && i <= B.x1 && j >= B.y0 && j <= B.y1 && k >= B.z0 && k <= B.z1) {
    b = (i - B.x0) + B.nx * ((j - B.y0) + B.ny * (k - B.z0));
    if (live[b]) atomicAdd(&cnt[b], 1);
    else b = -1;
  }
  pidx[t] = b;
}

// pairs into their cell's segment (any order; k_ibm_bin_sort orders them)
__global__ void k_ibm_bin_fill(int nm, const int32_t *__restrict__ pidx, const int32_t *__restrict__ bandpos,
                               const int32_t *__restrict__ bstart, int32_t *__restrict__ cursor,
                               int32_t *__restrict__ ent) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nm * 64) return;
  const int b = pidx[t];
  if (b < 0) return;
  ent[bstart[bandpos[b]] + atomicAdd(&cursor[b], 1)] = t;
}

// One thread per band cell: sort its pairs by marker (a marker has at most
// one point on a cell), then replace each pair by its marker index and
// store the pair's three 1D weights.
__global__ void k_ibm_bin_sort(int nband, const int32_t *__restrict__ bstart, Stencils st,
                               int32_t *__restrict__ ent, double *__restrict__ ew) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nband) return;
  const int e0 = bstart[i], e1 = bstart[i + 1];
  for (int e = e0 + 1; e < e1; e++) {
    const int32_t v = ent[e];
    int q = e - 1;
    while (q >= e0 && ent[q] > v) {
      ent[q + 1] = ent[q];
      q--;
    }
    ent[q + 1] = v;
  }
  for (int e = e0; e < e1; e++) {
    const int t = ent[e], n = t >> 6, s = t & 63;
    ew[3 * (size_t)e + 0] = st.w[12 * n + (s >> 4)];
    ew[3 * (size_t)e + 1] = st.w[12 * n + 4 + ((s >> 2) & 3)];
    ew[3 * (size_t)e + 2] = st.w[12 * n + 8 + (s & 3)];
    ent[e] = n;
  }
}

// Unforced moments of the box cells, lbm.cpp:1262-1285, summed in l order
// in FP64.  PULL: the buffer holds f* of the previous step and the cell's
// f is pulled first (the fused path); otherwise the buffer holds f.  Box
// cells that are not fluid keep their set-up rho, u (their density is read
// by the nearest-point force and the interior average).  Each workgroup
// also writes the sum of rho over its interior cells (fixed tree order).
template <class S, bool PULL>
__global__ __launch_bounds__(256) void k_ibm_box_moments(const Planes<S> P, const uint8_t *__restrict__ mask,
                                                         KArgs<typename Store<S>::C> a, Box B,
                                                         BoxArrays o, const uint8_t *__restrict__ inside,
                                                         double *__restrict__ part) {
  using C = typename Store<S>::C;
  __shared__ double sh[256];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double r_in = 0.0;
  if (b < B.ncell()) {
    const int x = B.x0 + b % B.nx, y = B.y0 + (b / B.nx) % B.ny, z = B.z0 + b / (B.nx * B.ny);
    const uint32_t idx = lattice_idx(a.g, x, y, z);
    const uint8_t mk = mask[idx];
    o.Fx[b] = 0.0;
    o.Fy[b] = 0.0;
    o.Fz[b] = 0.0;
    double rho, ux, uy, uz;
    if (!(mk & M_FLUID)) {
      rho = (double)a.rho[idx];
      ux = (double)a.ux[idx];
      uy = (double)a.uy[idx];
      uz = (double)a.uz[idx];
    } else {
      C f[Q];
      if constexpr (PULL) {
        pull_cell<S>(P, mk, idx, a, f);
      } else {
        for_dirs([&](auto M) {
          constexpr int m = decltype(M)::value;
          f[m] = Store<S>::ld(P.in[m] + idx);
        });
      }
      double r0 = 0, rx = 0, ry = 0, rz = 0;
      for_dirs([&](auto M) {
        constexpr int m = decltype(M)::value;
        double v = (double)f[m];
        if constexpr (Store<S>::shifted) v = v / (double)Store<S>::K + W[m];
        r0 += v;
        rx += v * CX[m];
        ry += v * CY[m];
        rz += v * CZ[m];
      });
      rho = r0;
      ux = rx / r0;
      uy = ry / r0;
      uz = rz / r0;
    }
    o.rho[b] = rho;
    o.ux[b] = ux;
    o.uy[b] = uy;
    o.uz[b] = uz;
    if (inside[b]) r_in = rho;
  }
  sh[threadIdx.x] = r_in;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

struct ForceArgs {
  const int32_t *near;   // box index of each marker's density point
  const double *ds;
  double *Fb, *Fds;      // F_b and F_b ds (3 per marker)
  double *dsum;          // running drag, lift per marker (2 per marker)
  double u0, u1, u2;     // u_des = V_CM
  int force;             // 0: interpolation only (last pass); 1: also the next force
  int first;             // 1: start the running drag/lift
  const double *ud;      // per-marker u_des (3 per marker) in place of V_CM, or null
  double *fsum;          // running -F_b per marker (3 per marker), or null
};

// Interpolation of one marker per wavefront (lbm.cpp:1295-1306 and
// 1370-1384) and the marker force of the next pass (lbm.cpp:1314-1323).
// Lane s = (a*4 + b)*4 + c is stencil point (base + (a, b, c)); points
// outside the box or on non-IBM cells add nothing.  Lanes 0..2 sum the
// 64 terms of one component serially in lane order, which is the
// reference's i, j, k loop order.
__global__ __launch_bounds__(256) void k_ibm_interp(int nm, const int32_t *__restrict__ pidx, Stencils st,
                                                    BoxArrays o, double *__restrict__ ub, ForceArgs fa) {
  __shared__ double term[4][3][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = blockIdx.x * 4 + wv;
  const bool valid = n < nm;
  double t0 = 0, t1 = 0, t2 = 0;
  const int b = valid ? pidx[64 * n + lane] : -1;
  // density at the marker's nearest point, read early (lanes 0..2)
  double rho_n = 0;
  if (fa.force && valid && lane < 3) {
    const int nb = fa.near[n];
    rho_n = nb >= 0 ? o.rho[nb] : nan("");
  }
  if (b >= 0) {
    const double kx = st.w[12 * n + (lane >> 4)], ky = st.w[12 * n + 4 + ((lane >> 2) & 3)],
                 kz = st.w[12 * n + 8 + (lane & 3)];
    t0 = o.ux[b] * kx * ky * kz;
    t1 = o.uy[b] * kx * ky * kz;
    t2 = o.uz[b] * kx * ky * kz;
  }
  term[wv][0][lane] = t0;
  term[wv][1][lane] = t1;
  term[wv][2][lane] = t2;
  __syncthreads();
  if (valid && lane < 3) {
    double acc = 0.0;
    for (int q = 0; q < 64; q++) acc += term[wv][lane][q];
    ub[3 * n + lane] = acc;
    if (fa.force) {
      const double dt = 1.0, dx = 1.0, dy = 1.0;
      const double rho = rho_n;
      const double ud = fa.ud ? fa.ud[3 * n + lane] : lane == 0 ? fa.u0 : lane == 1 ? fa.u1 : fa.u2;
      const double F = 2 * rho * (ud - acc) / dt;
      fa.Fb[3 * n + lane] = F;
      fa.Fds[3 * n + lane] = F * fa.ds[n];
      if (fa.fsum) fa.fsum[3 * n + lane] = fa.first ? -F : fa.fsum[3 * n + lane] - F;
      if (lane < 2) {
        const double d = -F * (lane == 0 ? dx : dy) * fa.ds[n];
        fa.dsum[2 * n + lane] = fa.first ? d : fa.dsum[2 * n + lane] + d;
      }
    }
  }
}

// ---- the implicit passes in marker space (rigid bodies) --------------------
// The m_max passes of lbm.cpp:1311-1387 are linear in the markers' forces:
// the cells' rho does not change during an IBM step, so one pass moves the
// interpolated velocity of marker n by
//   sum_x phi_n(x) Fm(x)/(2 rho(x)) = sum_k H_nk (u_des,k - u_b,k),
//   H_nk = rho_k ds_k sum_x phi_n(x) phi_k(x)/rho(x),
// with x over the live box cells that both stencils reach, and rho_k the
// density point of marker k (F_b,k = 2 rho_k (u_des,k - u_b,k)/dt).  The
// passes therefore run on the markers alone (H has a few dozen entries per
// row: only markers closer than 4 cells on every axis share stencil cells),
// one launch per pass, and the cells receive the sum of all passes' forces
// A_k = sum_m F_b,k^m ds_k in one spreading at the end: F = -sum_k phi_k A_k,
// u += sum_k phi_k A_k/(2 rho), which is the reference's field after its
// m_max spreadings.  Drag and lift still sum -F_b ds over every pass (Q12).

// Rows of H are stored ELL-style: R entries per marker (R a multiple of 8),
// the row's markers first (sorted), then padding entries (the marker
// itself with H = 0).

// The first interpolation (u_b of the unforced field, summed as k_ibm_interp
// sums), each marker's rho_n ds_n (rho_n its density point's value), and
// row n of G (H_nk = G_nk rho_k ds_k): one wavefront per marker.  The row's
// markers and their stencils are loaded with the cell values; phi_n/rho of
// the 64 stencil points goes to LDS; then each lane takes row entries and
// sums over all 64 points with k's weights shifted to n's stencil (zero off
// the overlap), so every lane reads the same LDS word at a time.
__global__ __launch_bounds__(64 * MK_WG) void k_ibm_interp_h(int nm, int R, const int32_t *__restrict__ pidx, Stencils st,
                                                      BoxArrays o, const int32_t *__restrict__ near,
                                                      const double *__restrict__ ds,
                                                      const int32_t *__restrict__ nlen,
                                                      const int32_t *__restrict__ nbr, double *__restrict__ G,
                                                      double *__restrict__ rn, double *__restrict__ cds,
                                                      double *__restrict__ ub) {
  __shared__ double term[MK_WG][3][64];
  __shared__ double qn[MK_WG][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = blockIdx.x * (blockDim.x >> 6) + wv;  // up to MK_WG wavefronts per workgroup
  const bool valid = n < nm;
  const int b = valid ? pidx[64 * n + lane] : -1;
  const int len = valid ? nlen[n] : 0;
  // up to two row entries per lane (rows of at most 128 entries take this
  // path; longer ones loop below)
  const int k0 = lane < len ? nbr[(size_t)n * R + lane] : -1;
  const int k1 = lane + 64 < len ? nbr[(size_t)n * R + lane + 64] : -1;
  double t0 = 0, t1 = 0, t2 = 0, q = 0;
  if (b >= 0) {
    const double kx = st.w[12 * n + (lane >> 4)], ky = st.w[12 * n + 4 + ((lane >> 2) & 3)],
                 kz = st.w[12 * n + 8 + (lane & 3)];
    t0 = o.ux[b] * kx * ky * kz;
    t1 = o.uy[b] * kx * ky * kz;
    t2 = o.uz[b] * kx * ky * kz;
    q = kx * ky * kz / o.rho[b];
  }
  term[wv][0][lane] = t0;
  term[wv][1][lane] = t1;
  term[wv][2][lane] = t2;
  qn[wv][lane] = q;
  __syncthreads();
  if (!valid) return;
  if (lane < 3) {
    double acc = 0.0;
    for (int s = 0; s < 64; s++) acc += term[wv][lane][s];
    ub[3 * n + lane] = acc;
  }
  if (lane == 3) {
    const int nn = near[n];
    const double r = nn >= 0 ? o.rho[nn] : nan("");
    rn[n] = r;
    cds[n] = r * ds[n];
  }
  const int bx = st.base[3 * n + 0], by = st.base[3 * n + 1], bz = st.base[3 * n + 2];
  for (int r = lane; r < len; r += 64) {
    const int k = r == lane ? k0 : r == lane + 64 ? k1 : nbr[(size_t)n * R + r];
    const int dx = st.base[3 * k + 0] - bx, dy = st.base[3 * k + 1] - by, dz = st.base[3 * k + 2] - bz;
    // k's weights on n's stencil points (0 where k's stencil does not reach)
    double px[4], py[4], pz[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int ax = a - dx, ay = a - dy, az = a - dz;
      px[a] = (ax >= 0 && ax < 4) ? st.w[12 * k + ax] : 0.0;
      py[a] = (ay >= 0 && ay < 4) ? st.w[12 * k + 4 + ay] : 0.0;
      pz[a] = (az >= 0 && az < 4) ? st.w[12 * k + 8 + az] : 0.0;
    }
    // four partial sums (one per z offset) keep the dependent chains short
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double pac = px[a] * py[c];
#pragma unroll
        for (int z = 0; z < 4; z++) acc[z] += qn[wv][(a * 4 + c) * 4 + z] * (pac * pz[z]);
      }
    G[(size_t)n * R + r] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }
}

// Pass m of the implicit scheme on the markers: F_b = 2 rho_n (u_des - u_b)
// from the previous pass's u_b (lbm.cpp:1316-1320), its drag and lift
// (lbm.cpp:1322-1323) and its share of A; then u_b after the pass's
// spreading and re-interpolation, u_b += sum_k H_nk (u_des - u_b,k).  A
// group of L lanes per marker (64/L markers per wavefront): each lane takes
// every L-th entry of the row, all of its row loads issued before the
// gathers of the entries' u_b and rho_k ds_k, and a fixed butterfly over
// the group sums them.
// PM: per-marker desired velocities udv (a coupled filament's nodes): the
// running fsum = -sum F_b takes the place of the drag/lift pairs.
template <bool PM, int L>
__global__ __launch_bounds__(64 * MK_WG) void k_ibm_pass(int nm, int R, const double *__restrict__ rn,
                                                  const double *__restrict__ cds,
                                                  const double *__restrict__ ds, const int32_t *__restrict__ nbr,
                                                  const double *__restrict__ G, const double *__restrict__ ub_in,
                                                  double *__restrict__ ub_out, double u0, double u1, double u2,
                                                  const double *__restrict__ udv, double *__restrict__ A,
                                                  double *__restrict__ dsum, double *__restrict__ Fb, int first,
                                                  int last) {
  static_assert(L >= 4 && L <= 64 && (L & (L - 1)) == 0, "lanes per marker: a power of two, 4..64");
  const int lane = threadIdx.x & 63, j = lane & (L - 1);
  const int n0 = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 / L);
  if (n0 >= nm) return;  // wave-uniform
  const int nr = n0 + lane / L;
  const bool valid = nr < nm;  // the group's lanes keep shuffling when not
  const int n = valid ? nr : nm - 1;
  // the marker's own values, loaded with the row
  const int l3 = j < 3 ? j : 0;
  const double ubn = ub_in[3 * n + l3], dsn = ds[n], rho = rn[n];
  const double udn = PM ? udv[3 * n + l3] : (j == 0 ? u0 : j == 1 ? u1 : u2);
  double Ao = 0, Do = 0;
  if (!first && j < 3) {
    Ao = A[3 * n + j];
    if (PM) Do = dsum[3 * n + j];
    else if (j < 2) Do = dsum[2 * n + j];
  }
  double s0 = 0, s1 = 0, s2 = 0;
  const size_t row = (size_t)n * R;
  constexpr int U = L == 64 ? 2 : 8;  // row entries per lane per round
  for (int r0 = 0; r0 < R; r0 += U * L) {
    int k[U];
    double g[U];
#pragma unroll
    for (int q = 0; q < U; q++) {
      const int r = r0 + q * L + j;
      k[q] = r < R ? nbr[row + r] : n;
      g[q] = r < R ? G[row + r] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < U; q++) {
      const int kk = k[q];
      const double h = g[q] * cds[kk];
      const double v0 = PM ? udv[3 * kk + 0] : u0, v1 = PM ? udv[3 * kk + 1] : u1, v2 = PM ? udv[3 * kk + 2] : u2;
      s0 += h * (v0 - ub_in[3 * kk + 0]);
      s1 += h * (v1 - ub_in[3 * kk + 1]);
      s2 += h * (v2 - ub_in[3 * kk + 2]);
    }
  }
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off);
    s1 += __shfl_xor(s1, off);
    s2 += __shfl_xor(s2, off);
  }
  if (!valid || j >= 3) return;
  const double sum = j == 0 ? s0 : j == 1 ? s1 : s2;
  const double dt = 1.0, dx = 1.0, dy = 1.0;
  const double F = 2 * rho * (udn - ubn) / dt;
  const double a = F * dsn;
  A[3 * n + j] = first ? a : Ao + a;
  if (PM) {
    dsum[3 * n + j] = first ? -F : Do - F;  // fsum
  } else if (j < 2) {
    const double dd = -F * (j == 0 ? dx : dy) * dsn;
    dsum[2 * n + j] = first ? dd : Do + dd;
  }
  if (last) Fb[3 * n + j] = F;
  ub_out[3 * n + j] = ubn + sum;
}

// The passes of one IBM step, one wavefront per marker (groups of fewer
// lanes, several markers per wavefront, measured slower: 6.9 against 5.1 us
// per pass on 4096 markers with 72-entry rows), `waves` wavefronts per
// workgroup, the u_b buffers alternating so that the last pass writes ub[0].
template <bool PM>
static void launch_passes(hipStream_t st, int nm, int R, int m_max, const double *rn, const double *cds,
                          const double *ds, const int32_t *nbr, const double *G, double *const ub[2],
                          const double u[3], const double *udv, double *A, double *dsum, double *Fb,
                          int waves = MK_WG) {
  const int s0 = m_max % 2;
  const int L = 64;
  const int per_wg = waves * (64 / L);
  const dim3 grid((unsigned)((nm + per_wg - 1) / per_wg)), blk(64 * waves);
  for (int it = 1; it <= m_max; it++) {
    const double *in = ub[(it - 1 + s0) % 2];
    double *out = ub[(it + s0) % 2];
    const int f = it == 1 ? 1 : 0, l = it == m_max ? 1 : 0;
#define LBM_PASS(LL)                                                                                                \
  hipLaunchKernelGGL((k_ibm_pass<PM, LL>), grid, blk, 0, st, nm, R, rn, cds, ds, nbr, G, in, out, u[0], u[1], u[2], \
                     udv, A, dsum, Fb, f, l)
    LBM_PASS(64);
#undef LBM_PASS
  }
}

// ---- paired passes: the implicit iteration with the squared operator -------
// With e = u_des - u_b and C = I - H, pass m maps e to C e, so after an
// optional single pass (m_max odd) the remaining 2q passes are q steps of
// Q = C^2 = I - 2H + H^2 (exact in exact arithmetic; the sums round
// differently, within a few ulps):
//   t_0 = e, t_{i+1} = Q t_i, W = t_0 + ... + t_{q-1},
//   sum of the 2q passes' e = W + C W,  e before the last pass = C t_{q-1},
//   e after it = Q t_{q-1}.
// The forces follow as in the single passes (F_b = 2 rho_n e per pass,
// A = sum F_b ds, drag/lift -F_b ds, the last F_b kept, u_b = u_des - e),
// with the q - 1 products by Q and one closing launch instead of 2q
// dependent launches.  Rows of Q cover the markers that share a stencil
// cell with a marker that shares one with n (sorted, R2 entries).

// Q = I - 2H + H^2 per step (H_nk = G_nk rho_k ds_k changes with the
// cells' rho): one workgroup per row, SQ_WAVES wavefronts taking every
// SQ_WAVES-th entry j of row n and adding (G_nj cds_j) G_jk over row j into
// their own window of marker indices [base_n, base_n + SQ_WIN) in LDS, in
// entry order (LDS atomic adds: a wavefront's adds to one word land in
// program order, and no two lanes of one add share a word); wavefront 0
// first adds -2 G_nk.  Then one thread per entry of Q's row: Q_nk = [k == n]
// + cds_k (sum of the windows at k).  Entries with G = 0 (the rows'
// padding among them) add nothing.  Measured on 4096 markers with 72-entry
// rows: 32.5 us, bound by the 21 M scattered adds (staging each workgroup's
// union of rows in LDS and giving a wavefront a whole row took 57 us).
constexpr int SQ_WIN = 1024;
constexpr int SQ_WAVES = 4;
constexpr int SQ_UNROLL = 9;  // rows j loaded together per wavefront
template <bool WIDE>  // rows of H longer than 64 entries (R <= 128)
__global__ __launch_bounds__(64 * SQ_WAVES) void k_ibm_square(int R, const int32_t *__restrict__ nbr,
                                                              const double *__restrict__ G,
                                                              const double *__restrict__ cds, int R2,
                                                              const int32_t *__restrict__ nbr2,
                                                              const int32_t *__restrict__ len2,
                                                              const int32_t *__restrict__ base2,
                                                              double *__restrict__ Q) {
  __shared__ double win[SQ_WAVES][SQ_WIN];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = blockIdx.x;
  const int base = base2[n];
  // row n: lane l holds entries l and l + 64 (R <= 128); the thread's entry
  // of Q's row and its cds, loaded with them
  const size_t rn = (size_t)n * R;
  const int ja = lane < R ? nbr[rn + lane] : n, jb = WIDE && lane + 64 < R ? nbr[rn + lane + 64] : n;
  const double ga = lane < R ? G[rn + lane] : 0.0, gb = WIDE && lane + 64 < R ? G[rn + lane + 64] : 0.0;
  const int p = threadIdx.x;
  const size_t e2 = (size_t)n * R2 + p;
  const bool real = p < R2 && p < len2[n];
  const int k2 = real ? nbr2[e2] : n;
  const double ca = ga * cds[ja], cb = gb * cds[jb], c2 = cds[k2];
  // each wavefront writes only its own window until the barrier below
  for (int i = lane; i < SQ_WIN; i += 64) win[wv][i] = 0.0;
  if (wv == 0) {
    if (ga != 0.0) atomicAdd(&win[0][ja - base], -2.0 * ga);
    if (WIDE && gb != 0.0) atomicAdd(&win[0][jb - base], -2.0 * gb);
  }
  for (int t0 = wv; t0 < R; t0 += SQ_WAVES * SQ_UNROLL) {
    int ka[SQ_UNROLL], kb[SQ_UNROLL];
    double gka[SQ_UNROLL], gkb[SQ_UNROLL], c[SQ_UNROLL];
#pragma unroll
    for (int u = 0; u < SQ_UNROLL; u++) {
      const int t = t0 + SQ_WAVES * u;
      const int tl = t & 63;
      const int j = !WIDE || t < 64 ? __shfl(ja, tl) : __shfl(jb, tl);
      const double cj = !WIDE || t < 64 ? __shfl(ca, tl) : __shfl(cb, tl);
      c[u] = t < R ? cj : 0.0;
      const size_t rj = (size_t)j * R;
      ka[u] = lane < R ? nbr[rj + lane] : j;
      kb[u] = WIDE && lane + 64 < R ? nbr[rj + lane + 64] : j;
      gka[u] = lane < R ? G[rj + lane] : 0.0;
      gkb[u] = WIDE && lane + 64 < R ? G[rj + lane + 64] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < SQ_UNROLL; u++) {
      if (c[u] == 0.0) continue;  // uniform over the wavefront
      if (gka[u] != 0.0) atomicAdd(&win[wv][ka[u] - base], c[u] * gka[u]);
      if (WIDE && gkb[u] != 0.0) atomicAdd(&win[wv][kb[u] - base], c[u] * gkb[u]);
    }
  }
  __syncthreads();
  if (p < R2) {
    double q = 0.0;
    if (real) {
      const int w = k2 - base;
      const double v = (win[0][w] + win[1][w]) + (win[2][w] + win[3][w]);
      q = (k2 == n ? 1.0 : 0.0) + c2 * v;
    }
    Q[e2] = q;
  }
}

// t_out = Q t_in and W += t_in, one wavefront per marker.  UB: t_in is
// u_des - u_b of the u_b array `tin` (the first product).
template <bool UB>
__global__ __launch_bounds__(64 * MK_WG) void k_ibm_qpass(int nm, int R2, const int32_t *__restrict__ nbr2,
                                                   const double *__restrict__ Q, const double *__restrict__ tin,
                                                   double *__restrict__ tout, double *__restrict__ W, double u0,
                                                   double u1, double u2, int first) {
  const int lane = threadIdx.x & 63;
  const int n = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (n >= nm) return;  // wave-uniform
  const int l3 = lane < 3 ? lane : 0;
  const double ul = lane == 0 ? u0 : lane == 1 ? u1 : u2;
  const double tn = UB ? ul - tin[3 * n + l3] : tin[3 * n + l3];
  const double wo = first ? 0.0 : W[3 * n + l3];
  double s0 = 0, s1 = 0, s2 = 0;
  const size_t row = (size_t)n * R2;
  for (int r0 = 0; r0 < R2; r0 += 4 * 64) {
    int k[4];
    double q[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = r0 + 64 * u + lane;
      k[u] = r < R2 ? nbr2[row + r] : n;
      q[u] = r < R2 ? Q[row + r] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int kk = k[u];
      const double a0 = tin[3 * kk + 0], a1 = tin[3 * kk + 1], a2 = tin[3 * kk + 2];
      s0 += q[u] * (UB ? u0 - a0 : a0);
      s1 += q[u] * (UB ? u1 - a1 : a1);
      s2 += q[u] * (UB ? u2 - a2 : a2);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off);
    s1 += __shfl_xor(s1, off);
    s2 += __shfl_xor(s2, off);
  }
  if (lane >= 3) return;
  tout[3 * n + lane] = lane == 0 ? s0 : lane == 1 ? s1 : s2;
  W[3 * n + lane] = first ? tn : wo + tn;
}

// The closing launch of the paired passes, from t = t_{q-1} (UB: from
// u_des - u_b of the u_b array `t`) and W = t_0 + ... + t_{q-2} (HW = 0:
// none): s = 2 (W + t) - H (W + t), e_last = t - H t; then A += 2 rho_n s
// ds_n, the drag/lift pairs -2 rho_n s ds_n, F_b = 2 rho_n e_last.
// `first`: A and the pairs start here.  The markers' final u_b = u_des -
// Q t is an output only (the next step interpolates afresh): k_ibm_ub_last
// computes it when it is read.
template <bool UB, bool WIDE>  // WIDE: rows of H longer than 64 entries (R <= 128)
__global__ __launch_bounds__(64 * MK_WG) void k_ibm_qfinal(int nm, int R, const int32_t *__restrict__ nbr,
                                                    const double *__restrict__ G, const double *__restrict__ cds,
                                                    const double *__restrict__ t, const double *__restrict__ W,
                                                    int HW, const double *__restrict__ rn,
                                                    const double *__restrict__ ds, double u0, double u1, double u2,
                                                    double *__restrict__ A, double *__restrict__ dsum,
                                                    double *__restrict__ Fb, int first) {
  const int lane = threadIdx.x & 63;
  const int n = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (n >= nm) return;  // wave-uniform
  const int l3 = lane < 3 ? lane : 0;
  const double ul = lane == 0 ? u0 : lane == 1 ? u1 : u2;
  const double tn = UB ? ul - t[3 * n + l3] : t[3 * n + l3];
  const double Wn = (HW ? W[3 * n + l3] : 0.0) + tn;
  const double dsn = ds[n], rho = rn[n];
  double Ao = 0, Do = 0;
  if (!first && lane < 3) {
    Ao = A[3 * n + lane];
    if (lane < 2) Do = dsum[2 * n + lane];
  }
  // H (W + t) and H t over row n of H: the row's entries first, then the
  // gathers
  constexpr int NH = WIDE ? 2 : 1;
  int kh[NH];
  double gh[NH];
#pragma unroll
  for (int u = 0; u < NH; u++) {
    const int r = 64 * u + lane;
    kh[u] = r < R ? nbr[(size_t)n * R + r] : n;
    gh[u] = r < R ? G[(size_t)n * R + r] : 0.0;
  }
  double h0 = 0, h1 = 0, h2 = 0, g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
  for (int u = 0; u < NH; u++) {
    const int kk = kh[u];
    const double h = gh[u] * cds[kk];
    const double a0 = t[3 * kk + 0], a1 = t[3 * kk + 1], a2 = t[3 * kk + 2];
    const double t0 = UB ? u0 - a0 : a0, t1 = UB ? u1 - a1 : a1, t2 = UB ? u2 - a2 : a2;
    const double w0 = (HW ? W[3 * kk + 0] : 0.0) + t0, w1 = (HW ? W[3 * kk + 1] : 0.0) + t1,
                 w2 = (HW ? W[3 * kk + 2] : 0.0) + t2;
    h0 += h * w0;
    h1 += h * w1;
    h2 += h * w2;
    g0 += h * t0;
    g1 += h * t1;
    g2 += h * t2;
  }
  for (int off = 32; off > 0; off >>= 1) {
    h0 += __shfl_xor(h0, off);
    h1 += __shfl_xor(h1, off);
    h2 += __shfl_xor(h2, off);
    g0 += __shfl_xor(g0, off);
    g1 += __shfl_xor(g1, off);
    g2 += __shfl_xor(g2, off);
  }
  if (lane >= 3) return;
  const double hw = lane == 0 ? h0 : lane == 1 ? h1 : h2;
  const double ht = lane == 0 ? g0 : lane == 1 ? g1 : g2;
  const double dt = 1.0, dx = 1.0, dy = 1.0;
  const double sm = 2.0 * Wn - hw;  // the 2q passes' summed e
  const double F = 2 * rho * sm / dt;
  const double a = F * dsn;
  A[3 * n + lane] = first ? a : Ao + a;
  if (lane < 2) {
    const double dd = -F * (lane == 0 ? dx : dy) * dsn;
    dsum[2 * n + lane] = first ? dd : Do + dd;
  }
  Fb[3 * n + lane] = 2 * rho * (tn - ht) / dt;
}

// The markers' u_b after the last of the paired passes: u_des - Q t_{q-1}
// (UB: t from u_des - u_b of the u_b array `t`), one wavefront per marker.
template <bool UB>
__global__ __launch_bounds__(64 * MK_WG) void k_ibm_ub_last(int nm, int R2, const int32_t *__restrict__ nbr2,
                                                     const double *__restrict__ Q, const double *__restrict__ t,
                                                     double u0, double u1, double u2, double *__restrict__ ub) {
  const int lane = threadIdx.x & 63;
  const int n = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (n >= nm) return;  // wave-uniform
  double q0 = 0, q1 = 0, q2 = 0;
  for (int r = lane; r < R2; r += 64) {
    const int kk = nbr2[(size_t)n * R2 + r];
    const double qv = Q[(size_t)n * R2 + r];
    const double a0 = t[3 * kk + 0], a1 = t[3 * kk + 1], a2 = t[3 * kk + 2];
    q0 += qv * (UB ? u0 - a0 : a0);
    q1 += qv * (UB ? u1 - a1 : a1);
    q2 += qv * (UB ? u2 - a2 : a2);
  }
  for (int off = 32; off > 0; off >>= 1) {
    q0 += __shfl_xor(q0, off);
    q1 += __shfl_xor(q1, off);
    q2 += __shfl_xor(q2, off);
  }
  if (lane >= 3) return;
  const double ul = lane == 0 ? u0 : lane == 1 ? u1 : u2;
  ub[3 * n + lane] = ul - (lane == 0 ? q0 : lane == 1 ? q1 : q2);
}

template <int T = FIN_THREADS>
__device__ __forceinline__ double fin_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = T / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// The drag/lift/added-mass sums of one IBM step, by one workgroup of T
// threads: fixed-order tree sums of the per-marker (or per-block) pairs and
// of the interior-density partials.
struct FinArgs {
  int block;  // k_ibm_spread: the workgroup that runs the sums (-1: none)
  int nm;     // pairs (drag, lift) in dsum
  const double *dsum;
  int npart;  // interior-density partials
  const double *part;
  double ninside, vsolid, d0, d1;
  double *out;
};

template <int T>
__device__ __forceinline__ void finish_sums(const FinArgs &f, double *sh) {
  double a = 0, l = 0, r = 0;
  for (int q = threadIdx.x; q < f.nm; q += T) {
    a += f.dsum[2 * q + 0];
    l += f.dsum[2 * q + 1];
  }
  for (int q = threadIdx.x; q < f.npart; q += T) r += f.part[q];
  const double drag = fin_sum<T>(a, sh), lift = fin_sum<T>(l, sh), rs = fin_sum<T>(r, sh);
  if (threadIdx.x == 0) {
    const double dt = 1.0;
    const double rho = rs / f.ninside;
    f.out[0] = drag + rho * f.vsolid * f.d0 / dt;
    f.out[1] = lift + rho * f.vsolid * f.d1 / dt;
    f.out[2] = rho;
  }
}

// Spreading and velocity correction, lbm.cpp:1332-1360.  One workgroup per
// chunk of band cells (cells that some marker stencil reaches): the terms
// F_b ds phi phi phi of the chunk's (marker, point) pairs are computed in
// parallel into LDS, then each cell sums its own terms serially in marker
// order, as the reference's loop over n does, and updates F and u.
__global__ __launch_bounds__(256) void k_ibm_spread(const int32_t *__restrict__ chunk,
                                                    const int32_t *__restrict__ band,
                                                    const int32_t *__restrict__ bstart,
                                                    const int32_t *__restrict__ en,
                                                    const double *__restrict__ ew,
                                                    const double *__restrict__ Fds, BoxArrays o,
                                                    double fsign, FinArgs fin) {
  __shared__ double term[3][CHUNK_E];
  if ((int)blockIdx.x == fin.block) {
    // the step's drag/lift sums, which read only the passes' per-marker
    // sums and the box-moment partials, in the same launch
    finish_sums<256>(fin, &term[0][0]);
    return;
  }
  // chunk[2q] = first band cell, chunk[2q+1] = first pair of chunk q
  const int c0 = chunk[2 * blockIdx.x], c1 = chunk[2 * blockIdx.x + 2];
  const int e0 = chunk[2 * blockIdx.x + 1], e1 = chunk[2 * blockIdx.x + 3];
  // the cell's own state, loaded while the pair terms are computed
  const int i = c0 + (int)threadIdx.x;
  const bool cell = i < c1;
  int s0 = 0, s1 = 0, b = 0;
  double r = 1, fx = 0, fy = 0, fz = 0, ux = 0, uy = 0, uz = 0;
  if (cell) {
    s0 = bstart[i] - e0;
    s1 = bstart[i + 1] - e0;
    b = band[i];
    r = o.rho[b];
    fx = o.Fx[b]; fy = o.Fy[b]; fz = o.Fz[b];
    ux = o.ux[b]; uy = o.uy[b]; uz = o.uz[b];
  }
  for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
    const int n = en[e];
    const double kx = ew[3 * (size_t)e + 0], ky = ew[3 * (size_t)e + 1], kz = ew[3 * (size_t)e + 2];
    term[0][e - e0] = Fds[3 * n + 0] * kx * ky * kz;
    term[1][e - e0] = Fds[3 * n + 1] * kx * ky * kz;
    term[2][e - e0] = Fds[3 * n + 2] * kx * ky * kz;
  }
  __syncthreads();
  if (cell) {
    double m0 = 0, m1 = 0, m2 = 0;
    for (int e = s0; e < s1; e++) {
      m0 += term[0][e];
      m1 += term[1][e];
      m2 += term[2][e];
    }
    // the implicit branch subtracts the spread force, the explicit one adds
    // it (lbm.cpp:1347-1349 against 1244-1246); u gains +Fm/(2 rho) in both
    const double dt = 1.0;
    o.Fx[b] = fx + fsign * m0;
    o.Fy[b] = fy + fsign * m1;
    o.Fz[b] = fz + fsign * m2;
    o.ux[b] = ux + dt * m0 / (2 * r);
    o.uy[b] = uy + dt * m1 / (2 * r);
    o.uz[b] = uz + dt * m2 / (2 * r);
  }
}

// The explicit branch's drag and lift, summed over the box's F/IN/OUT cells
// (lbm.cpp:1252-1253: -F_x dx^2, -F_y dx^2): per block a fixed-order tree
// sum, stored as (drag, lift) pairs that k_ibm_finish reduces like the
// per-marker sums of the implicit branch.
__global__ __launch_bounds__(256) void k_ibm_cell_force(Box B, BoxArrays o, double *__restrict__ pairs) {
  __shared__ double sh[256];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double fx = 0, fy = 0;
  if (b < B.ncell() && o.live[b]) {
    const double dx = 1.0;
    fx = -o.Fx[b] * (dx * dx);
    fy = -o.Fy[b] * (dx * dx);
  }
  for (int c = 0; c < 2; c++) {
    sh[threadIdx.x] = c == 0 ? fx : fy;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) pairs[2 * blockIdx.x + c] = sh[0];
    __syncthreads();
  }
}

// Drag, lift and the added-mass term, lbm.cpp:1105-1132, 1390-1398.  One
// workgroup, fixed summation order: the per-marker running sums and the
// interior-density partials of k_ibm_box_moments.  ninside = the number of
// interior box cells (0: the reference divides by zero, NaN here too).
__global__ __launch_bounds__(FIN_THREADS) void k_ibm_finish(FinArgs f) {
  __shared__ double sh[FIN_THREADS];
  finish_sums<FIN_THREADS>(f, sh);
}

// Density point of a thin body's marker: the box cell holding the marker
// (nearest lattice point), -1 outside the box.
__global__ void k_ibm_near_cell(Box B, int nm, const double *__restrict__ mx, const double *__restrict__ my,
                                const double *__restrict__ mz, int32_t *__restrict__ near) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  const int i = (int)floor(mx[n] + 0.5), j = (int)floor(my[n] + 0.5), k = (int)floor(mz[n] + 0.5);
  const bool in = i >= B.x0 && i <= B.x1 && j >= B.y0 && j <= B.y1 && k >= B.z0 && k <= B.z1;
  near[n] = in ? (i - B.x0) + B.nx * ((j - B.y0) + B.ny * (k - B.z0)) : -1;
}

// Markers of a coupled filament: its node positions and, as desired
// velocities, the displacement over the last step (dt = 1).
__global__ void k_fsi_markers(int nm, const double *__restrict__ x, const double *__restrict__ xp,
                              double *__restrict__ mx, double *__restrict__ my, double *__restrict__ mz,
                              double *__restrict__ ud) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  mx[n] = x[3 * n + 0];
  my[n] = x[3 * n + 1];
  mz[n] = x[3 * n + 2];
  for (int d = 0; d < 3; d++) ud[3 * n + d] = x[3 * n + d] - xp[3 * n + d];
}

// ---- the IBM passes of a coupled filament -------------------------------------
// A filament's markers move every step, so its kernels take the marker
// positions straight from the filament's nodes and derive the stencil base,
// the weights and the density cell themselves (the same expressions as
// k_ibm_weights and k_ibm_near_cell, hence the same values); no binning and
// no host round trip.
struct FsiArgs {
  Box B;
  int nm;
  const double *x, *xp;  // the nodes now and one step before (3 per node)
  const double *ds;
  const uint8_t *live;
  double *ub, *Fb, *Fds, *fsum;  // 3 per marker
  // written by the first pass, for the marker getters
  double *mx, *my, *mz, *ud;
  Stencils st;
  int32_t *near;
};

__device__ __forceinline__ int box_cell(const Box &B, int i, int j, int k) {
  const bool in = i >= B.x0 && i <= B.x1 && j >= B.y0 && j <= B.y1 && k >= B.z0 && k <= B.z1;
  return in ? (i - B.x0) + B.nx * ((j - B.y0) + B.ny * (k - B.z0)) : -1;
}

// Interpolation and the next pass's marker force for a filament: as
// k_ibm_interp (one wavefront per marker, lane = stencil point, lanes 0..2
// sum in lane order), with u_des = the node's displacement over the last
// step, rho at the box cell holding the marker, and the running
// fsum = -sum F_b.  `first` starts fsum and writes the markers' positions,
// u_des, stencils and density cells.
__global__ __launch_bounds__(256) void k_fsi_interp(FsiArgs a, BoxArrays o, int force, int first) {
  __shared__ double term[4][3][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = blockIdx.x * 4 + wv;
  const bool valid = n < a.nm;
  double p0 = 0, p1 = 0, p2 = 0, xpl = 0, dsn = 0, fs = 0;
  if (valid) {
    p0 = a.x[3 * n + 0];
    p1 = a.x[3 * n + 1];
    p2 = a.x[3 * n + 2];
    if (lane < 3) {
      xpl = a.xp[3 * n + lane];
      dsn = a.ds[n];
      if (!first) fs = a.fsum[3 * n + lane];
    }
  }
  const int b0 = (int)floor(p0) - 1, b1 = (int)floor(p1) - 1, b2 = (int)floor(p2) - 1;
  const int pa = lane >> 4, pb = (lane >> 2) & 3, pc = lane & 3;
  int b = valid ? box_cell(a.B, b0 + pa, b1 + pb, b2 + pc) : -1;
  const int nb = box_cell(a.B, (int)floor(p0 + 0.5), (int)floor(p1 + 0.5), (int)floor(p2 + 0.5));
  double rho_n = 0;
  if (force && valid && lane < 3) rho_n = nb >= 0 ? o.rho[nb] : nan("");
  // the cell's velocity is loaded with its live flag (box arrays cover
  // every box cell), the term kept for live cells only
  double t0 = 0, t1 = 0, t2 = 0;
  if (b >= 0) {
    const bool lv = o.live[b] != 0;
    const double vx = o.ux[b], vy = o.uy[b], vz = o.uz[b];
    const double kx = peskin(((b0 + pa) * 1.0 - p0) / 1.0), ky = peskin(((b1 + pb) * 1.0 - p1) / 1.0),
                 kz = peskin(((b2 + pc) * 1.0 - p2) / 1.0);
    if (lv) {
      t0 = vx * kx * ky * kz;
      t1 = vy * kx * ky * kz;
      t2 = vz * kx * ky * kz;
    }
  }
  term[wv][0][lane] = t0;
  term[wv][1][lane] = t1;
  term[wv][2][lane] = t2;
  __syncthreads();
  if (!valid) return;
  const double pl = lane == 0 ? p0 : lane == 1 ? p1 : p2;
  if (lane < 3) {
    double acc = 0.0;
    for (int q = 0; q < 64; q++) acc += term[wv][lane][q];
    a.ub[3 * n + lane] = acc;
    if (force) {
      const double dt = 1.0;
      const double ud = pl - xpl;
      const double F = 2 * rho_n * (ud - acc) / dt;
      a.Fb[3 * n + lane] = F;
      a.Fds[3 * n + lane] = F * dsn;
      a.fsum[3 * n + lane] = first ? -F : fs - F;
    }
  }
  if (first) {
    if (lane < 3) {
      (lane == 0 ? a.mx : lane == 1 ? a.my : a.mz)[n] = pl;
      a.ud[3 * n + lane] = pl - xpl;
      a.st.base[3 * n + lane] = lane == 0 ? b0 : lane == 1 ? b1 : b2;
    }
    if (lane < 12) {
      const int ax = lane >> 2, q = lane & 3;
      const int bb = ax == 0 ? b0 : ax == 1 ? b1 : b2;
      const double p = ax == 0 ? p0 : ax == 1 ? p1 : p2;
      a.st.w[12 * n + lane] = peskin(((bb + q) * 1.0 - p) / 1.0);
    }
    if (lane == 0) a.near[n] = nb;
  }
}

// The filament's IBM passes in marker space (as the rigid body's, with the
// nodes' own desired velocities): the first interpolation with the markers'
// geometry (k_fsi_interp with first = 1, force = 0: the same stencils,
// weights, density cells and u_b), rho_n and rho_n ds_n, and row n of G over
// every marker k (entry r = k, R >= nm):
//   G_nk = sum_x phi_n(x) phi_k(x)/rho(x)
// over the live box cells both stencils reach, 0 where they share none.  The
// nodes move every step, so each wave takes marker k's stencil base and
// weights from its position, with k_fsi_interp's expressions.
__global__ __launch_bounds__(64 * MK_WG) void k_fsi_interp_h(FsiArgs a, BoxArrays o, int R, double *__restrict__ G,
                                                          double *__restrict__ rn, double *__restrict__ cds) {
  __shared__ double term[MK_WG][3][64];
  __shared__ double qn[MK_WG][64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n = blockIdx.x * (blockDim.x >> 6) + wv;  // up to MK_WG wavefronts per workgroup
  const bool valid = n < a.nm;
  double p0 = 0, p1 = 0, p2 = 0, xpl = 0;
  if (valid) {
    p0 = a.x[3 * n + 0];
    p1 = a.x[3 * n + 1];
    p2 = a.x[3 * n + 2];
    if (lane < 3) xpl = a.xp[3 * n + lane];
  }
  const int b0 = (int)floor(p0) - 1, b1 = (int)floor(p1) - 1, b2 = (int)floor(p2) - 1;
  const int pa = lane >> 4, pb = (lane >> 2) & 3, pc = lane & 3;
  const int b = valid ? box_cell(a.B, b0 + pa, b1 + pb, b2 + pc) : -1;
  const int nb = box_cell(a.B, (int)floor(p0 + 0.5), (int)floor(p1 + 0.5), (int)floor(p2 + 0.5));
  double t0 = 0, t1 = 0, t2 = 0, q = 0;
  if (b >= 0) {
    const bool lv = o.live[b] != 0;
    const double vx = o.ux[b], vy = o.uy[b], vz = o.uz[b], rb = o.rho[b];
    const double kx = peskin(((b0 + pa) * 1.0 - p0) / 1.0), ky = peskin(((b1 + pb) * 1.0 - p1) / 1.0),
                 kz = peskin(((b2 + pc) * 1.0 - p2) / 1.0);
    if (lv) {
      t0 = vx * kx * ky * kz;
      t1 = vy * kx * ky * kz;
      t2 = vz * kx * ky * kz;
      q = kx * ky * kz / rb;
    }
  }
  term[wv][0][lane] = t0;
  term[wv][1][lane] = t1;
  term[wv][2][lane] = t2;
  qn[wv][lane] = q;
  __syncthreads();
  if (!valid) return;
  const double pl = lane == 0 ? p0 : lane == 1 ? p1 : p2;
  if (lane < 3) {
    double acc = 0.0;
    for (int s = 0; s < 64; s++) acc += term[wv][lane][s];
    a.ub[3 * n + lane] = acc;
    (lane == 0 ? a.mx : lane == 1 ? a.my : a.mz)[n] = pl;
    a.ud[3 * n + lane] = pl - xpl;
    a.st.base[3 * n + lane] = lane == 0 ? b0 : lane == 1 ? b1 : b2;
  }
  if (lane < 12) {
    const int ax = lane >> 2, qq = lane & 3;
    const int bb = ax == 0 ? b0 : ax == 1 ? b1 : b2;
    const double p = ax == 0 ? p0 : ax == 1 ? p1 : p2;
    a.st.w[12 * n + lane] = peskin(((bb + qq) * 1.0 - p) / 1.0);
  }
  if (lane == 0) a.near[n] = nb;
  if (lane == 3) {
    const double r = nb >= 0 ? o.rho[nb] : nan("");
    rn[n] = r;
    cds[n] = r * a.ds[n];
  }
  for (int r = lane; r < a.nm; r += 64) {
    const double x0 = a.x[3 * r + 0], x1 = a.x[3 * r + 1], x2 = a.x[3 * r + 2];
    const int c0 = (int)floor(x0) - 1, c1 = (int)floor(x1) - 1, c2 = (int)floor(x2) - 1;
    const int dx = c0 - b0, dy = c1 - b1, dz = c2 - b2;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (dx > -4 && dx < 4 && dy > -4 && dy < 4 && dz > -4 && dz < 4) {
      // k's weights on n's stencil points (0 where k's stencil does not reach)
      double px[4], py[4], pz[4];
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const int ax = s - dx, ay = s - dy, az = s - dz;
        px[s] = (ax >= 0 && ax < 4) ? peskin(((c0 + ax) * 1.0 - x0) / 1.0) : 0.0;
        py[s] = (ay >= 0 && ay < 4) ? peskin(((c1 + ay) * 1.0 - x1) / 1.0) : 0.0;
        pz[s] = (az >= 0 && az < 4) ? peskin(((c2 + az) * 1.0 - x2) / 1.0) : 0.0;
      }
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const double pac = px[s] * py[c];
#pragma unroll
          for (int z = 0; z < 4; z++) acc[z] += qn[wv][(s * 4 + c) * 4 + z] * (pac * pz[z]);
        }
    }
    G[(size_t)n * R + r] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }
}

// Spreading and velocity correction for a filament (at most FSI_MAX
// markers): one thread per box cell gathers F_b ds phi phi phi over the
// markers whose stencil holds the cell, in marker order (the reference's
// loop over n, lbm.cpp:1332-1360; the terms and their order equal
// k_ibm_spread's).  A wavefront takes a 4x4x4 tile of box cells.  The
// markers' stencils (from the step's first k_fsi_interp) and forces are
// staged in LDS, their loads issued together with the cell's own state, and
// markers go in batches of FSI_BATCH with the batch's stencil hull: a
// wavefront whose tile the hull misses skips the batch on one test, the
// others add their hit terms in marker order.
constexpr int FSI_MAX = 256;
constexpr int FSI_BATCH = 8;
__global__ __launch_bounds__(256) void k_fsi_spread(FsiArgs a, BoxArrays o) {
  __shared__ int4 sb[FSI_MAX];
  __shared__ int4 hull[FSI_MAX / FSI_BATCH][2];  // lowest base, highest base + 3
  __shared__ double sw[FSI_MAX * 12];
  __shared__ double sf[FSI_MAX * 3];
  __shared__ int32_t sbf[FSI_MAX * 3];
  const int nm = a.nm;
  const int npad = (nm + FSI_BATCH - 1) / FSI_BATCH * FSI_BATCH;
  // the thread's cell: tile t of the box (x fastest), lane (x, y, z) in it
  const Box &B = a.B;
  const int tx = (B.nx + 3) >> 2, ty = (B.ny + 3) >> 2;
  const int t = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int ox = 4 * (t % tx), oy = 4 * ((t / tx) % ty), oz = 4 * (t / (tx * ty));
  const int cx = ox + (lane & 3), cy = oy + ((lane >> 2) & 3), cz = oz + (lane >> 4);
  const bool inb = cx < B.nx && cy < B.ny && cz < B.nz;
  const int b = inb ? cx + B.nx * (cy + B.ny * cz) : 0;
  // every global load first: the cell's state, the markers' positions and forces
  bool active = false;
  double r = 1, fx = 0, fy = 0, fz = 0, ux = 0, uy = 0, uz = 0;
  if (inb) {
    active = o.live[b] != 0;
    r = o.rho[b];
    fx = o.Fx[b];
    fy = o.Fy[b];
    fz = o.Fz[b];
    ux = o.ux[b];
    uy = o.uy[b];
    uz = o.uz[b];
  }
  // the stencils of this step's markers (written by its first k_fsi_interp)
#pragma unroll
  for (int u = 0; u < 3 * FSI_MAX / 256; u++) {
    const int q = (int)threadIdx.x + 256 * u;
    const bool real = q < 3 * nm;
    sbf[q] = real ? a.st.base[q] : -(1 << 28);  // no cell is within 4 of it
    sf[q] = real ? a.Fds[q] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 12 * FSI_MAX / 256; u++) {
    const int q = (int)threadIdx.x + 256 * u;
    sw[q] = q < 12 * nm ? a.st.w[q] : 0.0;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < npad; q += blockDim.x) sb[q] = make_int4(sbf[3 * q], sbf[3 * q + 1], sbf[3 * q + 2], 0);
  __syncthreads();
  for (int q = threadIdx.x; q < npad / FSI_BATCH; q += blockDim.x) {
    int4 lo = make_int4(1 << 29, 1 << 29, 1 << 29, 0), hi = make_int4(-(1 << 29), -(1 << 29), -(1 << 29), 0);
    for (int u = 0; u < FSI_BATCH; u++) {
      const int n = q * FSI_BATCH + u;
      if (n >= nm) break;
      const int4 v = sb[n];
      lo.x = min(lo.x, v.x);
      lo.y = min(lo.y, v.y);
      lo.z = min(lo.z, v.z);
      hi.x = max(hi.x, v.x + 3);
      hi.y = max(hi.y, v.y + 3);
      hi.z = max(hi.z, v.z + 3);
    }
    hull[q][0] = lo;
    hull[q][1] = hi;
  }
  __syncthreads();
  // the tile in reference indices
  const int i0 = B.x0 + ox, j0 = B.y0 + oy, k0 = B.z0 + oz;
  const int i = B.x0 + cx, j = B.y0 + cy, k = B.z0 + cz;
  // the batches whose hull meets the tile (uniform over the wavefront)
  uint32_t reach = 0;
#pragma unroll
  for (int q = 0; q < FSI_MAX / FSI_BATCH; q++) {
    const int4 lo = hull[q][0], hi = hull[q][1];
    const bool meets = q * FSI_BATCH < nm && lo.x <= i0 + 3 && hi.x >= i0 && lo.y <= j0 + 3 && hi.y >= j0 &&
                       lo.z <= k0 + 3 && hi.z >= k0;
    reach |= (uint32_t)meets << q;
  }
  reach = __builtin_amdgcn_readfirstlane(reach);
  double m0 = 0, m1 = 0, m2 = 0;
  bool any = false;
  while (reach) {
    const int n0 = FSI_BATCH * __builtin_ctz(reach);
    reach &= reach - 1;
    uint32_t di[FSI_BATCH], dj[FSI_BATCH], dk[FSI_BATCH];
    uint32_t hit = 0;
#pragma unroll
    for (int u = 0; u < FSI_BATCH; u++) {
      const int4 q = sb[n0 + u];
      di[u] = (uint32_t)(i - q.x);
      dj[u] = (uint32_t)(j - q.y);
      dk[u] = (uint32_t)(k - q.z);
      hit |= (uint32_t)((di[u] | dj[u] | dk[u]) < 4u) << u;
    }
    hit = active ? hit : 0u;
    if (!__ballot(hit != 0)) continue;
    any = any || hit != 0;
#pragma unroll
    for (int u = 0; u < FSI_BATCH; u++) {
      const int n = n0 + u;
      const double kx = sw[12 * n + (di[u] & 3)], ky = sw[12 * n + 4 + (dj[u] & 3)],
                   kz = sw[12 * n + 8 + (dk[u] & 3)];
      const double t0 = sf[3 * n + 0] * kx * ky * kz;
      const double t1 = sf[3 * n + 1] * kx * ky * kz;
      const double t2 = sf[3 * n + 2] * kx * ky * kz;
      const bool h = (hit >> u) & 1u;
      m0 = h ? m0 + t0 : m0;
      m1 = h ? m1 + t1 : m1;
      m2 = h ? m2 + t2 : m2;
    }
  }
  if (!any) return;
  const double dt = 1.0;
  o.Fx[b] = fx + -m0;
  o.Fy[b] = fy + -m1;
  o.Fz[b] = fz + -m2;
  o.ux[b] = ux + dt * m0 / (2 * r);
  o.uy[b] = uy + dt * m1 / (2 * r);
  o.uz[b] = uz + dt * m2 / (2 * r);
}

// 4x4x4 tiles of a box, 4 per workgroup
inline dim3 fsi_spread_grid(const Box &B) {
  const int tiles = ((B.nx + 3) >> 2) * ((B.ny + 3) >> 2) * ((B.nz + 3) >> 2);
  return dim3((unsigned)((tiles + 3) / 4));
}

__global__ void k_fill_i32(int32_t *p, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline dim3 blocks(int n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

}  // namespace

// ---------------------------------------------------------------------------
struct lbm_ibm {
  int nm = 0, m_max = 0, moving = 0, follow = 0;
  double vsolid = 0;
  double cx = 0, cy = 0, half = 0;  // body centre and box half-width (follow)
  Box B{};
  bool placed = false;
  double *mx = nullptr, *my = nullptr, *mz = nullptr, *nx = nullptr, *ny = nullptr, *nz = nullptr;
  double *ds = nullptr, *ub = nullptr, *Fb = nullptr, *Fds = nullptr, *dsum = nullptr;
  int32_t *near = nullptr;
  Stencils st{};
  double ninside = 0;  // interior box cells
  // box arrays (capacity box_cap cells)
  double *brho = nullptr, *bux = nullptr, *buy = nullptr, *buz = nullptr;
  double *bFx = nullptr, *bFy = nullptr, *bFz = nullptr;
  uint8_t *live = nullptr, *inside = nullptr;
  double *part = nullptr;  // interior-density partials, one per box-moments workgroup
  int32_t *cnt = nullptr, *cursor = nullptr, *bandpos = nullptr;  // per box cell
  size_t box_cap = 0;
  // (marker, point) binning: pidx (64 per marker), band cells and their
  // pair segments, spreading chunks
  int32_t *pidx = nullptr, *band = nullptr, *bstart = nullptr, *ent = nullptr, *chunk = nullptr;
  double *ew = nullptr;
  size_t band_cap = 0, bstart_cap = 0, ent_cap = 0, ew_cap = 0, chunk_cap = 0;
  int nband = 0, nchunk = 0;
  std::vector<int32_t> h_cnt, h_band, h_bandpos, h_bstart, h_chunk;
  std::vector<uint8_t> h_inside;
  double *out = nullptr;  // 3 per step of a run
  int out_cap = 0;
  int ready = 0;          // IBLBM done, the next collision uses force mode 3
  int displaced = 0;      // some step moved the markers (their positions are state)
  std::vector<void *> allocs;  // marker arrays, freed with the state
  // coupled filament (lbm_fsi_setup): its nodes are the markers, with their
  // own desired velocities and the density of the cell they sit in
  lbm_filament *fil = nullptr;
  int near_cell = 0;
  double *ud = nullptr, *fsum = nullptr;
  int scheme = 0;              // 0 implicit, 1 explicit (one pass, lbm.cpp:1171-1258)
  // marker-space passes (rigid implicit): markers closer than 4 cells on
  // every axis (CSR rows, the marker itself included), H per entry, the
  // second u_b buffer of the passes
  int32_t *nlen = nullptr, *nbr = nullptr;  // row lengths; R entries per row
  double *H = nullptr, *ub2 = nullptr;  // H: G_nk (times rho_k ds_k in the passes)
  double *rn = nullptr, *cds = nullptr;  // density point value rho_n per marker, rho_n ds_n
  int R = 0;
  int rows_static = 0;  // rows by stencil bases (valid while the markers stay put)
  // paired passes (m_max >= 2): rows of Q = (I - H)^2, R2 entries each
  // (len2 real, the rest padding), the LDS window base of each row, Q per
  // step, the products' t buffers, W and the u_b of a leading single pass
  int R2 = 0;
  int32_t *nbr2 = nullptr, *len2 = nullptr, *base2 = nullptr;
  double *Q2 = nullptr, *sq_t[2] = {nullptr, nullptr}, *sq_w = nullptr, *sq_ub = nullptr;
  // the last paired passes' u_b, computed when read: from t (UB: the u_b
  // array t) and u_des
  int ub_pending = 0, ub_from_ub = 0;
  const double *ub_t = nullptr;
  double ub_u[3] = {0, 0, 0};
  // a coupled filament's passes in marker space: R >= nm entries per row,
  // entry r = marker r (the nodes move, so G is recomputed every step)
  int fsi_rows = 0;
  double *cell_pairs = nullptr;  // explicit: per-block (drag, lift) of the box cells
  size_t pairs_cap = 0;
};

static void *dalloc(lbm_ibm *m, size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  m->allocs.push_back(p);
  return p;
}

template <class T>
static void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

static void free_box_arrays(lbm_ibm *m) {
  dfree(m->brho); dfree(m->bux); dfree(m->buy); dfree(m->buz);
  dfree(m->bFx); dfree(m->bFy); dfree(m->bFz);
  dfree(m->live); dfree(m->inside); dfree(m->part);
  dfree(m->cnt); dfree(m->cursor); dfree(m->bandpos);
  m->box_cap = 0;
}

static void free_bin_arrays(lbm_ibm *m) {
  dfree(m->band); dfree(m->bstart); dfree(m->ent); dfree(m->chunk); dfree(m->ew);
  m->band_cap = m->bstart_cap = m->ent_cap = m->ew_cap = m->chunk_cap = 0;
}

// grows a device array to hold n elements of T (contents not kept)
template <class T>
static bool grow(T *&p, size_t &cap, size_t n) {
  if (n <= cap && p) return true;
  dfree(p);
  cap = 0;
  if (hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return false;
  cap = std::max<size_t>(n, 1);
  return true;
}

void ibm_free(lbm_ctx *c) {
  if (!c || !c->ibm) return;
  lbm_ibm *m = c->ibm;
  for (void *p : m->allocs) (void)hipFree(p);
  free_box_arrays(m);
  free_bin_arrays(m);
  dfree(m->out);
  dfree(m->cell_pairs);
  delete m;
  c->ibm = nullptr;
}

static BoxArrays box_arrays(const lbm_ibm *m) {
  return BoxArrays{m->brho, m->bux, m->buy, m->buz, m->bFx, m->bFy, m->bFz, m->live};
}

template <class S>
static Planes<S> planes_of(const lbm_ctx *c, const void *in) {
  Planes<S> P;
  const size_t ps = c->g.ps;
  const S *b = (const S *)in;
  for (int q = 0; q < Q; q++) {
    const ptrdiff_t off = (ptrdiff_t)CX[q] + (ptrdiff_t)CY[q] * c->g.px + (ptrdiff_t)CZ[q] * c->g.plane;
    P.in[q] = b + q * ps;
    P.src[q] = b + (ptrdiff_t)(q * ps) - off;
    P.out[q] = nullptr;
  }
  return P;
}

template <class S>
static KArgs<typename Store<S>::C> kargs_of(const LaunchCtx &L) {
  using C = typename Store<S>::C;
  KArgs<C> a{};
  a.g = L.g;
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
  a.wall_slots = L.wall_slots;
  a.slot_links = L.slot_links;
  return a;
}

template <class S>
static hipError_t box_moments_s(lbm_ctx *c, bool pull, hipStream_t st) {
  lbm_ibm *m = c->ibm;
  const LaunchCtx L = launch_ctx(c);
  const Planes<S> P = planes_of<S>(c, c->cur);
  const auto a = kargs_of<S>(L);
  if (pull)
    hipLaunchKernelGGL((k_ibm_box_moments<S, true>), blocks(m->B.ncell()), dim3(256), 0, st, P, c->mask, a, m->B,
                       box_arrays(m), m->inside, m->part);
  else
    hipLaunchKernelGGL((k_ibm_box_moments<S, false>), blocks(m->B.ncell()), dim3(256), 0, st, P, c->mask, a, m->B,
                       box_arrays(m), m->inside, m->part);
  return hipGetLastError();
}

static hipError_t box_moments(lbm_ctx *c, bool pull, hipStream_t st = nullptr) {
  if (!st) st = c->stream;
  switch (c->storage) {
    case LBM_STORE_F64: return box_moments_s<double>(c, pull, st);
    case LBM_STORE_F32: return box_moments_s<float>(c, pull, st);
    case LBM_STORE_F16S: return box_moments_s<__half>(c, pull, st);
    default: return box_moments_s<f16c_t>(c, pull, st);
  }
}

// The box of the reference's IBLBM around a centre: [floor(c - half),
// ceil(c + half)] in x and y (lbm.cpp:1165-1166).
static void follow_box(const lbm_ibm *m, int box[6]) {
  box[0] = (int)floor(m->cx - m->half);
  box[1] = (int)ceil(m->cx + m->half);
  box[2] = (int)floor(m->cy - m->half);
  box[3] = (int)ceil(m->cy + m->half);
  box[4] = m->B.z0;
  box[5] = m->B.z1;
}

// Checks a box against the lattice: inside it, and only TYPE_F or TYPE_S
// cells (inlet/outlet populations are not evolved on the device).
static int check_box(const lbm_ctx *c, const int b[6]) {
  if (b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] >= (int)c->g.nx || b[3] >= (int)c->g.ny ||
      b[5] >= (int)c->g.nz || b[0] > b[1] || b[2] > b[3] || b[4] > b[5])
    return LBM_ERR_ARG;
  for (int k = b[4]; k <= b[5]; k++)
    for (int j = b[2]; j <= b[3]; j++)
      for (int i = b[0]; i <= b[1]; i++) {
        const int16_t t = c->h_type[((size_t)k * c->g.ny + j) * c->g.px + i];
        if (t != LBM_TYPE_F && t != LBM_TYPE_S) return LBM_ERR_UNSUPPORTED;
      }
  return LBM_OK;
}

// Puts the IBM box at b: clears the old box's mask bits, grows the box
// arrays when needed, marks the new box and points the collision's
// box arguments at it.
static int place_box(lbm_ctx *c, const int b[6]) {
  lbm_ibm *m = c->ibm;
  int e = check_box(c, b);
  if (e != LBM_OK) return e;
  if (m->placed) {
    hipLaunchKernelGGL(k_ibm_mark, blocks(m->B.ncell()), dim3(256), 0, c->stream, c->mask, c->g, m->B, m->live, 0);
    HIPCHK(hipGetLastError());
    m->placed = false;
  }
  Box &B = m->B;
  B.x0 = b[0]; B.x1 = b[1]; B.y0 = b[2]; B.y1 = b[3]; B.z0 = b[4]; B.z1 = b[5];
  B.nx = b[1] - b[0] + 1; B.ny = b[3] - b[2] + 1; B.nz = b[5] - b[4] + 1;
  const size_t nb = (size_t)B.ncell();
  if (nb > m->box_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    free_box_arrays(m);
    double **bp[7] = {&m->brho, &m->bux, &m->buy, &m->buz, &m->bFx, &m->bFy, &m->bFz};
    for (int q = 0; q < 7; q++)
      if (hipMalloc((void **)bp[q], 8 * nb) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipMalloc((void **)&m->live, nb) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipMalloc((void **)&m->inside, nb) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipMalloc((void **)&m->part, 8 * ((nb + 255) / 256)) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipMalloc((void **)&m->cnt, 4 * nb) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipMalloc((void **)&m->cursor, 4 * nb) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipMalloc((void **)&m->bandpos, 4 * nb) != hipSuccess) return LBM_ERR_NOMEM;
    for (int q = 0; q < 7; q++) HIPCHK(hipMemsetAsync(*bp[q], 0, 8 * nb, c->stream));
    m->box_cap = nb;
  }
  hipLaunchKernelGGL(k_ibm_mark, blocks((int)nb), dim3(256), 0, c->stream, c->mask, c->g, B, m->live, 1);
  HIPCHK(hipGetLastError());
  m->placed = true;
  // the collision of IBM steps reads the box (force mode 3)
  c->box = BoxArgs{B.x0, B.y0, B.z0, B.nx, B.ny, B.nz, m->brho, m->bux, m->buy, m->buz, m->bFx, m->bFy, m->bFz};
  return LBM_OK;
}

// Everything that depends on the marker positions: stencil weights, nearest
// points, interior cells, and the binning of (marker, point) pairs to box
// cells with the spreading chunks.  Ends with the stream synchronised.
static int ibm_geometry(lbm_ctx *c) {
  lbm_ibm *m = c->ibm;
  const Box &B = m->B;
  const int nb = B.ncell(), np = m->nm * 64;
  hipLaunchKernelGGL(k_ibm_weights, blocks(m->nm), dim3(256), 0, c->stream, m->nm, m->mx, m->my, m->mz, m->st);
  if (m->near_cell)
    hipLaunchKernelGGL(k_ibm_near_cell, blocks(m->nm), dim3(256), 0, c->stream, B, m->nm, m->mx, m->my, m->mz, m->near);
  else if (m->scheme == 1)
    hipLaunchKernelGGL(k_ibm_near_local, blocks(m->nm), dim3(256), 0, c->stream, B, m->nm, m->mx, m->my, m->mz, m->nx,
                       m->ny, m->nz, m->near);
  else
    hipLaunchKernelGGL(k_ibm_nearest, dim3(m->nm), dim3(256), 0, c->stream, B, m->mx, m->my, m->mz, m->nx, m->ny, m->nz, m->near);
  hipLaunchKernelGGL(k_ibm_inside, blocks(nb), dim3(256), 0, c->stream, B, m->nm, m->mx, m->my, m->mz, m->nx, m->ny, m->nz, m->inside);
  hipLaunchKernelGGL(k_fill_i32, blocks(nb), dim3(256), 0, c->stream, m->cnt, nb, 0);
  hipLaunchKernelGGL(k_fill_i32, blocks(nb), dim3(256), 0, c->stream, m->cursor, nb, 0);
  hipLaunchKernelGGL(k_ibm_bin_count, blocks(np), dim3(256), 0, c->stream, B, m->nm, m->st, m->live, m->cnt, m->pidx);
  HIPCHK(hipGetLastError());
  m->h_cnt.resize(nb);
  m->h_inside.resize(nb);
  HIPCHK(hipMemcpyAsync(m->h_cnt.data(), m->cnt, 4 * (size_t)nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(m->h_inside.data(), m->inside, (size_t)nb, hipMemcpyDeviceToHost, c->stream));
  std::vector<int32_t> h_near;
  if (m->scheme == 1 && !m->near_cell) {
    h_near.resize(m->nm);
    HIPCHK(hipMemcpyAsync(h_near.data(), m->near, 4 * (size_t)m->nm, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  // the explicit branch reads rho at the marker's local nearest point, or at
  // (0, 0, 1) when no neighbour lies ahead along its normal (lbm.cpp:1216-
  // 1225, e.g. a marker on a lattice point): a cell outside the box, whose
  // IBM-step density the device does not compute
  for (int32_t v : h_near)
    if (v < 0) return LBM_ERR_UNSUPPORTED;
  size_t nin = 0;
  for (uint8_t v : m->h_inside) nin += v;
  m->ninside = (double)nin;
  // band cells in box order, their pair segments, and chunks of at most
  // CHUNK_C cells and CHUNK_E pairs
  m->h_band.clear();
  m->h_bandpos.assign(nb, -1);
  m->h_bstart.assign(1, 0);
  m->h_chunk.assign(1, 0);
  int in_chunk_c = 0, in_chunk_e = 0;
  for (int b = 0; b < nb; b++) {
    const int k = m->h_cnt[b];
    if (k == 0) continue;
    if (k > CHUNK_E) return LBM_ERR_UNSUPPORTED;  // more than CHUNK_E markers reach one cell
    const int i = (int)m->h_band.size();
    if (in_chunk_c + 1 > CHUNK_C || in_chunk_e + k > CHUNK_E) {
      m->h_chunk.push_back(i);
      in_chunk_c = 0;
      in_chunk_e = 0;
    }
    m->h_bandpos[b] = i;
    m->h_band.push_back(b);
    m->h_bstart.push_back(m->h_bstart.back() + k);
    in_chunk_c++;
    in_chunk_e += k;
  }
  m->nband = (int)m->h_band.size();
  m->h_chunk.push_back(m->nband);
  if (m->nband == 0) m->h_chunk.assign(1, 0);
  m->nchunk = (int)m->h_chunk.size() - 1;
  {  // (first band cell, first pair) per chunk boundary
    std::vector<int32_t> cp(2 * m->h_chunk.size());
    for (size_t q = 0; q < m->h_chunk.size(); q++) {
      cp[2 * q] = m->h_chunk[q];
      cp[2 * q + 1] = m->h_bstart[m->h_chunk[q]];
    }
    m->h_chunk.swap(cp);
  }
  const size_t nent = (size_t)m->h_bstart.back();
  if (!grow(m->band, m->band_cap, (size_t)m->nband + 1) ||
      !grow(m->bstart, m->bstart_cap, (size_t)m->nband + 1) || !grow(m->ent, m->ent_cap, nent) ||
      !grow(m->ew, m->ew_cap, 3 * nent))
    return LBM_ERR_NOMEM;
  if (!grow(m->chunk, m->chunk_cap, m->h_chunk.size())) return LBM_ERR_NOMEM;
  if (m->nband) HIPCHK(hipMemcpyAsync(m->band, m->h_band.data(), 4 * (size_t)m->nband, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(m->bstart, m->h_bstart.data(), 4 * m->h_bstart.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(m->bandpos, m->h_bandpos.data(), 4 * (size_t)nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(m->chunk, m->h_chunk.data(), 4 * m->h_chunk.size(), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_ibm_bin_fill, blocks(np), dim3(256), 0, c->stream, m->nm, m->pidx, m->bandpos, m->bstart, m->cursor, m->ent);
  hipLaunchKernelGGL(k_ibm_bin_sort, blocks(std::max(m->nband, 1)), dim3(256), 0, c->stream, m->nband, m->bstart, m->st, m->ent, m->ew);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

// One IBLBM call's device work after the box moments: interpolation, m_max
// force/spread/interpolate passes, and the drag/lift/added-mass sums into
// out[0..2].
static int ibm_iterate(lbm_ctx *c, const double vcm[3], const double vb[3], double *out, hipStream_t st = nullptr) {
  lbm_ibm *m = c->ibm;
  m->ub_pending = 0;  // every path below but the paired passes writes m->ub
  if (!st) st = c->stream;
  const Box &B = m->B;
  const BoxArrays o = box_arrays(m);
  const bool expl = m->scheme == 1;
  ForceArgs fa{m->near, m->ds, m->Fb, m->Fds, m->dsum, vcm[0], vcm[1], vcm[2], expl || m->m_max > 0, 1, m->ud, m->fsum};
  const dim3 gi((unsigned)((m->nm + 3) / 4));
  const int npart = (B.ncell() + 255) / 256;
  FinArgs none{};
  none.block = -1;
  const FinArgs fin{-1, m->nm, m->dsum, npart, m->part, m->ninside, m->vsolid, vcm[0] - vb[0], vcm[1] - vb[1], out};
  if (expl) {
    hipLaunchKernelGGL(k_ibm_interp, gi, dim3(256), 0, st, m->nm, m->pidx, m->st, o, m->ub, fa);
    // one pass: the unforced interpolation's force, spread with + (the
    // markers keep the unforced u_b), drag and lift summed over the cells
    if (m->nchunk > 0)
      hipLaunchKernelGGL(k_ibm_spread, dim3(m->nchunk), dim3(256), 0, st, m->chunk, m->band, m->bstart, m->ent,
                         m->ew, m->Fds, o, 1.0, none);
    if (out) {
      if ((size_t)npart > m->pairs_cap) {
        HIPCHK(hipStreamSynchronize(st));
        dfree(m->cell_pairs);
        HIPCHK(hipMalloc((void **)&m->cell_pairs, 16 * (size_t)npart));
        m->pairs_cap = npart;
      }
      hipLaunchKernelGGL(k_ibm_cell_force, dim3((unsigned)npart), dim3(256), 0, st, B, o, m->cell_pairs);
      FinArgs cf = fin;
      cf.nm = npart;  // per-block (drag, lift) pairs of the box cells
      cf.dsum = m->cell_pairs;
      hipLaunchKernelGGL(k_ibm_finish, dim3(1), dim3(FIN_THREADS), 0, st, cf);
    }
    HIPCHK(hipGetLastError());
    return LBM_OK;
  }
  if (m->nbr && !m->ud) {
    // the passes in marker space, then one spreading of their summed force
    // (u_b buffers alternate so that the last pass writes m->ub)
    double *buf[2] = {m->ub, m->ub2};
    const int s0 = m->m_max % 2;
    const dim3 gm((unsigned)((m->nm + MK_WG - 1) / MK_WG));
    if (m->nbr2 && m->m_max >= 2) {
      // paired passes: the first interpolation into ub2, Q's rows, a single
      // pass when m_max is odd, q - 1 products by Q, the closing launch
      hipLaunchKernelGGL(k_ibm_interp_h, gm, dim3(64 * MK_WG), 0, st, m->nm, m->R, m->pidx, m->st, o, m->near,
                         m->ds, m->nlen, m->nbr, m->H, m->rn, m->cds, m->ub2);
      if (m->R > 64)
        hipLaunchKernelGGL(k_ibm_square<true>, dim3((unsigned)m->nm), dim3(64 * SQ_WAVES), 0, st, m->R, m->nbr, m->H,
                           m->cds, m->R2, m->nbr2, m->len2, m->base2, m->Q2);
      else
        hipLaunchKernelGGL(k_ibm_square<false>, dim3((unsigned)m->nm), dim3(64 * SQ_WAVES), 0, st, m->R, m->nbr, m->H,
                           m->cds, m->R2, m->nbr2, m->len2, m->base2, m->Q2);
      const int lead = m->m_max % 2, q = (m->m_max - lead) / 2;
      double *src = m->ub2;
      if (lead) {
        hipLaunchKernelGGL((k_ibm_pass<false, 64>), gm, dim3(64 * MK_WG), 0, st, m->nm, m->R, m->rn, m->cds, m->ds,
                           m->nbr, m->H, m->ub2, m->sq_ub, vcm[0], vcm[1], vcm[2], nullptr, m->Fds, m->dsum, m->Fb,
                           1, 0);
        src = m->sq_ub;
      }
      for (int i = 0; i + 1 < q; i++) {
        if (i == 0)
          hipLaunchKernelGGL(k_ibm_qpass<true>, gm, dim3(64 * MK_WG), 0, st, m->nm, m->R2, m->nbr2, m->Q2, src,
                             m->sq_t[0], m->sq_w, vcm[0], vcm[1], vcm[2], 1);
        else
          hipLaunchKernelGGL(k_ibm_qpass<false>, gm, dim3(64 * MK_WG), 0, st, m->nm, m->R2, m->nbr2, m->Q2,
                             m->sq_t[(i - 1) % 2], m->sq_t[i % 2], m->sq_w, vcm[0], vcm[1], vcm[2], 0);
      }
      const bool wide = m->R > 64;
      const double *tq = q == 1 ? src : m->sq_t[(q - 2) % 2];
      const int hw = q == 1 ? 0 : 1, first = lead ? 0 : 1;
#define LBM_QFINAL(UBV, WV)                                                                                        \
  hipLaunchKernelGGL((k_ibm_qfinal<UBV, WV>), gm, dim3(64 * MK_WG), 0, st, m->nm, m->R, m->nbr, m->H, m->cds, tq, \
                     m->sq_w, hw, m->rn, m->ds, vcm[0], vcm[1], vcm[2], m->Fds, m->dsum, m->Fb, first)
      if (q == 1) {
        if (wide) LBM_QFINAL(true, true);
        else LBM_QFINAL(true, false);
      } else {
        if (wide) LBM_QFINAL(false, true);
        else LBM_QFINAL(false, false);
      }
#undef LBM_QFINAL
      // the final u_b when it is read (lbm_ibm_markers)
      m->ub_pending = 1;
      m->ub_t = tq;
      m->ub_from_ub = q == 1;
      for (int c3 = 0; c3 < 3; c3++) m->ub_u[c3] = vcm[c3];
    } else {
      hipLaunchKernelGGL(k_ibm_interp_h, gm, dim3(64 * MK_WG), 0, st, m->nm, m->R, m->pidx, m->st, o, m->near,
                         m->ds, m->nlen, m->nbr, m->H, m->rn, m->cds, buf[s0]);
      launch_passes<false>(st, m->nm, m->R, m->m_max, m->rn, m->cds, m->ds, m->nbr, m->H, buf, vcm, nullptr,
                           m->Fds, m->dsum, m->Fb);
    }
    if (m->m_max > 0 && m->nchunk > 0) {
      // the drag/lift sums as one more workgroup of the spreading launch
      FinArgs f = fin;
      f.block = out ? m->nchunk : -1;
      hipLaunchKernelGGL(k_ibm_spread, dim3(m->nchunk + (out ? 1 : 0)), dim3(256), 0, st, m->chunk, m->band,
                         m->bstart, m->ent, m->ew, m->Fds, o, -1.0, f);
      HIPCHK(hipGetLastError());
      return LBM_OK;
    }
  } else {
    hipLaunchKernelGGL(k_ibm_interp, gi, dim3(256), 0, st, m->nm, m->pidx, m->st, o, m->ub, fa);
    fa.first = 0;
    for (int it = 1; it <= m->m_max; it++) {
      if (m->nchunk > 0)
        hipLaunchKernelGGL(k_ibm_spread, dim3(m->nchunk), dim3(256), 0, st, m->chunk, m->band, m->bstart,
                           m->ent, m->ew, m->Fds, o, -1.0, none);
      fa.force = it < m->m_max;
      hipLaunchKernelGGL(k_ibm_interp, gi, dim3(256), 0, st, m->nm, m->pidx, m->st, o, m->ub, fa);
    }
  }

  if (out)
    hipLaunchKernelGGL(k_ibm_finish, dim3(1), dim3(FIN_THREADS), 0, st, fin);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// The IBM passes of a coupled filament's step after the box moments: the
// first interpolation (with the markers' geometry), then m_max spread /
// interpolate passes; the last interpolation computes no force.
static int fsi_iterate(lbm_ctx *c) {
  lbm_ibm *m = c->ibm;
  m->ub_pending = 0;
  double *fx, *fxp, fds;
  int n, fdev;
  filament_state(m->fil, &fx, &fxp, &n, &fds, &fdev);
  const FsiArgs a{m->B, n, fx, fxp, m->ds, m->live, m->ub, m->Fb, m->Fds, m->fsum,
                  m->mx, m->my, m->mz, m->ud, m->st, m->near};
  const BoxArrays o = box_arrays(m);
  if (m->fsi_rows && m->m_max > 0) {
    // the passes in marker space, then one spreading of their summed force
    // (u_b buffers alternate so that the last pass writes m->ub)
    double *buf[2] = {m->ub, m->ub2};
    const int s0 = m->m_max % 2;
    FsiArgs a0 = a;
    a0.ub = buf[s0];
    // one wavefront per workgroup: a filament's few markers spread over as
    // many compute units (k_fsi_interp_h computes every row entry's weights)
    const dim3 gm((unsigned)n);
    hipLaunchKernelGGL(k_fsi_interp_h, gm, dim3(64), 0, c->stream, a0, o, m->R, m->H, m->rn, m->cds);
    const double zero[3] = {0, 0, 0};
    launch_passes<true>(c->stream, n, m->R, m->m_max, m->rn, m->cds, m->ds, m->nbr, m->H, buf, zero, m->ud, m->Fds,
                        m->fsum, m->Fb, 1);
    hipLaunchKernelGGL(k_fsi_spread, fsi_spread_grid(m->B), dim3(256), 0, c->stream, a, o);
    HIPCHK(hipGetLastError());
    return LBM_OK;
  }
  const dim3 gi((unsigned)((n + 3) / 4));
  hipLaunchKernelGGL(k_fsi_interp, gi, dim3(256), 0, c->stream, a, o, m->m_max > 0 ? 1 : 0, 1);
  for (int it = 1; it <= m->m_max; it++) {
    hipLaunchKernelGGL(k_fsi_spread, fsi_spread_grid(m->B), dim3(256), 0, c->stream, a, o);
    hipLaunchKernelGGL(k_fsi_interp, gi, dim3(256), 0, c->stream, a, o, it < m->m_max ? 1 : 0, 0);
  }
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// Marker motion of one IBLBM call (lbm.cpp:1154-1166): the trapezoid rule
// for the markers and, when the box follows the body, for its centre; then
// the box and the geometry that depends on the marker positions.
static int build_neighbours(lbm_ibm *m, const double *x, const double *y, const double *z, bool by_base);

static int ibm_move(lbm_ctx *c, const double vcm[3], const double vb[3]) {
  lbm_ibm *m = c->ibm;
  const bool moved = vcm[0] + vb[0] != 0 || vcm[1] + vb[1] != 0 || vcm[2] + vb[2] != 0;
  if (moved) m->displaced = 1;
  if (moved && m->rows_static) {
    // a body set up at rest was given a velocity: its rows by stencil bases
    // no longer cover every overlapping pair; the rows of a rigid motion
    // (markers closer than 4 cells) replace them
    m->nbr = nullptr;
    m->nbr2 = nullptr;
    m->rows_static = 0;
    std::vector<double> x(m->nm), y(m->nm), z(m->nm);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(x.data(), m->mx, 8 * (size_t)m->nm, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(y.data(), m->my, 8 * (size_t)m->nm, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(z.data(), m->mz, 8 * (size_t)m->nm, hipMemcpyDeviceToHost));
    if (int e = build_neighbours(m, x.data(), y.data(), z.data(), false)) return e;
  }
  // (V_CM + V_CM_before) dt/2 = 0 leaves every marker where it is
  if (moved) {
    hipLaunchKernelGGL(k_ibm_move, blocks(m->nm), dim3(256), 0, c->stream, m->nm, m->mx, m->my, m->mz, vcm[0], vcm[1],
                       vcm[2], vb[0], vb[1], vb[2]);
    HIPCHK(hipGetLastError());
  }
  if (m->follow && moved) {
    const double dt = 1.0;
    m->cx += (vcm[0] + vb[0]) * dt / 2;
    m->cy += (vcm[1] + vb[1]) * dt / 2;
    int b[6];
    follow_box(m, b);
    const Box &B = m->B;
    if (b[0] != B.x0 || b[1] != B.x1 || b[2] != B.y0 || b[3] != B.y1) {
      int e = place_box(c, b);
      if (e != LBM_OK) return e;
    }
  }
  // the stencils, nearest points, interior cells and bins follow a body that moves
  if (m->moving || moved) return ibm_geometry(c);
  return LBM_OK;
}

// The rows of Q = (I - H)^2 for the paired passes: for each marker the
// union of its neighbours' rows, sorted.  Kept only when every row spans
// fewer than SQ_WIN marker indices (the square kernel's LDS window) and has
// at most 64 SQ_WAVES entries (one thread per entry); otherwise the single
// passes run alone.
static int build_square(lbm_ibm *m, const std::vector<std::vector<int32_t>> &rows) {
  const int nm = m->nm;
  std::vector<int32_t> stamp(nm, -1), base(nm), len(nm);
  std::vector<std::vector<int32_t>> r2(nm);
  size_t maxlen = 0;
  for (int n = 0; n < nm; n++) {
    std::vector<int32_t> &row = r2[n];
    for (int32_t j : rows[n])
      for (int32_t k : rows[j])
        if (stamp[k] != n) {
          stamp[k] = n;
          row.push_back(k);
        }
    std::sort(row.begin(), row.end());
    if (row.back() - row.front() >= SQ_WIN) return LBM_OK;
    base[n] = row.front();
    len[n] = (int32_t)row.size();
    maxlen = std::max(maxlen, row.size());
  }
  const int R2 = (int)((maxlen + 7) / 8 * 8);
  if (R2 > 64 * SQ_WAVES) return LBM_OK;
  std::vector<int32_t> idx((size_t)nm * R2);
  for (int n = 0; n < nm; n++)
    for (int r = 0; r < R2; r++) idx[(size_t)n * R2 + r] = r < len[n] ? r2[n][r] : n;
  m->nbr2 = (int32_t *)dalloc(m, 4 * idx.size());
  m->len2 = (int32_t *)dalloc(m, 4 * (size_t)nm);
  m->base2 = (int32_t *)dalloc(m, 4 * (size_t)nm);
  m->Q2 = (double *)dalloc(m, 8 * idx.size());
  m->sq_t[0] = (double *)dalloc(m, 24 * (size_t)nm);
  m->sq_t[1] = (double *)dalloc(m, 24 * (size_t)nm);
  m->sq_w = (double *)dalloc(m, 24 * (size_t)nm);
  m->sq_ub = (double *)dalloc(m, 24 * (size_t)nm);
  if (!m->nbr2 || !m->len2 || !m->base2 || !m->Q2 || !m->sq_t[0] || !m->sq_t[1] || !m->sq_w || !m->sq_ub) {
    m->nbr2 = nullptr;
    return LBM_ERR_NOMEM;
  }
  HIPCHK(hipMemcpy(m->nbr2, idx.data(), 4 * idx.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->len2, len.data(), 4 * (size_t)nm, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->base2, base.data(), 4 * (size_t)nm, hipMemcpyHostToDevice));
  m->R2 = R2;
  return LBM_OK;
}

// The rows of H for the marker-space passes: for each marker the markers
// closer than 4 cells on every axis (itself included), sorted.  Stencils
// (floor(x) - 1 .. floor(x) + 2 per axis) of markers 4 or more cells apart
// on some axis never share a cell, and a rigid motion keeps the list.
// by_base (a body that stays put): only markers whose stencil bases are at
// most 3 apart on every axis, the pairs whose stencils do share cells.
// Buckets of 4 cells; rows longer than 256 entries on average (markers far
// denser than the lattice) keep the pass-by-pass path.
static int build_neighbours(lbm_ibm *m, const double *x, const double *y, const double *z, bool by_base) {
  const int nm = m->nm;
  std::vector<std::pair<uint64_t, int>> key(nm);
  auto bucket = [](double v) { return (int64_t)std::floor(v / 4.0); };
  auto pack = [](int64_t a, int64_t b, int64_t c) {
    return ((uint64_t)(a & 0x1fffff) << 42) | ((uint64_t)(b & 0x1fffff) << 21) | (uint64_t)(c & 0x1fffff);
  };
  for (int n = 0; n < nm; n++) key[n] = {pack(bucket(x[n]), bucket(y[n]), bucket(z[n])), n};
  std::sort(key.begin(), key.end());
  std::vector<std::vector<int32_t>> rows(nm);
  size_t total = 0, maxlen = 0;
  for (int n = 0; n < nm; n++) {
    std::vector<int32_t> &row = rows[n];
    const int64_t bx = bucket(x[n]), by = bucket(y[n]), bz = bucket(z[n]);
    for (int64_t i = bx - 1; i <= bx + 1; i++)
      for (int64_t j = by - 1; j <= by + 1; j++)
        for (int64_t k = bz - 1; k <= bz + 1; k++) {
          const uint64_t kk = pack(i, j, k);
          auto it = std::lower_bound(key.begin(), key.end(), std::make_pair(kk, -1));
          for (; it != key.end() && it->first == kk; ++it) {
            const int q = it->second;
            if (std::fabs(x[q] - x[n]) < 4.0 && std::fabs(y[q] - y[n]) < 4.0 && std::fabs(z[q] - z[n]) < 4.0 &&
                (!by_base || (std::fabs(std::floor(x[q]) - std::floor(x[n])) <= 3.0 &&
                              std::fabs(std::floor(y[q]) - std::floor(y[n])) <= 3.0 &&
                              std::fabs(std::floor(z[q]) - std::floor(z[n])) <= 3.0)))
              row.push_back(q);
          }
        }
    std::sort(row.begin(), row.end());
    total += row.size();
    maxlen = std::max(maxlen, row.size());
    if (total > (size_t)256 * nm) return LBM_OK;  // too dense: keep the pass-by-pass path
  }
  const int R = (int)((maxlen + 7) / 8 * 8);
  if (m->m_max >= 2 && R <= 128) {
    int e = build_square(m, rows);
    if (e != LBM_OK) return e;
  }
  std::vector<int32_t> len(nm), idx((size_t)nm * R);
  for (int n = 0; n < nm; n++) {
    len[n] = (int32_t)rows[n].size();
    for (int r = 0; r < R; r++) idx[(size_t)n * R + r] = r < len[n] ? rows[n][r] : n;
  }
  m->R = R;
  m->nlen = (int32_t *)dalloc(m, 4 * (size_t)nm);
  m->nbr = (int32_t *)dalloc(m, 4 * idx.size());
  m->H = (double *)dalloc(m, 8 * idx.size());
  m->ub2 = (double *)dalloc(m, 24 * (size_t)nm);
  m->rn = (double *)dalloc(m, 8 * (size_t)nm);
  m->cds = (double *)dalloc(m, 8 * (size_t)nm);
  if (!m->nlen || !m->nbr || !m->H || !m->ub2 || !m->rn || !m->cds) return LBM_ERR_NOMEM;
  HIPCHK(hipMemcpy(m->nlen, len.data(), 4 * (size_t)nm, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->nbr, idx.data(), 4 * idx.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(m->H, 0, 8 * idx.size()));  // the padding entries stay 0
  return LBM_OK;
}

static int ibm_setup(lbm_ctx *c, const lbm_ibm_desc *d, const double *x, const double *y,
                     const double *z, const double *nx, const double *ny, const double *nz,
                     const double *ds) {
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->ibm) {
    if (c->ibm->placed) {
      hipLaunchKernelGGL(k_ibm_mark, blocks(c->ibm->B.ncell()), dim3(256), 0, c->stream, c->mask, c->g, c->ibm->B, c->ibm->live, 0);
      HIPCHK(hipStreamSynchronize(c->stream));
    }
    ibm_free(c);
    c->box = BoxArgs{};
  }
  lbm_ibm *m = new lbm_ibm();
  c->ibm = m;
  m->nm = d->nmarkers;
  m->m_max = d->m_max;
  m->moving = d->moving;
  m->follow = d->follow;
  m->vsolid = d->vsolid;
  m->cx = d->center[0];
  m->cy = d->center[1];
  m->half = d->half;
  m->scheme = d->scheme;
  int b[6] = {d->box[0], d->box[1], d->box[2], d->box[3], d->box[4], d->box[5]};
  if (m->follow) {
    m->B.z0 = b[4];
    m->B.z1 = b[5];
    follow_box(m, b);
  }
  int e = check_box(c, b);
  if (e != LBM_OK) return e;
  const size_t nm = (size_t)m->nm;
  double **mp[7] = {&m->mx, &m->my, &m->mz, &m->nx, &m->ny, &m->nz, &m->ds};
  const double *src[7] = {x, y, z, nx, ny, nz, ds};
  for (int q = 0; q < 7; q++) {
    *mp[q] = (double *)dalloc(m, 8 * nm);
    if (!*mp[q]) return LBM_ERR_NOMEM;
    HIPCHK(hipMemcpy(*mp[q], src[q], 8 * nm, hipMemcpyHostToDevice));
  }
  m->ub = (double *)dalloc(m, 24 * nm);
  m->Fb = (double *)dalloc(m, 24 * nm);
  m->Fds = (double *)dalloc(m, 24 * nm);
  m->dsum = (double *)dalloc(m, 16 * nm);
  m->near = (int32_t *)dalloc(m, 4 * nm);
  m->st.base = (int32_t *)dalloc(m, 12 * nm);
  m->st.w = (double *)dalloc(m, 96 * nm);
  m->pidx = (int32_t *)dalloc(m, 256 * nm);
  if (!m->pidx) return LBM_ERR_NOMEM;
  if (!m->ub || !m->Fb || !m->Fds || !m->dsum || !m->near || !m->st.base || !m->st.w) return LBM_ERR_NOMEM;
  HIPCHK(hipMemsetAsync(m->ub, 0, 24 * nm, c->stream));
  HIPCHK(hipMemsetAsync(m->Fb, 0, 24 * nm, c->stream));
  HIPCHK(hipMemsetAsync(m->Fds, 0, 24 * nm, c->stream));
  HIPCHK(hipMemsetAsync(m->dsum, 0, 16 * nm, c->stream));
  // rigid bodies of the implicit scheme run the passes in marker space
  if (m->scheme == 0) {
    // a body at rest keeps its stencil bases: only markers whose stencils
    // overlap (bases at most 3 apart on every axis) enter a row
    const bool by_base = !m->moving && !m->follow;
    e = build_neighbours(m, x, y, z, by_base);
    m->rows_static = by_base;
    if (e != LBM_OK) return e;
  }
  e = place_box(c, b);
  if (e != LBM_OK) return e;
  e = ibm_geometry(c);
  if (e != LBM_OK) return e;
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

extern "C" {

int lbm_ibm_setup(lbm_ctx *c, const lbm_ibm_desc *d, const double *x, const double *y,
                  const double *z, const double *nx, const double *ny, const double *nz,
                  const double *ds) {
  if (!c || !d || !x || !y || !z || !nx || !ny || !nz || !ds) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  // the reference's IBM cases have no ENERGY; the D3Q7 branch couples (its
  // buoyancy added to the IBM force)
  if (c->thermal == 1) return LBM_ERR_UNSUPPORTED;
  if (d->nmarkers < 1 || d->m_max < 0) return LBM_ERR_ARG;
  if (d->follow && !(d->half >= 0)) return LBM_ERR_ARG;
  if (d->scheme != 0 && d->scheme != 1) return LBM_ERR_ARG;
  // the explicit branch leaves F of cells outside the box as it was (the
  // set-up force, or the last IBM force of a cell the box has left); the
  // IBM collision takes F = 0 there, so only unforced set-ups with a box
  // that stays in place are accepted
  if (d->scheme == 1 && (d->follow || c->force_mode != 0)) return LBM_ERR_UNSUPPORTED;
  const int e = ibm_setup(c, d, x, y, z, nx, ny, nz, ds);
  if (e != LBM_OK && c->ibm) {
    // leave no half-built body behind: clear its marks, drop the state
    if (c->ibm->placed) {
      const Box &B = c->ibm->B;
      hipLaunchKernelGGL(k_ibm_mark, blocks(B.ncell()), dim3(256), 0, c->stream, c->mask, c->g, B, c->ibm->live, 0);
      (void)hipStreamSynchronize(c->stream);
    }
    ibm_free(c);
    c->box = BoxArgs{};
  }
  return e;
}

// IBLBM (lbm.cpp:1135-1417) on the lattice as it stands after Init or
// Streaming; the next lbm_collide uses its result.
int lbm_ibm_apply(lbm_ctx *c, const double vcm[3], const double vcm_before[3], double out[3]) {
  if (!c || !vcm || !vcm_before) return LBM_ERR_ARG;
  if (!c->ibm) return LBM_ERR_STATE;
  if (c->state != ST_F) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  lbm_ibm *m = c->ibm;
  if (m->out_cap < 1) {
    dfree(m->out);
    HIPCHK(hipMalloc((void **)&m->out, 3 * sizeof(double)));
    m->out_cap = 1;
  }
  int e = ibm_move(c, vcm, vcm_before);
  if (e != LBM_OK) return e;
  // the IBM step's moments of the inlet/outlet cells (lbm.cpp:1262-1285)
  if ((e = faces_macro(c, true)) != LBM_OK) return e;
  HIPCHK(box_moments(c, false));
  e = ibm_iterate(c, vcm, vcm_before, m->out);
  if (e != LBM_OK) return e;
  double h[3];
  HIPCHK(hipMemcpyAsync(h, m->out, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (out) memcpy(out, h, sizeof h);
  m->ready = 1;
  return LBM_OK;
}

// The lattice launches of an IBM step: force mode 3, the D3Q7 kernels when
// the context carries D3Q7 temperature.
static int ibm_collide(lbm_ctx *c, LaunchCtx &L) {
  if (c->thermal == 2) return q7_launch(c, Q7_COLLIDE, 3);
  HIPCHK(launch_collide(L, c->cur));
  return LBM_OK;
}
static int ibm_stream(lbm_ctx *c, LaunchCtx &L, int mode) {
  if (c->thermal == 2) {
    const int tm = mode == STREAM_COLLIDE ? Q7_FUSED : mode == STREAM_COLLIDE_MACRO ? Q7_FUSED_EMIT : Q7_STREAM;
    return q7_launch(c, tm, 3);
  }
  HIPCHK(launch_stream(L, c->cur, c->oth, mode));
  return LBM_OK;
}

// nsteps of main.cpp:199-205 (IBLBM, Collision, Streaming).  vcm holds the
// centre-of-mass velocity before step 1 and after each step
// ((nsteps+1) x 3 values) or is NULL for a body at rest; out receives
// Drag, Lift and the interior density of every step (nsteps x 3).
int lbm_ibm_run(lbm_ctx *c, int32_t nsteps, const double *vcm, double *out) {
  if (!c || nsteps < 0) return LBM_ERR_ARG;
  if (!c->ibm) return LBM_ERR_STATE;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  if (nsteps == 0) return LBM_OK;
  HIPCHK(hipSetDevice(c->dev));
  lbm_ibm *m = c->ibm;
  if (m->out_cap < nsteps) {
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(m->out);
    HIPCHK(hipMalloc((void **)&m->out, 3 * sizeof(double) * nsteps));
    m->out_cap = nsteps;
  }
  // the first IBM step's moments of the inlet/outlet cells; later steps take
  // them after each streaming (lbm.cpp:1262-1285)
  if (int e = faces_macro(c, true)) return e;
  const double zero[3] = {0, 0, 0};
  for (int s = 0; s < nsteps; s++) {
    const double *vb = vcm ? vcm + 3 * s : zero;
    const double *vn = vcm ? vcm + 3 * (s + 1) : zero;
    int e = ibm_move(c, vn, vb);
    if (e != LBM_OK) return e;
    LaunchCtx L = launch_ctx(c);
    L.force_mode = 3;
    if (c->state == ST_F) {
      // f is in the buffer: IBM, then an in-place collision
      HIPCHK(box_moments(c, false));
      e = ibm_iterate(c, vn, vb, m->out + 3 * s);
      if (e != LBM_OK) return e;
      if ((e = ibm_collide(c, L)) != LBM_OK) return e;
    } else {
      // f* is in the buffer: the box pulls f, the fused launch streams,
      // collides with the IBM result and stores the next f*
      HIPCHK(box_moments(c, true));
      e = ibm_iterate(c, vn, vb, m->out + 3 * s);
      if (e != LBM_OK) return e;
      if ((e = ibm_stream(c, L, s == nsteps - 1 ? STREAM_COLLIDE_MACRO : STREAM_COLLIDE)) != LBM_OK) return e;
      std::swap(c->cur, c->oth);
      c->fstar_ready = 1;
      // the face loops of this launch's streaming and the next IBM step's
      // moments of the inlets (of every inlet/outlet cell on the last step:
      // the values the reference holds after the run)
      if ((e = faces_step(c, true, s == nsteps - 1 ? 2 : 1)) != LBM_OK) return e;
      e = after_stream(c);
      if (e != LBM_OK) return e;
    }
    c->state = ST_FSTAR;
  }
  // the last Streaming of the run
  {
    LaunchCtx L = launch_ctx(c);
    if (int e = ibm_stream(c, L, STREAM_ONLY)) return e;
  }
  std::swap(c->cur, c->oth);
  c->fstar_ready = 1;
  c->state = ST_F;
  int e = faces_step(c, true, 0);  // no IBM step follows inside the run
  if (e == LBM_OK) e = after_stream(c);
  if (e != LBM_OK) return e;
  m->ready = 0;
  if (out) {
    HIPCHK(hipMemcpyAsync(out, m->out, 3 * sizeof(double) * nsteps, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

// ---- a flexible filament coupled through the IBM ---------------------------

int lbm_fsi_setup(lbm_ctx *c, lbm_filament *f, const lbm_ibm_desc *d) {
  if (!c || !f || !d) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  if (c->thermal == 1) return LBM_ERR_UNSUPPORTED;
  if (d->m_max < 0 || d->follow) return LBM_ERR_ARG;
  if (d->scheme != 0) return LBM_ERR_UNSUPPORTED;  // the filament couples through the implicit branch
  double *fx, *fxp, fds;
  int n, fdev;
  filament_state(f, &fx, &fxp, &n, &fds, &fdev);
  if (fdev != c->dev) return LBM_ERR_ARG;
  if (n > FSI_MAX) return LBM_ERR_UNSUPPORTED;  // k_fsi_spread's LDS tables
  HIPCHK(hipSetDevice(c->dev));
  std::vector<double> h(3 * (size_t)n), x(n), y(n), z(n), zero(n, 0.0), ds(n, fds);
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(h.data(), fx, 24 * (size_t)n, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++) {
    x[i] = h[3 * i];
    y[i] = h[3 * i + 1];
    z[i] = h[3 * i + 2];
  }
  ds[0] = ds[n - 1] = fds / 2;  // trapezoid weights of the node spacing
  lbm_ibm_desc dd = *d;
  dd.nmarkers = n;
  dd.moving = 1;
  dd.vsolid = 0;
  int e = lbm_ibm_setup(c, &dd, x.data(), y.data(), z.data(), zero.data(), zero.data(), zero.data(), ds.data());
  if (e != LBM_OK) return e;
  lbm_ibm *m = c->ibm;
  m->fil = f;
  m->near_cell = 1;
  m->ud = (double *)dalloc(m, 24 * (size_t)n);
  m->fsum = (double *)dalloc(m, 24 * (size_t)n);
  if (!m->ud || !m->fsum) return LBM_ERR_NOMEM;
  HIPCHK(hipMemsetAsync(m->fsum, 0, 24 * (size_t)n, c->stream));
  {
    // marker-space rows over all nodes: row n, entry r = node r; the padding
    // entries name the node itself with G = 0
    const int R = (n + 7) / 8 * 8;
    std::vector<int32_t> idx((size_t)n * R);
    for (int i = 0; i < n; i++)
      for (int r = 0; r < R; r++) idx[(size_t)i * R + r] = r < n ? r : i;
    m->R = R;
    m->nbr = (int32_t *)dalloc(m, 4 * idx.size());
    m->H = (double *)dalloc(m, 8 * idx.size());
    if (!m->ub2) m->ub2 = (double *)dalloc(m, 24 * (size_t)n);
    if (!m->rn) m->rn = (double *)dalloc(m, 8 * (size_t)n);
    if (!m->cds) m->cds = (double *)dalloc(m, 8 * (size_t)n);
    if (!m->nbr || !m->H || !m->ub2 || !m->rn || !m->cds) return LBM_ERR_NOMEM;
    HIPCHK(hipMemcpy(m->nbr, idx.data(), 4 * idx.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(m->H, 0, 8 * idx.size()));
    m->fsi_rows = 1;
  }
  hipLaunchKernelGGL(k_fsi_markers, blocks(n), dim3(256), 0, c->stream, n, fx, fxp, m->mx, m->my, m->mz, m->ud);
  HIPCHK(hipGetLastError());
  return ibm_geometry(c);
}

// nsteps coupled steps: markers from the filament, the IBM step with their
// velocities, the lattice step (force mode 3), the filament step under the
// fluid's force -sum F_b per unit length.  out: nsteps x 6 (total force on
// the filament, free-end position).
int lbm_fsi_run(lbm_ctx *c, int32_t nsteps, double *out) {
  if (!c || nsteps < 0) return LBM_ERR_ARG;
  if (!c->ibm || !c->ibm->fil) return LBM_ERR_STATE;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  if (nsteps == 0) return LBM_OK;
  HIPCHK(hipSetDevice(c->dev));
  lbm_ibm *m = c->ibm;
  double *fx, *fxp, fds;
  int n, fdev;
  filament_state(m->fil, &fx, &fxp, &n, &fds, &fdev);
  if (m->out_cap < 2 * nsteps) {
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(m->out);
    HIPCHK(hipMalloc((void **)&m->out, 3 * sizeof(double) * 2 * nsteps));
    m->out_cap = 2 * nsteps;
  }
  if (int e = faces_macro(c, true)) return e;
  for (int s = 0; s < nsteps; s++) {
    int e;
    LaunchCtx L = launch_ctx(c);
    L.force_mode = 3;
    if (c->state == ST_F) {
      HIPCHK(box_moments(c, false));
      if ((e = fsi_iterate(c)) != LBM_OK) return e;
      if ((e = ibm_collide(c, L)) != LBM_OK) return e;
    } else {
      HIPCHK(box_moments(c, true));
      if ((e = fsi_iterate(c)) != LBM_OK) return e;
      if ((e = ibm_stream(c, L, s == nsteps - 1 ? STREAM_COLLIDE_MACRO : STREAM_COLLIDE)) != LBM_OK) return e;
      std::swap(c->cur, c->oth);
      c->fstar_ready = 1;
      // the face loops of this launch's streaming and the next IBM step's
      // moments of the inlets (of every inlet/outlet cell on the last step:
      // the values the reference holds after the run)
      if ((e = faces_step(c, true, s == nsteps - 1 ? 2 : 1)) != LBM_OK) return e;
      if ((e = after_stream(c)) != LBM_OK) return e;
    }
    c->state = ST_FSTAR;
    // the filament step, which also writes the step's total force and tip
    if ((e = filament_step_dev(m->fil, m->fsum, 1.0, 1, c->stream, m->ds, m->out + 6 * s)) != LBM_OK) return e;
  }
  {
    LaunchCtx L = launch_ctx(c);
    if (int e = ibm_stream(c, L, STREAM_ONLY)) return e;
  }
  std::swap(c->cur, c->oth);
  c->fstar_ready = 1;
  c->state = ST_F;
  int e = faces_step(c, true, 0);  // no IBM step follows inside the run
  if (e == LBM_OK) e = after_stream(c);
  if (e != LBM_OK) return e;
  m->ready = 0;
  if (out) HIPCHK(hipMemcpyAsync(out, m->out, 6 * sizeof(double) * nsteps, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

int lbm_ibm_markers(lbm_ctx *c, double *x, double *y, double *z, double *ub, double *Fb) {
  if (!c) return LBM_ERR_ARG;
  if (!c->ibm) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  lbm_ibm *m = c->ibm;
  const size_t nm = (size_t)m->nm;
  if (m->ub_pending) {
    const dim3 gm((unsigned)((m->nm + MK_WG - 1) / MK_WG));
    if (m->ub_from_ub)
      hipLaunchKernelGGL(k_ibm_ub_last<true>, gm, dim3(64 * MK_WG), 0, c->stream, m->nm, m->R2, m->nbr2, m->Q2,
                         m->ub_t, m->ub_u[0], m->ub_u[1], m->ub_u[2], m->ub);
    else
      hipLaunchKernelGGL(k_ibm_ub_last<false>, gm, dim3(64 * MK_WG), 0, c->stream, m->nm, m->R2, m->nbr2, m->Q2,
                         m->ub_t, m->ub_u[0], m->ub_u[1], m->ub_u[2], m->ub);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    m->ub_pending = 0;
  }
  if (x) HIPCHK(hipMemcpy(x, m->mx, 8 * nm, hipMemcpyDeviceToHost));
  if (y) HIPCHK(hipMemcpy(y, m->my, 8 * nm, hipMemcpyDeviceToHost));
  if (z) HIPCHK(hipMemcpy(z, m->mz, 8 * nm, hipMemcpyDeviceToHost));
  if (ub) HIPCHK(hipMemcpy(ub, m->ub, 24 * nm, hipMemcpyDeviceToHost));
  if (Fb) HIPCHK(hipMemcpy(Fb, m->Fb, 24 * nm, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_ibm_box(lbm_ctx *c, int32_t box[6], double center[2]) {
  if (!c) return LBM_ERR_ARG;
  if (!c->ibm) return LBM_ERR_STATE;
  const Box &B = c->ibm->B;
  if (box) {
    box[0] = B.x0; box[1] = B.x1; box[2] = B.y0; box[3] = B.y1; box[4] = B.z0; box[5] = B.z1;
  }
  if (center) {
    center[0] = c->ibm->cx;
    center[1] = c->ibm->cy;
  }
  return LBM_OK;
}

}  // extern "C"

// lbm_init rebuilds the kernel mask: mark the box again
int ibm_remark(lbm_ctx *c) {
  if (!c || !c->ibm || !c->ibm->placed) return LBM_OK;
  const Box &B = c->ibm->B;
  hipLaunchKernelGGL(k_ibm_mark, blocks(B.ncell()), dim3(256), 0, c->stream, c->mask, c->g, B, c->ibm->live, 1);
  HIPCHK(hipGetLastError());
  c->ibm->ready = 0;
  return LBM_OK;
}

bool ibm_active(const lbm_ctx *c) { return c && c->ibm; }
bool ibm_ready(const lbm_ctx *c) { return c && c->ibm && c->ibm->ready; }
void ibm_consume(lbm_ctx *c) {
  if (c && c->ibm) c->ibm->ready = 0;
}

// Whether a checkpoint of the lattice alone resumes this context's IBM runs:
// a rigid body that never moved, between IBLBM calls (every IBM step
// recomputes its markers' velocities and forces from the populations).
bool ibm_checkpointable(const lbm_ctx *c) {
  const lbm_ibm *m = c->ibm;
  return !m || (!m->fil && !m->displaced && !m->ready && !m->moving);
}
