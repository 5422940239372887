// lbm_ibm.hip -- the implicit diffuse-interface immersed-boundary step
// (IBLBM, lbm.cpp:1135-1417, branch Implicit_Diffuse_Interface_Scheme) on
// gfx950, generalised from the reference's 2D slab (box z = {1}) to a 3D box.
//
// Per step:
//   k_ibm_move        trapezoidal marker motion, u_des = V_CM (lbm.cpp:1154-1161)
//   k_ibm_box_moments unforced rho, u = sum f c / rho, F = 0 of the box cells
//                     (lbm.cpp:1262-1285), pulled from f* on the fused path
//   k_ibm_interp      u_b = sum u phi phi phi over the 4x4x4 stencil clipped to
//                     the box (lbm.cpp:1288-1309), in the reference's i,j,k order
//   m = 1..m_max:
//     k_ibm_force     F_b = 2 rho (u_des - u_b)/dt with rho at the box point
//                     farthest along the marker normal; drag/lift partials
//                     (lbm.cpp:1314-1329)
//     k_ibm_spread    Fm = sum F_b ds phi phi phi per box cell, F -= Fm,
//                     u += Fm/(2 rho) (lbm.cpp:1332-1360).  Atomic-free: each
//                     workgroup owns a tile of the box, stages the markers
//                     whose stencil reaches the tile in LDS, sorted by marker
//                     index so every cell sums in the reference's n order
//     k_ibm_interp    (lbm.cpp:1363-1387)
//   k_ibm_finish      FP64 drag, lift (tree sums of the partials), the mean
//                     density of the interior box cells and the added-mass
//                     term (lbm.cpp:1105-1132, 1390-1398)
// The collision then takes rho, u, F from the box arrays (force mode 3).
//
// The marker arithmetic is FP64 with a*b+c contraction off, so each marker
// and box value is rounded as the reference's C++ rounds it.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_step.h"

#pragma clang fp contract(off)

using namespace lbm;

namespace {

constexpr int TX = 8, TY = 8, TZ = 4;  // spreading tile (256 cells)
constexpr int TILE_MAX = 2048;         // markers reaching one tile (LDS sort bound)
constexpr int FIN_THREADS = 256;

struct Box {
  int x0, x1, y0, y1, z0, z1;  // inclusive, reference indices (i,j,k)
  int nx, ny, nz;
  int tnx, tny, tnz;
  __host__ __device__ int ncell() const { return nx * ny * nz; }
  __host__ __device__ int ntile() const { return tnx * tny * tnz; }
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

// stencil range of a marker along one axis, clipped to the box:
// floor(m)-1 .. floor(m)+2 holds every point with phi(r) != 0
__device__ __forceinline__ void stencil(double m, int lo, int hi, int &a, int &b) {
  const int f = (int)floor(m);
  a = max(f - 1, lo);
  b = min(f + 2, hi);
}

__global__ void k_ibm_tile_count(Box B, int nm, const double *__restrict__ mx,
                                 const double *__restrict__ my, const double *__restrict__ mz,
                                 int32_t *__restrict__ count) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  int a0, a1, b0, b1, c0, c1;
  stencil(mx[n], B.x0, B.x1, a0, a1);
  stencil(my[n], B.y0, B.y1, b0, b1);
  stencil(mz[n], B.z0, B.z1, c0, c1);
  if (a0 > a1 || b0 > b1 || c0 > c1) return;
  for (int tz = (c0 - B.z0) / TZ; tz <= (c1 - B.z0) / TZ; tz++)
    for (int ty = (b0 - B.y0) / TY; ty <= (b1 - B.y0) / TY; ty++)
      for (int tx = (a0 - B.x0) / TX; tx <= (a1 - B.x0) / TX; tx++)
        atomicAdd(&count[tx + B.tnx * (ty + B.tny * tz)], 1);
}

__global__ void k_ibm_tile_fill(Box B, int nm, const double *__restrict__ mx,
                                const double *__restrict__ my, const double *__restrict__ mz,
                                const int32_t *__restrict__ start, int32_t *__restrict__ cursor,
                                int32_t *__restrict__ list) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  int a0, a1, b0, b1, c0, c1;
  stencil(mx[n], B.x0, B.x1, a0, a1);
  stencil(my[n], B.y0, B.y1, b0, b1);
  stencil(mz[n], B.z0, B.z1, c0, c1);
  if (a0 > a1 || b0 > b1 || c0 > c1) return;
  for (int tz = (c0 - B.z0) / TZ; tz <= (c1 - B.z0) / TZ; tz++)
    for (int ty = (b0 - B.y0) / TY; ty <= (b1 - B.y0) / TY; ty++)
      for (int tx = (a0 - B.x0) / TX; tx <= (a1 - B.x0) / TX; tx++) {
        const int t = tx + B.tnx * (ty + B.tny * tz);
        list[start[t] + atomicAdd(&cursor[t], 1)] = n;
      }
}

// Sorts each tile's marker list ascending (bitonic sort in LDS), so that
// the spreading sums follow the reference's marker order.
__global__ __launch_bounds__(256) void k_ibm_tile_sort(const int32_t *__restrict__ start,
                                                       int32_t *__restrict__ list) {
  __shared__ int32_t s[TILE_MAX];
  const int t = blockIdx.x;
  const int b = start[t], len = start[t + 1] - b;
  if (len <= 1) return;
  int p2 = 1;
  while (p2 < len) p2 <<= 1;
  for (int i = threadIdx.x; i < p2; i += blockDim.x) s[i] = i < len ? list[b + i] : 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= p2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < p2; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          if ((s[i] > s[l]) == up) {
            const int32_t tmp = s[i];
            s[i] = s[l];
            s[l] = tmp;
          }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < len; i += blockDim.x) list[b + i] = s[i];
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

// Unforced moments of the box cells, lbm.cpp:1262-1285, summed in l order
// in FP64.  PULL: the buffer holds f* of the previous step and the cell's
// f is pulled first (the fused path); otherwise the buffer holds f.  Box
// cells that are not fluid keep their set-up rho, u (their density is read
// by the nearest-point force and the interior average).
template <class S, bool PULL>
__global__ __launch_bounds__(256) void k_ibm_box_moments(const Planes<S> P, const uint8_t *__restrict__ mask,
                                                         KArgs<typename Store<S>::C> a, Box B,
                                                         BoxArrays o) {
  using C = typename Store<S>::C;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B.ncell()) return;
  const int x = B.x0 + b % B.nx, y = B.y0 + (b / B.nx) % B.ny, z = B.z0 + b / (B.nx * B.ny);
  const uint32_t idx = lattice_idx(a.g, x, y, z);
  const uint8_t mk = mask[idx];
  o.Fx[b] = 0.0;
  o.Fy[b] = 0.0;
  o.Fz[b] = 0.0;
  if (!(mk & M_FLUID)) {
    o.rho[b] = (double)a.rho[idx];
    o.ux[b] = (double)a.ux[idx];
    o.uy[b] = (double)a.uy[idx];
    o.uz[b] = (double)a.uz[idx];
    return;
  }
  C f[Q];
  if constexpr (PULL) {
    pull_cell<S>(P, mk, idx, a, f);
  } else {
    for_dirs([&](auto M) {
      constexpr int m = decltype(M)::value;
      f[m] = Store<S>::ld(P.in[m] + idx);
    });
  }
  double rho = 0, rx = 0, ry = 0, rz = 0;
  for_dirs([&](auto M) {
    constexpr int m = decltype(M)::value;
    double v = (double)f[m];
    if constexpr (Store<S>::shifted) v = v / (double)Store<S>::K + W[m];
    rho += v;
    rx += v * CX[m];
    ry += v * CY[m];
    rz += v * CZ[m];
  });
  o.rho[b] = rho;
  o.ux[b] = rx / rho;
  o.uy[b] = ry / rho;
  o.uz[b] = rz / rho;
}

// u_b(n) = sum over the clipped stencil of u * phi(rx) phi(ry) phi(rz) in
// i, j, k order (lbm.cpp:1295-1306); non-IBM cells (solids) are skipped.
__global__ void k_ibm_interp(Box B, int nm, const double *__restrict__ mx,
                             const double *__restrict__ my, const double *__restrict__ mz,
                             BoxArrays o, double *__restrict__ ub) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  const double px = mx[n], py = my[n], pz = mz[n];
  int a0, a1, b0, b1, c0, c1;
  stencil(px, B.x0, B.x1, a0, a1);
  stencil(py, B.y0, B.y1, b0, b1);
  stencil(pz, B.z0, B.z1, c0, c1);
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = a0; i <= a1; i++) {
    const double kx = peskin((i * 1.0 - px) / 1.0);
    for (int j = b0; j <= b1; j++) {
      const double ky = peskin((j * 1.0 - py) / 1.0);
      for (int k = c0; k <= c1; k++) {
        const int b = (i - B.x0) + B.nx * ((j - B.y0) + B.ny * (k - B.z0));
        if (!o.live[b]) continue;
        const double kz = peskin((k * 1.0 - pz) / 1.0);
        s0 += o.ux[b] * kx * ky * kz;
        s1 += o.uy[b] * kx * ky * kz;
        s2 += o.uz[b] * kx * ky * kz;
      }
    }
  }
  ub[3 * n + 0] = s0;
  ub[3 * n + 1] = s1;
  ub[3 * n + 2] = s2;
}

// marker force and drag/lift partials of iteration m, lbm.cpp:1314-1323
__global__ void k_ibm_force(int nm, const int32_t *__restrict__ near, const double *__restrict__ brho,
                            const double *__restrict__ ub, const double *__restrict__ ds, double u0,
                            double u1, double u2, double *__restrict__ Fb, double *__restrict__ dl) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nm) return;
  const double dt = 1.0, dx = 1.0, dy = 1.0;
  const int nb = near[n];
  const double rho = nb >= 0 ? brho[nb] : nan("");
  const double f0 = 2 * rho * (u0 - ub[3 * n + 0]) / dt;
  const double f1 = 2 * rho * (u1 - ub[3 * n + 1]) / dt;
  const double f2 = 2 * rho * (u2 - ub[3 * n + 2]) / dt;
  Fb[3 * n + 0] = f0;
  Fb[3 * n + 1] = f1;
  Fb[3 * n + 2] = f2;
  dl[2 * n + 0] = -f0 * dx * ds[n];
  dl[2 * n + 1] = -f1 * dy * ds[n];
}

// Spreading and velocity correction, lbm.cpp:1332-1360.  One workgroup per
// tile of TX x TY x TZ box cells; the markers reaching the tile (sorted by
// index) are staged in LDS in chunks and every live cell sums their
// contributions in marker order.
__global__ __launch_bounds__(256) void k_ibm_spread(Box B, const int32_t *__restrict__ start,
                                                    const int32_t *__restrict__ list,
                                                    const double *__restrict__ mx,
                                                    const double *__restrict__ my,
                                                    const double *__restrict__ mz,
                                                    const double *__restrict__ Fb,
                                                    const double *__restrict__ ds, BoxArrays o) {
  constexpr int CH = 256;
  __shared__ double sx[CH], sy[CH], sz[CH], sf0[CH], sf1[CH], sf2[CH], sds[CH];
  const int t = blockIdx.x;
  const int tx = t % B.tnx, ty = (t / B.tnx) % B.tny, tz = t / (B.tnx * B.tny);
  const int lx = threadIdx.x % TX, ly = (threadIdx.x / TX) % TY, lz = threadIdx.x / (TX * TY);
  const int bx = tx * TX + lx, by = ty * TY + ly, bz = tz * TZ + lz;
  const bool mine = bx < B.nx && by < B.ny && bz < B.nz;
  const int b = bx + B.nx * (by + B.ny * bz);
  const bool live = mine && o.live[b];
  const double i = B.x0 + bx, j = B.y0 + by, k = B.z0 + bz;
  double m0 = 0, m1 = 0, m2 = 0;
  const int l0 = start[t], l1 = start[t + 1];
  for (int c = l0; c < l1; c += CH) {
    const int len = min(CH, l1 - c);
    __syncthreads();
    if ((int)threadIdx.x < len) {
      const int n = list[c + threadIdx.x];
      sx[threadIdx.x] = mx[n];
      sy[threadIdx.x] = my[n];
      sz[threadIdx.x] = mz[n];
      sf0[threadIdx.x] = Fb[3 * n + 0];
      sf1[threadIdx.x] = Fb[3 * n + 1];
      sf2[threadIdx.x] = Fb[3 * n + 2];
      sds[threadIdx.x] = ds[n];
    }
    __syncthreads();
    if (live) {
      for (int q = 0; q < len; q++) {
        const double rx = (i * 1.0 - sx[q]) / 1.0, ry = (j * 1.0 - sy[q]) / 1.0,
                     rz = (k * 1.0 - sz[q]) / 1.0;
        if (fabs(rx) >= 2 || fabs(ry) >= 2 || fabs(rz) >= 2) continue;  // phi = 0
        const double kx = peskin(rx), ky = peskin(ry), kz = peskin(rz);
        m0 += sf0[q] * sds[q] * kx * ky * kz;
        m1 += sf1[q] * sds[q] * kx * ky * kz;
        m2 += sf2[q] * sds[q] * kx * ky * kz;
      }
    }
  }
  if (live) {
    const double dt = 1.0;
    o.Fx[b] += -m0;
    o.Fy[b] += -m1;
    o.Fz[b] += -m2;
    const double r = o.rho[b];
    o.ux[b] += dt * m0 / (2 * r);
    o.uy[b] += dt * m1 / (2 * r);
    o.uz[b] += dt * m2 / (2 * r);
  }
}

__device__ __forceinline__ double fin_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = FIN_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Drag, lift and the added-mass term, lbm.cpp:1105-1132, 1390-1398.  One
// workgroup, fixed summation order.
__global__ __launch_bounds__(FIN_THREADS) void k_ibm_finish(int nterms, const double *__restrict__ dl,
                                                            Box B, const uint8_t *__restrict__ inside,
                                                            const double *__restrict__ brho,
                                                            double vsolid, double d0, double d1,
                                                            double *__restrict__ out) {
  __shared__ double sh[FIN_THREADS];
  double a = 0, l = 0, r = 0, c = 0;
  for (int q = threadIdx.x; q < nterms; q += FIN_THREADS) {
    a += dl[2 * q + 0];
    l += dl[2 * q + 1];
  }
  for (int b = threadIdx.x; b < B.ncell(); b += FIN_THREADS)
    if (inside[b]) {
      r += brho[b];
      c += 1.0;
    }
  const double drag = fin_sum(a, sh), lift = fin_sum(l, sh), rs = fin_sum(r, sh),
               cnt = fin_sum(c, sh);
  if (threadIdx.x == 0) {
    const double dt = 1.0;
    const double rho = rs / cnt;  // NaN when no cell is interior, as lbm.cpp:1129
    out[0] = drag + rho * vsolid * d0 / dt;
    out[1] = lift + rho * vsolid * d1 / dt;
    out[2] = rho;
  }
}

__global__ void k_fill_i32(int32_t *p, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(1024) void k_scan_i32(const int32_t *__restrict__ count, int n,
                                                   int32_t *__restrict__ start) {
  // single block, n is small (tiles of the IBM box): sequential chunks
  __shared__ int32_t part[1024];
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int lo = threadIdx.x * per, hi = min(n, lo + per);
  int32_t s = 0;
  for (int i = lo; i < hi; i++) s += count[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t acc = 0;
    for (int i = 0; i < (int)blockDim.x; i++) {
      const int32_t v = part[i];
      part[i] = acc;
      acc += v;
    }
    start[n] = acc;
  }
  __syncthreads();
  int32_t acc = part[threadIdx.x];
  for (int i = lo; i < hi; i++) {
    start[i] = acc;
    acc += count[i];
  }
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
  double *ds = nullptr, *ub = nullptr, *Fb = nullptr, *dl = nullptr;
  int32_t *near = nullptr;
  // box arrays (capacity box_cap cells) and tile arrays (tile_cap tiles)
  double *brho = nullptr, *bux = nullptr, *buy = nullptr, *buz = nullptr;
  double *bFx = nullptr, *bFy = nullptr, *bFz = nullptr;
  uint8_t *live = nullptr, *inside = nullptr;
  int32_t *tcount = nullptr, *tstart = nullptr, *tcursor = nullptr, *tlist = nullptr;
  size_t box_cap = 0, tile_cap = 0, tlist_cap = 0;
  double *out = nullptr;  // 3 per step of a run
  int out_cap = 0;
  int ready = 0;          // IBLBM done, the next collision uses force mode 3
  std::vector<void *> allocs;  // marker arrays, freed with the state
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
  dfree(m->live); dfree(m->inside);
  m->box_cap = 0;
}

static void free_tile_arrays(lbm_ibm *m) {
  dfree(m->tcount); dfree(m->tcursor); dfree(m->tstart);
  m->tile_cap = 0;
}

void ibm_free(lbm_ctx *c) {
  if (!c || !c->ibm) return;
  lbm_ibm *m = c->ibm;
  for (void *p : m->allocs) (void)hipFree(p);
  free_box_arrays(m);
  free_tile_arrays(m);
  dfree(m->tlist);
  dfree(m->out);
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
  return a;
}

template <class S>
static hipError_t box_moments_s(lbm_ctx *c, bool pull) {
  lbm_ibm *m = c->ibm;
  const LaunchCtx L = launch_ctx(c);
  const Planes<S> P = planes_of<S>(c, c->cur);
  const auto a = kargs_of<S>(L);
  if (pull)
    hipLaunchKernelGGL((k_ibm_box_moments<S, true>), blocks(m->B.ncell()), dim3(256), 0, c->stream, P, c->mask, a, m->B, box_arrays(m));
  else
    hipLaunchKernelGGL((k_ibm_box_moments<S, false>), blocks(m->B.ncell()), dim3(256), 0, c->stream, P, c->mask, a, m->B, box_arrays(m));
  return hipGetLastError();
}

static hipError_t box_moments(lbm_ctx *c, bool pull) {
  switch (c->storage) {
    case LBM_STORE_F64: return box_moments_s<double>(c, pull);
    case LBM_STORE_F32: return box_moments_s<float>(c, pull);
    default: return box_moments_s<__half>(c, pull);
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

// Puts the IBM box at b: clears the old box's mask bits, grows the box and
// tile arrays when needed, marks the new box and points the collision's
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
  B.tnx = (B.nx + TX - 1) / TX; B.tny = (B.ny + TY - 1) / TY; B.tnz = (B.nz + TZ - 1) / TZ;
  const size_t nb = (size_t)B.ncell(), nt = (size_t)B.ntile();
  if (nb > m->box_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    free_box_arrays(m);
    double **bp[7] = {&m->brho, &m->bux, &m->buy, &m->buz, &m->bFx, &m->bFy, &m->bFz};
    for (int q = 0; q < 7; q++)
      if (hipMalloc((void **)bp[q], 8 * nb) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipMalloc((void **)&m->live, nb) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipMalloc((void **)&m->inside, nb) != hipSuccess) return LBM_ERR_NOMEM;
    for (int q = 0; q < 7; q++) HIPCHK(hipMemsetAsync(*bp[q], 0, 8 * nb, c->stream));
    m->box_cap = nb;
  }
  if (nt > m->tile_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    free_tile_arrays(m);
    if (hipMalloc((void **)&m->tcount, 4 * nt) != hipSuccess ||
        hipMalloc((void **)&m->tcursor, 4 * nt) != hipSuccess ||
        hipMalloc((void **)&m->tstart, 4 * (nt + 1)) != hipSuccess)
      return LBM_ERR_NOMEM;
    m->tile_cap = nt;
  }
  hipLaunchKernelGGL(k_ibm_mark, blocks((int)nb), dim3(256), 0, c->stream, c->mask, c->g, B, m->live, 1);
  HIPCHK(hipGetLastError());
  m->placed = true;
  // the collision of IBM steps reads the box (force mode 3)
  c->box = BoxArgs{B.x0, B.y0, B.z0, B.nx, B.ny, B.nz, m->brho, m->bux, m->buy, m->buz, m->bFx, m->bFy, m->bFz};
  return LBM_OK;
}

// nearest points, interior cells, spreading tiles for the current markers
static int ibm_geometry(lbm_ctx *c) {
  lbm_ibm *m = c->ibm;
  const Box &B = m->B;
  hipLaunchKernelGGL(k_ibm_nearest, dim3(m->nm), dim3(256), 0, c->stream, B, m->mx, m->my, m->mz, m->nx, m->ny, m->nz, m->near);
  hipLaunchKernelGGL(k_ibm_inside, blocks(B.ncell()), dim3(256), 0, c->stream, B, m->nm, m->mx, m->my, m->mz, m->nx, m->ny, m->nz, m->inside);
  const int nt = B.ntile();
  hipLaunchKernelGGL(k_fill_i32, blocks(nt), dim3(256), 0, c->stream, m->tcount, nt, 0);
  hipLaunchKernelGGL(k_fill_i32, blocks(nt), dim3(256), 0, c->stream, m->tcursor, nt, 0);
  hipLaunchKernelGGL(k_ibm_tile_count, blocks(m->nm), dim3(256), 0, c->stream, B, m->nm, m->mx, m->my, m->mz, m->tcount);
  hipLaunchKernelGGL(k_scan_i32, dim3(1), dim3(1024), 0, c->stream, m->tcount, nt, m->tstart);
  HIPCHK(hipGetLastError());
  // list capacity and the per-tile sort bound are checked on the host
  std::vector<int32_t> cnt(nt);
  HIPCHK(hipMemcpyAsync(cnt.data(), m->tcount, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  size_t total = 0;
  for (int t = 0; t < nt; t++) {
    if (cnt[t] > TILE_MAX) return LBM_ERR_UNSUPPORTED;
    total += (size_t)cnt[t];
  }
  if (total > m->tlist_cap) {
    dfree(m->tlist);
    const size_t cap = std::max<size_t>(total, 1);
    if (hipMalloc((void **)&m->tlist, sizeof(int32_t) * cap) != hipSuccess) return LBM_ERR_NOMEM;
    m->tlist_cap = cap;
  }
  hipLaunchKernelGGL(k_ibm_tile_fill, blocks(m->nm), dim3(256), 0, c->stream, B, m->nm, m->mx, m->my, m->mz, m->tstart, m->tcursor, m->tlist);
  hipLaunchKernelGGL(k_ibm_tile_sort, dim3(nt), dim3(256), 0, c->stream, m->tstart, m->tlist);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// One IBLBM call's device work after the box moments: interpolation, m_max
// force/spread/interpolate passes, and the drag/lift/added-mass sums into
// out[0..2].
static int ibm_iterate(lbm_ctx *c, const double vcm[3], const double vb[3], double *out) {
  lbm_ibm *m = c->ibm;
  const Box &B = m->B;
  const BoxArrays o = box_arrays(m);
  hipLaunchKernelGGL(k_ibm_interp, blocks(m->nm, 64), dim3(64), 0, c->stream, B, m->nm, m->mx, m->my, m->mz, o, m->ub);
  for (int it = 0; it < m->m_max; it++) {
    hipLaunchKernelGGL(k_ibm_force, blocks(m->nm, 64), dim3(64), 0, c->stream, m->nm, m->near, m->brho, m->ub, m->ds,
                       vcm[0], vcm[1], vcm[2], m->Fb, m->dl + (size_t)2 * m->nm * it);
    hipLaunchKernelGGL(k_ibm_spread, dim3(B.ntile()), dim3(256), 0, c->stream, B, m->tstart, m->tlist, m->mx, m->my, m->mz, m->Fb, m->ds, o);
    hipLaunchKernelGGL(k_ibm_interp, blocks(m->nm, 64), dim3(64), 0, c->stream, B, m->nm, m->mx, m->my, m->mz, o, m->ub);
  }
  hipLaunchKernelGGL(k_ibm_finish, dim3(1), dim3(FIN_THREADS), 0, c->stream, m->nm * m->m_max, m->dl, B, m->inside, m->brho,
                     m->vsolid, vcm[0] - vb[0], vcm[1] - vb[1], out);
  HIPCHK(hipGetLastError());
  return LBM_OK;
}

// Marker motion of one IBLBM call (lbm.cpp:1154-1166): the trapezoid rule
// for the markers and, when the box follows the body, for its centre; then
// the box and the geometry that depends on the marker positions.
static int ibm_move(lbm_ctx *c, const double vcm[3], const double vb[3]) {
  lbm_ibm *m = c->ibm;
  hipLaunchKernelGGL(k_ibm_move, blocks(m->nm), dim3(256), 0, c->stream, m->nm, m->mx, m->my, m->mz, vcm[0], vcm[1], vcm[2], vb[0], vb[1], vb[2]);
  HIPCHK(hipGetLastError());
  const bool moved = vcm[0] + vb[0] != 0 || vcm[1] + vb[1] != 0 || vcm[2] + vb[2] != 0;
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
  // the nearest points, interior cells and tiles follow a body that moves
  if (m->moving || moved) return ibm_geometry(c);
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
  m->dl = (double *)dalloc(m, 16 * nm * std::max(1, m->m_max));
  m->near = (int32_t *)dalloc(m, 4 * nm);
  if (!m->ub || !m->Fb || !m->dl || !m->near) return LBM_ERR_NOMEM;
  HIPCHK(hipMemsetAsync(m->ub, 0, 24 * nm, c->stream));
  HIPCHK(hipMemsetAsync(m->Fb, 0, 24 * nm, c->stream));
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
  if (d->nmarkers < 1 || d->m_max < 0) return LBM_ERR_ARG;
  if (d->follow && !(d->half >= 0)) return LBM_ERR_ARG;
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
      HIPCHK(launch_collide(L, c->cur));
    } else {
      // f* is in the buffer: the box pulls f, the fused launch streams,
      // collides with the IBM result and stores the next f*
      HIPCHK(box_moments(c, true));
      e = ibm_iterate(c, vn, vb, m->out + 3 * s);
      if (e != LBM_OK) return e;
      HIPCHK(launch_stream(L, c->cur, c->oth, s == nsteps - 1 ? STREAM_COLLIDE_MACRO : STREAM_COLLIDE));
      std::swap(c->cur, c->oth);
      e = after_stream(c);
      if (e != LBM_OK) return e;
    }
    c->state = ST_FSTAR;
  }
  // the last Streaming of the run
  HIPCHK(launch_stream(launch_ctx(c), c->cur, c->oth, STREAM_ONLY));
  std::swap(c->cur, c->oth);
  c->state = ST_F;
  int e = after_stream(c);
  if (e != LBM_OK) return e;
  m->ready = 0;
  if (out) {
    HIPCHK(hipMemcpyAsync(out, m->out, 3 * sizeof(double) * nsteps, hipMemcpyDeviceToHost, c->stream));
  }
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
