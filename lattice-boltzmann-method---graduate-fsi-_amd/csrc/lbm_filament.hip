// lbm_filament.hip -- flexible filament structure solver, FP64 on the
// device (no reference counterpart: the reference's flexible-filament
// structure solver is absent).
//
// Lattice units:  m_l X_tt = d/ds (T dX/ds) - EI d4X/ds4 + G - gamma X_t + F,
// |dX/ds| = 1, nodes i = 0..N-1 at spacing ds = L/(N-1); node 0 clamped at
// X0 with slope e0, the last node free (d2X/ds2 = d3X/ds3 = 0, T = 0, half
// a segment of mass).  One structure step of size h:
//   tangents a_j = (X_{j+1} - X_j)/ds at time n;
//   predictor P_i = 2X_i - Xp_i + h^2/m_l (B_i + G + F_i - gamma (X_i - Xp_i)/h),
//     B = -EI D4 X with ghost nodes X_{-1} = X_1 - 2 ds e0,
//     X_N = 2X_{N-1} - X_{N-2}, X_{N+1} = 2X_N - 2X_{N-2} + X_{N-3};
//   X'_i = P_i + kappa c_i (T_i a_i - T_{i-1} a_{i-1}), kappa = h^2/(m_l ds),
//     c_{N-1} = 2 (half mass), T_{N-1} = 0, X'_0 = X0;
//   segment tensions T_j from |X'_{j+1} - X'_j|^2 = ds^2: three Newton steps
//     from the previous tensions; the Jacobian is tridiagonal (d_j depends
//     on T_{j-1}, T_j, T_{j+1} only) and is solved by parallel cyclic
//     reduction in LDS.
// Bending is explicit; a call advancing dt takes K sub-steps h = dt/K, K
// chosen below the leapfrog limit h <= ds^2 sqrt(m_l/EI)/2 unless given.
//
// One workgroup, one node per lane (N <= FIL_MAX), every sub-step of a call
// in one launch.  The operations and their order are those of the oracle's
// restatement, with a*b+c contraction off, so FP64 results agree bitwise.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/lbm_hip.h"

#pragma clang fp contract(off)

namespace {

constexpr int FIL_MAX = 256;

struct FilArgs {
  int n, m, nsub;
  double ds, ml, ei, gamma, h;
  double G[3], X0[3], e0[3];
  double *x, *xp, *T;   // 3n, 3n, n-1 (device)
  const double *F;      // 3n per unit length, or null
  // a coupled step's output (or null): the total force sum_i F_i w_i with
  // the nodes' weights w, and the free end's position after the step
  const double *w;
  double *out;
};

// parallel cyclic reduction of rows 0..m-1 (a x_{i-1} + b x_i + c x_{i+1} =
// d), the system in LDS buffer 0; returns x_i for lane i < m
__device__ double pcr(int m, double (*A)[FIL_MAX], double (*B)[FIL_MAX], double (*Cc)[FIL_MAX],
                      double (*D)[FIL_MAX]) {
  const int i = threadIdx.x;
  int cur = 0;
  for (int s = 1; s < m; s <<= 1) {
    if (i < m) {
      double a2 = 0, c2 = 0, bb = B[cur][i], dd = D[cur][i];
      if (i - s >= 0) {
        const double al = -A[cur][i] / B[cur][i - s];
        a2 = al * A[cur][i - s];
        bb = bb + al * Cc[cur][i - s];
        dd = dd + al * D[cur][i - s];
      }
      if (i + s < m) {
        const double ga = -Cc[cur][i] / B[cur][i + s];
        c2 = ga * Cc[cur][i + s];
        bb = bb + ga * A[cur][i + s];
        dd = dd + ga * D[cur][i + s];
      }
      A[cur ^ 1][i] = a2;
      B[cur ^ 1][i] = bb;
      Cc[cur ^ 1][i] = c2;
      D[cur ^ 1][i] = dd;
    }
    __syncthreads();
    cur ^= 1;
  }
  double x = 0;
  if (i < m) x = D[cur][i] / B[cur][i];
  __syncthreads();
  return x;
}

__global__ __launch_bounds__(FIL_MAX) void k_fil_step(FilArgs f, int ncalls) {
  __shared__ double sX[3][FIL_MAX], sA[3][FIL_MAX], sP[3][FIL_MAX];
  __shared__ double sT[FIL_MAX];
  __shared__ double pa[2][FIL_MAX], pb[2][FIL_MAX], pc[2][FIL_MAX], pd[2][FIL_MAX];
  const int i = threadIdx.x, n = f.n, m = f.m;
  const double ds = f.ds, ds4 = ds * ds * ds * ds, ml = f.ml, h = f.h;
  const double kap = h * h / (ml * ds);
  double X[3] = {0, 0, 0}, Xp[3] = {0, 0, 0}, Fi[3] = {0, 0, 0};
  if (i < n) {
    for (int d = 0; d < 3; d++) {
      X[d] = f.x[3 * i + d];
      Xp[d] = f.xp[3 * i + d];
      if (f.F) Fi[d] = f.F[3 * i + d];
    }
  }
  double T = (i < m) ? f.T[i] : 0.0;
  for (int step = 0; step < ncalls * f.nsub; step++) {
    if (i < n)
      for (int d = 0; d < 3; d++) sX[d][i] = X[d];
    if (i < m) sT[i] = T;
    __syncthreads();
    if (i < m)
      for (int d = 0; d < 3; d++) sA[d][i] = (sX[d][i + 1] - sX[d][i]) / ds;
    // predictor
    if (i < n) {
      for (int d = 0; d < 3; d++) {
        if (i == 0) {
          sP[d][0] = f.X0[d];
          continue;
        }
        // x(k) with the ghost nodes -1, N, N+1
        auto xg = [&](int k) -> double {
          if (k == -1) return sX[d][1] - 2 * ds * f.e0[d];
          if (k == n) return 2 * sX[d][n - 1] - sX[d][n - 2];
          if (k == n + 1) return 2 * (2 * sX[d][n - 1] - sX[d][n - 2]) - 2 * sX[d][n - 2] + sX[d][n - 3];
          return sX[d][k];
        };
        const double d4 = (xg(i - 2) - 4 * xg(i - 1) + 6 * xg(i) - 4 * xg(i + 1) + xg(i + 2)) / ds4;
        const double B = -f.ei * d4;
        const double v = (X[d] - Xp[d]) / h;
        sP[d][i] = 2 * X[d] - Xp[d] + h * h / ml * (B + f.G[d] + Fi[d] - f.gamma * v);
      }
    }
    __syncthreads();
    // Newton on the segment lengths
    for (int it = 0; it < 3; it++) {
      if (i < m) {
        const int j = i;
        const double cj1 = (j + 1 == n - 1) ? 2.0 : 1.0, cj = 1.0;
        const double Tj1 = (j + 1 < m) ? sT[j + 1] : 0.0, Tj = sT[j], Tjm = (j >= 1) ? sT[j - 1] : 0.0;
        double dv[3];
        for (int d = 0; d < 3; d++) {
          const double a1 = (j + 1 < m) ? sA[d][j + 1] : 0.0, a0 = sA[d][j], am = (j >= 1) ? sA[d][j - 1] : 0.0;
          const double xn1 = sP[d][j + 1] + kap * cj1 * (Tj1 * a1 - Tj * a0);
          const double xn0 = (j == 0) ? sP[d][0] : sP[d][j] + kap * cj * (Tj * a0 - Tjm * am);
          dv[d] = xn1 - xn0;
        }
        const double rr = -((dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]) - ds * ds);
        const double cjp = (j == 0) ? 0.0 : cj;
        double s1 = 0, s0 = 0, sm = 0;
        for (int d = 0; d < 3; d++) {
          const double a1 = (j + 1 < m) ? sA[d][j + 1] : 0.0, a0 = sA[d][j], am = (j >= 1) ? sA[d][j - 1] : 0.0;
          s1 += dv[d] * a1;
          s0 += dv[d] * a0;
          sm += dv[d] * am;
        }
        pc[0][j] = (j + 1 < m) ? 2 * kap * cj1 * s1 : 0.0;
        pb[0][j] = -2 * kap * (cj1 + cjp) * s0;
        pa[0][j] = (j >= 1) ? 2 * kap * cj * sm : 0.0;
        pd[0][j] = rr;
      }
      __syncthreads();
      const double dT = pcr(m, pa, pb, pc, pd);
      if (i < m) {
        T += dT;
        sT[i] = T;
      }
      __syncthreads();
    }
    // corrector
    if (i < n) {
      if (i == 0) {
        for (int d = 0; d < 3; d++) {
          Xp[d] = X[d];
          X[d] = f.X0[d];
        }
      } else {
        const double ci = (i == n - 1) ? 2.0 : 1.0;
        const double Ti = (i < m) ? sT[i] : 0.0, Tim = sT[i - 1];
        for (int d = 0; d < 3; d++) {
          const double ai = (i < m) ? sA[d][i] : 0.0, aim = sA[d][i - 1];
          const double xn = sP[d][i] + kap * ci * (Ti * ai - Tim * aim);
          Xp[d] = X[d];
          X[d] = xn;
        }
      }
    }
    __syncthreads();
  }
  if (i < n)
    for (int d = 0; d < 3; d++) {
      f.x[3 * i + d] = X[d];
      f.xp[3 * i + d] = Xp[d];
    }
  if (i < m) f.T[i] = T;
  if (f.out && f.F) {
    // lane-strided partial sums over the nodes, then a wavefront tree
    if (i < 64) {
      double s0 = 0, s1 = 0, s2 = 0;
      for (int k = i; k < n; k += 64) {
        const double wk = f.w[k];
        s0 += f.F[3 * k + 0] * wk;
        s1 += f.F[3 * k + 1] * wk;
        s2 += f.F[3 * k + 2] * wk;
      }
      for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_down(s0, o, 64);
        s1 += __shfl_down(s1, o, 64);
        s2 += __shfl_down(s2, o, 64);
      }
      if (i == 0) {
        f.out[0] = s0;
        f.out[1] = s1;
        f.out[2] = s2;
      }
    }
    if (i == n - 1)
      for (int d = 0; d < 3; d++) f.out[3 + d] = X[d];
  }
}

}  // namespace

struct lbm_filament {
  int dev = 0;
  lbm_filament_desc d{};
  int substeps = 1;
  double ds = 1;
  double *x = nullptr, *xp = nullptr, *T = nullptr, *F = nullptr;
  hipStream_t stream = nullptr;
};

static void fil_free(lbm_filament *f) {
  double **p[] = {&f->x, &f->xp, &f->T, &f->F};
  for (double **q : p) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
  if (f->stream) (void)hipStreamDestroy(f->stream);
  f->stream = nullptr;
}

#define FCHK(x)                                 \
  do {                                          \
    if ((x) != hipSuccess) return LBM_ERR_HIP;  \
  } while (0)

// the device state the lattice coupling reads (lbm_ibm.hip)
void filament_state(lbm_filament *f, double **x, double **xp, int *n, double *ds, int *dev) {
  *x = f->x;
  *xp = f->xp;
  *n = f->d.nnodes;
  *ds = f->ds;
  *dev = f->dev;
}

// device-pointer form for the coupled step: force already on the device
int filament_step_dev(lbm_filament *f, const double *F_dev, double dt, int32_t ncalls, hipStream_t s,
                      const double *w, double *out) {
  FilArgs a;
  a.n = f->d.nnodes;
  a.m = a.n - 1;
  a.nsub = f->substeps;
  a.ds = f->ds;
  a.ml = f->d.line_density;
  a.ei = f->d.bending;
  a.gamma = f->d.damping;
  a.h = dt / f->substeps;
  for (int q = 0; q < 3; q++) {
    a.G[q] = f->d.body_force[q];
    a.X0[q] = f->d.clamp[q];
    a.e0[q] = f->d.tangent[q];
  }
  a.x = f->x;
  a.xp = f->xp;
  a.T = f->T;
  a.F = F_dev;
  a.w = w;
  a.out = out;
  hipLaunchKernelGGL(k_fil_step, dim3(1), dim3(FIL_MAX), 0, s, a, (int)ncalls);
  return hipGetLastError() == hipSuccess ? LBM_OK : LBM_ERR_HIP;
}


extern "C" {

int lbm_filament_create(int32_t device, const lbm_filament_desc *d, lbm_filament **out) {
  if (!d || !out) return LBM_ERR_ARG;
  *out = nullptr;
  const int n = d->nnodes;
  if (n < 4 || !(d->length > 0) || !(d->line_density > 0) || !(d->bending >= 0) || !(d->damping >= 0))
    return LBM_ERR_ARG;
  if (n > FIL_MAX) return LBM_ERR_UNSUPPORTED;
  const double tn = sqrt(d->tangent[0] * d->tangent[0] + d->tangent[1] * d->tangent[1] + d->tangent[2] * d->tangent[2]);
  if (!(fabs(tn - 1.0) < 1e-12)) return LBM_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return LBM_ERR_HIP;
  lbm_filament *f = new lbm_filament();
  f->dev = device;
  f->d = *d;
  f->ds = d->length / (n - 1);
  int K = d->substeps;
  if (K <= 0) {
    const double hc = d->bending > 0 ? f->ds * f->ds * sqrt(d->line_density / d->bending) / 2 : 1.0;
    K = (int)ceil(1.0 / (0.8 * hc));
    if (K < 1) K = 1;
  }
  f->substeps = K;
  auto fail = [&](int code) {
    fil_free(f);
    delete f;
    return code;
  };
  if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) return fail(LBM_ERR_HIP);
  if (hipMalloc((void **)&f->x, 24 * (size_t)n) != hipSuccess || hipMalloc((void **)&f->xp, 24 * (size_t)n) != hipSuccess ||
      hipMalloc((void **)&f->T, 8 * (size_t)n) != hipSuccess || hipMalloc((void **)&f->F, 24 * (size_t)n) != hipSuccess)
    return fail(LBM_ERR_NOMEM);
  std::vector<double> h(3 * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int q = 0; q < 3; q++) h[3 * i + q] = d->clamp[q] + i * f->ds * d->tangent[q];
  if (hipMemcpy(f->x, h.data(), 24 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(f->xp, h.data(), 24 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(f->T, 0, 8 * (size_t)n) != hipSuccess)
    return fail(LBM_ERR_HIP);
  *out = f;
  return LBM_OK;
}

int lbm_filament_destroy(lbm_filament *f) {
  if (!f) return LBM_ERR_ARG;
  (void)hipSetDevice(f->dev);
  fil_free(f);
  delete f;
  return LBM_OK;
}

int lbm_filament_substeps(const lbm_filament *f) { return f ? f->substeps : -1; }

int lbm_filament_set(lbm_filament *f, const double *x_now, const double *x_prev) {
  if (!f || !x_now) return LBM_ERR_ARG;
  FCHK(hipSetDevice(f->dev));
  const size_t b = 24 * (size_t)f->d.nnodes;
  FCHK(hipStreamSynchronize(f->stream));
  FCHK(hipMemcpy(f->x, x_now, b, hipMemcpyHostToDevice));
  FCHK(hipMemcpy(f->xp, x_prev ? x_prev : x_now, b, hipMemcpyHostToDevice));
  return LBM_OK;
}

int lbm_filament_get(lbm_filament *f, double *x_now, double *x_prev, double *tension) {
  if (!f) return LBM_ERR_ARG;
  FCHK(hipSetDevice(f->dev));
  FCHK(hipStreamSynchronize(f->stream));
  const size_t b = 24 * (size_t)f->d.nnodes;
  if (x_now) FCHK(hipMemcpy(x_now, f->x, b, hipMemcpyDeviceToHost));
  if (x_prev) FCHK(hipMemcpy(x_prev, f->xp, b, hipMemcpyDeviceToHost));
  if (tension) FCHK(hipMemcpy(tension, f->T, 8 * (size_t)(f->d.nnodes - 1), hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_filament_step(lbm_filament *f, const double *force, double dt, int32_t nsteps) {
  if (!f || nsteps < 0 || !(dt > 0)) return LBM_ERR_ARG;
  if (nsteps == 0) return LBM_OK;
  FCHK(hipSetDevice(f->dev));
  if (force) FCHK(hipMemcpyAsync(f->F, force, 24 * (size_t)f->d.nnodes, hipMemcpyHostToDevice, f->stream));
  if (int e = filament_step_dev(f, force ? f->F : nullptr, dt, nsteps, f->stream, nullptr, nullptr)) return e;
  FCHK(hipStreamSynchronize(f->stream));
  return LBM_OK;
}

}  // extern "C"
