// lbm_probe.hip -- a plain device copy of known size, and the storage
// formats' round trip.
//
// The copy gives the practical HBM ceiling the step kernel is compared with,
// and the reference point for scaling hardware byte counters (a copy of B
// bytes must read B and write B).  16 bytes per lane, one element per thread.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm_hip.h"
#include "lbm_device.h"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(const f32x4 *__restrict__ a, f32x4 *__restrict__ b,
                                                uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// value -> storage encoding -> compute type, as the step kernels store and
// load populations
template <class S>
__global__ __launch_bounds__(256) void k_roundtrip(const double *__restrict__ in, double *__restrict__ out,
                                                   uint64_t n) {
  using C = typename lbm::Store<S>::C;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr double K = lbm::Store<S>::K;
  const S e = lbm::Store<S>::enc((C)(in[i] * K));
  out[i] = (double)lbm::Store<S>::ld(&e) / K;
}

}  // namespace

extern "C" int lbm_storage_roundtrip(int32_t device, int32_t storage, const double *in, double *out,
                                     int64_t n) {
  if (n < 0 || (n > 0 && (!in || !out)) || storage < LBM_STORE_F64 || storage > LBM_STORE_F16C)
    return LBM_ERR_ARG;
  if (n == 0) return LBM_OK;
  if (hipSetDevice(device) != hipSuccess) return LBM_ERR_HIP;
  double *d = nullptr;
  if (hipMalloc((void **)&d, (size_t)n * 16) != hipSuccess) return LBM_ERR_NOMEM;
  int rc = LBM_OK;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (hipMemcpy(d, in, (size_t)n * 8, hipMemcpyHostToDevice) != hipSuccess) {
    rc = LBM_ERR_HIP;
  } else {
    switch (storage) {
      case LBM_STORE_F64: hipLaunchKernelGGL(k_roundtrip<double>, grid, dim3(256), 0, 0, d, d + n, (uint64_t)n); break;
      case LBM_STORE_F32: hipLaunchKernelGGL(k_roundtrip<float>, grid, dim3(256), 0, 0, d, d + n, (uint64_t)n); break;
      case LBM_STORE_F16S: hipLaunchKernelGGL(k_roundtrip<__half>, grid, dim3(256), 0, 0, d, d + n, (uint64_t)n); break;
      default: hipLaunchKernelGGL(k_roundtrip<lbm::f16c_t>, grid, dim3(256), 0, 0, d, d + n, (uint64_t)n); break;
    }
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(out, d + n, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess)
      rc = LBM_ERR_HIP;
  }
  (void)hipFree(d);
  return rc;
}

extern "C" int lbm_copy_probe(int32_t device, int64_t bytes, int32_t reps, double *gbs) {
  if (bytes < 16 || reps < 1 || !gbs) return LBM_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return LBM_ERR_HIP;
  const uint64_t n = (uint64_t)bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return LBM_ERR_NOMEM;
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return LBM_ERR_NOMEM;
  }
  int rc = LBM_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = LBM_ERR_HIP;
  } else {
    hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, 0, (const f32x4 *)a, (f32x4 *)b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++)
      hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, 0, (const f32x4 *)a, (f32x4 *)b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = LBM_ERR_HIP;
  }
  if (rc == LBM_OK) *gbs = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
