// lbm_noalloc.hip -- lbm_reduce_fp64 for callers that cannot allocate device
// memory.
//
// lbm_reduce_fp64 sums into the partial and result buffers its context
// allocated at lbm_create.  This variant sums into one file-scope device
// array instead (`scratch`, NOALLOC_SCRATCH_DOUBLES doubles, held by every
// device the library runs on until the process ends), so the call itself
// allocates and frees nothing.  Its device address is looked up once per
// device, with that device current, on the first call there; kernels take
// that address as an argument (the array's own name is a host-side stand-in
// in host code, not a device address).
//
// Every call on a device shares that device's array and leaves its last
// partials there: the kernels write exactly the partial slots the sum reads,
// so nothing needs clearing, but two calls that overlap on one device (two
// host threads, or two contexts' streams) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lbm_ctx.h"
#include "lbm_launch.h"

using namespace lbm;

namespace {

// partial sums of up to NOALLOC_MAX_BLOCKS fixed 4096-cell chunks (three
// components for the momentum), then the 3 results: lattices of up to
// 65536 * 4096 = 268,435,456 padded cells (a 640^3 box)
constexpr int NOALLOC_MAX_BLOCKS = 65536;
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 3 * (size_t)NOALLOC_MAX_BLOCKS + 4;

__device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
double *g_scratch[MAX_DEVICES] = {};

// this device's address of `scratch` (the device must be current)
int scratch_of(int dev, double **out) {
  if (dev < 0 || dev >= MAX_DEVICES) return LBM_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LBM_ERR_HIP;
    g_scratch[dev] = (double *)p;
  }
  *out = g_scratch[dev];
  return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_reduce_fp64_noalloc(lbm_ctx *c, int32_t what, double *out) {
  if (!c || !out || what < 0 || what > 2) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  const int nb = reduce_blocks(c->g.ncell);
  if (nb > NOALLOC_MAX_BLOCKS) return LBM_ERR_UNSUPPORTED;  // beyond the array's size
  HIPCHK(hipSetDevice(c->dev));
  double *s = nullptr;
  if (int e = scratch_of(c->dev, &s)) return e;
  double *partial = s, *result = s + 3 * (size_t)NOALLOC_MAX_BLOCKS;
  HIPCHK(launch_reduce(launch_ctx(c), what, partial, result));
  const int nc = what == 1 ? 3 : 1;
  HIPCHK(hipMemcpyAsync(out, result, sizeof(double) * nc, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

}  // extern "C"
