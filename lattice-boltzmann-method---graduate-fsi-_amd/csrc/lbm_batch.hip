// lbm_batch.hip -- ensembles of independent lattices advanced by one launch
// per step, and the launch-shape tuners.
//
// A batch holds contexts of one storage precision and force mode.  The
// tile table maps workgroup tiles of blockDim.x cells to (lattice, first
// cell); the per-lattice plane pointers and kernel arguments of both buffer
// parities live in device memory (LatArgs).  The two launches of a step pair
// (parity 0 and 1) are captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_launch.h"

using namespace lbm;

struct lbm_batch {
  std::vector<lbm_ctx *> ctx;
  int dev = 0, storage = 0, force_mode = 0;
  int block = 256, per_cu = 0;  // launch shape
  int ncu = 256;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  size_t args_size = 0;
  std::vector<unsigned char> h_args;
  void *d_args = nullptr;
  uint32_t *tile_lat = nullptr, *tile_first = nullptr;
  uint32_t ntile = 0;
  int tiles_block = 0;  // block size the tile table was built for
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t nfluid = 0;
  int fold = 0;  // some lattice has folded or kept faces: the FOLD kernel
};

namespace {

void drop_graph(lbm_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int build_tiles(lbm_batch *b) {
  if (b->tiles_block == b->block && b->tile_lat) return LBM_OK;
  std::vector<uint32_t> lat, first;
  for (size_t l = 0; l < b->ctx.size(); l++) {
    const uint32_t n = b->ctx[l]->g.ncell;
    for (uint32_t f = 0; f < n; f += (uint32_t)b->block) {
      lat.push_back((uint32_t)l);
      first.push_back(f);
    }
  }
  if (b->tile_lat) (void)hipFree(b->tile_lat);
  if (b->tile_first) (void)hipFree(b->tile_first);
  b->tile_lat = b->tile_first = nullptr;
  HIPCHK(hipMalloc((void **)&b->tile_lat, 4 * lat.size()));
  HIPCHK(hipMalloc((void **)&b->tile_first, 4 * first.size()));
  HIPCHK(hipMemcpy(b->tile_lat, lat.data(), 4 * lat.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->tile_first, first.data(), 4 * first.size(), hipMemcpyHostToDevice));
  b->ntile = (uint32_t)lat.size();
  b->tiles_block = b->block;
  drop_graph(b);
  return LBM_OK;
}

// The kernel body of the batch: with folded or kept faces in some lattice,
// the FOLD body, which also serves lattices without them provided their
// offsets fit its 32-bit lane offsets.
int batch_layout(lbm_batch *b) {
  int fold = 0;
  for (lbm_ctx *c : b->ctx) fold |= c->g.fold != 0;
  if (fold)
    for (lbm_ctx *c : b->ctx)
      if (c->g.fold == 0 && !fold_fits(c)) return LBM_ERR_UNSUPPORTED;
  if (fold != b->fold) {
    b->fold = fold;
    drop_graph(b);
  }
  return LBM_OK;
}

BatchLaunch launch_of(const lbm_batch *b, int mode, int parity) {
  BatchLaunch B{};
  B.storage = b->storage;
  B.force_mode = b->force_mode;
  B.mode = mode;
  B.la = b->d_args;
  B.tile_lat = b->tile_lat;
  B.tile_first = b->tile_first;
  B.ntile = b->ntile;
  if (b->per_cu > 0) {
    const uint32_t g = (uint32_t)(b->per_cu * b->ncu);
    B.per = (b->ntile + g - 1) / g;
  } else {
    B.per = 1;
  }
  B.grid = (b->ntile + B.per - 1) / B.per;
  B.block = (unsigned)b->block;
  B.parity = parity;
  B.fold = b->fold;
  B.stream = b->stream;
  return B;
}

// kernel arguments of every lattice: parity 0 reads the context's current
// buffer, parity 1 the other one
int upload_args(lbm_batch *b) {
  if (int e = batch_layout(b)) return e;
  const size_t n = b->ctx.size();
  b->h_args.assign(2 * n * b->args_size, 0);
  for (size_t l = 0; l < n; l++) {
    lbm_ctx *c = b->ctx[l];
    LaunchCtx L = launch_ctx(c);
    L.block = b->block;
    batch_fill_args(L, c->cur, c->oth, &b->h_args[(2 * l + 0) * b->args_size]);
    batch_fill_args(L, c->oth, c->cur, &b->h_args[(2 * l + 1) * b->args_size]);
  }
  HIPCHK(hipMemcpyAsync(b->d_args, b->h_args.data(), b->h_args.size(), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return LBM_OK;
}

int ensure_graph(lbm_batch *b) {
  if (b->exec) return LBM_OK;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e0 = launch_stream_batch(launch_of(b, STREAM_COLLIDE, 0));
  hipError_t e1 = launch_stream_batch(launch_of(b, STREAM_COLLIDE, 1));
  hipGraph_t g = nullptr;
  hipError_t ec = hipStreamEndCapture(b->stream, &g);
  if (e0 != hipSuccess || e1 != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return LBM_ERR_HIP;
  }
  b->graph = g;
  HIPCHK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return LBM_OK;
}

// Every lattice: the steps whose MacroProp changes its wall tables run on the
// context itself (lbm_step); then all lattices are in ST_F and in lockstep.
int prologue(lbm_batch *b, int32_t &nsteps) {
  int k = 0;
  for (lbm_ctx *c : b->ctx) {
    HIPCHK(hipStreamSynchronize(c->stream));
    k = std::max(k, c->walls_from_setup + c->backeq_solid_pending);
  }
  k = std::min<int>(k, nsteps);
  if (k > 0) {
    for (lbm_ctx *c : b->ctx) {
      int e = lbm_step(c, k, 1);
      if (e != LBM_OK) return e;
      HIPCHK(hipStreamSynchronize(c->stream));
    }
    nsteps -= k;
  }
  return LBM_OK;
}

// nsteps >= 1 batched steps from ST_F: collide each lattice, nsteps - 1
// fused launches (graph pairs), one stream+macro launch.  ms (if not null)
// receives the event-timed milliseconds of the fused launches.
int batched(lbm_batch *b, int32_t nsteps, double *ms) {
  int e = build_tiles(b);
  if (e != LBM_OK) return e;
  for (lbm_ctx *c : b->ctx) {
    if (c->state != ST_F) return LBM_ERR_STATE;
    LaunchCtx L = launch_ctx(c);
    L.stream = b->stream;
    HIPCHK(launch_collide(L, c->cur));
  }
  if ((e = upload_args(b)) != LBM_OK) return e;
  if ((e = ensure_graph(b)) != LBM_OK) return e;
  const int K = nsteps - 1;
  EventPair ev;
  if (ms) {
    HIPCHK(ev.create());
    HIPCHK(hipEventRecord(ev.t0, b->stream));
  }
  for (int p = 0; p < K / 2; p++) HIPCHK(hipGraphLaunch(b->exec, b->stream));
  if (K % 2) HIPCHK(launch_stream_batch(launch_of(b, STREAM_COLLIDE, 0)));
  if (ms) HIPCHK(hipEventRecord(ev.t1, b->stream));
  HIPCHK(launch_stream_batch(launch_of(b, STREAM_MACRO, K % 2)));
  // K + 1 launches: an odd count leaves every lattice in its other buffer
  for (lbm_ctx *c : b->ctx) {
    if ((K + 1) % 2) std::swap(c->cur, c->oth);
    c->state = ST_F;
    c->fstar_ready = 1;  // the last launch streamed from oth
  }
  HIPCHK(hipEventRecord(b->done, b->stream));
  for (lbm_ctx *c : b->ctx) HIPCHK(hipStreamWaitEvent(c->stream, b->done, 0));
  if (ms) return ev.elapsed(ms);
  return LBM_OK;
}

// sizes already tuned in this process: (lattice cells, storage) -> (threads
// per block, cells per thread)
std::map<uint64_t, std::pair<int, int>> g_tuned;

// Dry runs of the launch-shape tuners.  The timed launches are the fused
// stream+collide kernel of each lattice reading its current buffer, as a
// real step does, but writing a scratch buffer of the same layout, with the
// status words in scratch too: the caller's populations, macroscopic
// arrays, status, buffer order and step state stay as they were.  (The
// current buffer may hold f rather than f*; the kernel does the same work
// either way.)
struct DryRun {
  std::vector<lbm_ctx *> ctx;
  std::vector<void *> out;
  int *status = nullptr;
  unsigned long long *first_bad = nullptr;
  hipStream_t stream = nullptr;
  ~DryRun() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : out) (void)hipFree(p);
    if (status) (void)hipFree(status);
    if (first_bad) (void)hipFree(first_bad);
  }
  int add(lbm_ctx *c) {
    if (!status) {
      HIPCHK(hipMalloc((void **)&status, sizeof(int)));
      HIPCHK(hipMalloc((void **)&first_bad, sizeof(unsigned long long)));
    }
    void *p = nullptr;
    if (hipMalloc(&p, c->esize * (size_t)c->g.ps * 19) != hipSuccess) return LBM_ERR_NOMEM;
    ctx.push_back(c);
    out.push_back(p);
    stream = c->stream;
    return LBM_OK;
  }
  LaunchCtx launch(size_t l, int block, int cpt) const {
    LaunchCtx L = launch_ctx(ctx[l]);
    L.block = block;
    L.cpt = cpt;
    L.status = status;
    L.first_bad = first_bad;
    return L;
  }
  // nsteps dry launches of one context's stream kernel, event-timed
  int time_ctx(lbm_ctx *c, int block, int cpt, int nsteps, double *ms) {
    HIPCHK(hipStreamSynchronize(c->stream));
    const LaunchCtx L = launch(0, block, cpt);
    HIPCHK(launch_stream(L, c->cur, out[0], STREAM_COLLIDE));  // warm-up
    EventPair ev;
    HIPCHK(ev.create());
    HIPCHK(hipEventRecord(ev.t0, c->stream));
    for (int s = 0; s < nsteps; s++) HIPCHK(launch_stream(L, c->cur, out[0], STREAM_COLLIDE));
    HIPCHK(hipEventRecord(ev.t1, c->stream));
    return ev.elapsed(ms);
  }
  // nsteps dry batched launches of shape (block, per_cu): a graph of a
  // launch pair replayed, as lbm_batch_step replays its own
  int time_batch(lbm_batch *b, int block, int per_cu, int nsteps, double *ms) {
    for (lbm_ctx *c : ctx) HIPCHK(hipStreamSynchronize(c->stream));
    int e = lbm_batch_set_launch(b, block, per_cu);
    if (e == LBM_OK) e = build_tiles(b);
    if (e != LBM_OK) return e;
    const size_t n = ctx.size();
    std::vector<unsigned char> h(2 * n * b->args_size, 0);
    for (size_t l = 0; l < n; l++) {
      const LaunchCtx L = launch(l, block, 1);
      batch_fill_args(L, ctx[l]->cur, out[l], &h[(2 * l + 0) * b->args_size]);
      batch_fill_args(L, ctx[l]->cur, out[l], &h[(2 * l + 1) * b->args_size]);
    }
    void *d = nullptr;
    HIPCHK(hipMalloc(&d, h.size()));
    struct Free {
      void *p;
      hipGraph_t g = nullptr;
      hipGraphExec_t x = nullptr;
      ~Free() {
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
        (void)hipFree(p);
      }
    } fr{d};
    HIPCHK(hipMemcpyAsync(d, h.data(), h.size(), hipMemcpyHostToDevice, b->stream));
    BatchLaunch B = launch_of(b, STREAM_COLLIDE, 0);
    B.la = d;
    HIPCHK(launch_stream_batch(B));  // warm-up
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e0 = launch_stream_batch(B);
    hipError_t e1 = launch_stream_batch(B);
    hipError_t ec = hipStreamEndCapture(b->stream, &fr.g);
    if (e0 != hipSuccess || e1 != hipSuccess || ec != hipSuccess) return LBM_ERR_HIP;
    HIPCHK(hipGraphInstantiate(&fr.x, fr.g, nullptr, nullptr, 0));
    EventPair ev;
    HIPCHK(ev.create());
    HIPCHK(hipEventRecord(ev.t0, b->stream));
    for (int s = 0; s < nsteps / 2; s++) HIPCHK(hipGraphLaunch(fr.x, b->stream));
    if (nsteps % 2) HIPCHK(launch_stream_batch(B));
    HIPCHK(hipEventRecord(ev.t1, b->stream));
    return ev.elapsed(ms);
  }
};

}  // namespace

extern "C" {

int lbm_batch_create(lbm_ctx *const *ctxs, int32_t n, lbm_batch **out) {
  if (!ctxs || n < 1 || !out) return LBM_ERR_ARG;
  *out = nullptr;
  lbm_ctx *c0 = ctxs[0];
  if (!c0) return LBM_ERR_ARG;
  for (int i = 0; i < n; i++) {
    lbm_ctx *c = ctxs[i];
    if (!c) return LBM_ERR_ARG;
    if (c->state == ST_NONE) return LBM_ERR_STATE;
    if (c->dev != c0->dev || c->storage != c0->storage || c->force_mode != c0->force_mode)
      return LBM_ERR_UNSUPPORTED;
    if (c->thermal || c->ibm || c->force_mode == 3) return LBM_ERR_UNSUPPORTED;
    // inlet/outlet (and other non-fluid, non-solid) cells: their face loops
    // run per lattice between the steps, which one batched launch has not
    if (c->fnslot) return LBM_ERR_UNSUPPORTED;
    for (int j = 0; j < i; j++)
      if (ctxs[j] == c) return LBM_ERR_ARG;
  }
  HIPCHK(hipSetDevice(c0->dev));
  lbm_batch *b = new lbm_batch();
  b->ctx.assign(ctxs, ctxs + n);
  b->dev = c0->dev;
  b->storage = c0->storage;
  b->force_mode = c0->force_mode;
  // records are packed at the struct's own size: the kernel indexes the
  // array as LatArgs<S>[2 * nlat]
  b->args_size = batch_args_size(b->storage);
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->dev) == hipSuccess && ncu > 0)
    b->ncu = ncu;
  for (lbm_ctx *c : b->ctx) b->nfluid += c->nfluid;
  if (int e = batch_layout(b)) {
    lbm_batch_destroy(b);
    return e;
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
      hipMalloc(&b->d_args, 2 * (size_t)n * b->args_size) != hipSuccess) {
    lbm_batch_destroy(b);
    return LBM_ERR_NOMEM;
  }
  *out = b;
  return LBM_OK;
}

int lbm_batch_destroy(lbm_batch *b) {
  if (!b) return LBM_ERR_ARG;
  (void)hipSetDevice(b->dev);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->tile_lat) (void)hipFree(b->tile_lat);
  if (b->tile_first) (void)hipFree(b->tile_first);
  if (b->d_args) (void)hipFree(b->d_args);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return LBM_OK;
}

int lbm_batch_step(lbm_batch *b, int32_t nsteps) {
  if (!b || nsteps < 0) return LBM_ERR_ARG;
  if (nsteps == 0) return LBM_OK;
  HIPCHK(hipSetDevice(b->dev));
  int e = prologue(b, nsteps);
  if (e != LBM_OK || nsteps == 0) return e;
  return batched(b, nsteps, nullptr);
}

int lbm_batch_time_steps(lbm_batch *b, int32_t nsteps, double *ms) {
  if (!b || nsteps < 1 || !ms) return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(b->dev));
  int32_t pre = 2;  // bring every lattice past its wall-table changes first
  int e = prologue(b, pre);
  if (e != LBM_OK) return e;
  e = batched(b, nsteps + 1, ms);
  if (e != LBM_OK) return e;
  HIPCHK(hipStreamSynchronize(b->stream));
  return LBM_OK;
}

int lbm_batch_set_launch(lbm_batch *b, int32_t block, int32_t per_cu) {
  if (!b || block < 64 || block > 256 || block % 64 != 0 || per_cu < 0 || per_cu > 64) return LBM_ERR_ARG;
  if (b->block != block || b->per_cu != per_cu) drop_graph(b);
  b->block = block;
  b->per_cu = per_cu;
  return LBM_OK;
}

int lbm_batch_get_launch(const lbm_batch *b, int32_t *block, int32_t *per_cu) {
  if (!b) return LBM_ERR_ARG;
  if (block) *block = b->block;
  if (per_cu) *per_cu = b->per_cu;
  return LBM_OK;
}

int64_t lbm_batch_nfluid(const lbm_batch *b) { return b ? b->nfluid : -1; }

int lbm_batch_autotune(lbm_batch *b, int32_t nsteps, int32_t *block, int32_t *per_cu, double *best_ms) {
  if (!b || nsteps < 2) return LBM_ERR_ARG;
  for (lbm_ctx *c : b->ctx)
    if (c->state == ST_NONE) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(b->dev));
  const int blocks[] = {64, 128, 256};
  const int pers[] = {0, 1, 2, 4, 8};
  const int b0 = b->block, p0 = b->per_cu;
  DryRun dry;
  for (lbm_ctx *c : b->ctx)
    if (int e = dry.add(c)) return e;
  double best = 1e30;
  int bb = b0, bp = p0;
  for (int blk : blocks)
    for (int p : pers) {
      double ms = 0;
      int e = dry.time_batch(b, blk, p, nsteps, &ms);
      if (e != LBM_OK) {
        b->block = b0;
        b->per_cu = p0;
        return e;
      }
      if (ms < best) {
        best = ms;
        bb = blk;
        bp = p;
      }
    }
  // the tile table and graph of the last timed shape are rebuilt for the pick
  b->block = b0;
  b->per_cu = p0;
  int e = lbm_batch_set_launch(b, bb, bp);
  if (e != LBM_OK) return e;
  if (block) *block = bb;
  if (per_cu) *per_cu = bp;
  if (best_ms) *best_ms = best;
  return LBM_OK;
}

int lbm_autotune(lbm_ctx *c, int32_t nsteps, int32_t *block, int32_t *cells_per_thread) {
  if (!c || nsteps < 1) return LBM_ERR_ARG;
  if (c->state == ST_NONE || c->ibm) return LBM_ERR_STATE;
  if (c->thermal) return LBM_ERR_UNSUPPORTED;
  HIPCHK(hipSetDevice(c->dev));
  const uint64_t key = (uint64_t)c->g.ncell * 4 + (uint64_t)c->storage;
  auto it = g_tuned.find(key);
  if (it == g_tuned.end()) {
    DryRun dry;
    if (int e = dry.add(c)) return e;
    double best = 1e30;
    std::pair<int, int> pick{c->block, c->cpt};
    for (int blk : {64, 128, 256})
      for (int cpt : {1, 2, 4}) {
        double ms = 0;
        if (int e = dry.time_ctx(c, blk, cpt, nsteps, &ms)) return e;
        if (ms < best) {
          best = ms;
          pick = {blk, cpt};
        }
      }
    it = g_tuned.emplace(key, pick).first;
  }
  c->block = it->second.first;
  c->cpt = it->second.second;
  if (block) *block = c->block;
  if (cells_per_thread) *cells_per_thread = c->cpt;
  return LBM_OK;
}

}  // extern "C"
