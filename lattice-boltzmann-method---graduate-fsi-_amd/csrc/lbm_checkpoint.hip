// lbm_checkpoint.hip -- checkpoint and resume of a lattice's evolving state.
//
// The reference has none (its only "resume" is the interactive prompt that
// extends the step count, main.cpp:397-410).  A checkpoint is the raw device
// state that stepping changes, in one host buffer: both population buffers
// (the current one and the one the last streaming pulled from), the
// macroscopic arrays, the wall tables (kernel mask, link words, wall
// velocities; thermal links and wall temperatures), the sticky status word, the
// per-cell force when it is a field, the inlet/outlet slots, the thermal
// arrays and the running statistics, with the context's state flags.  It is
// restored into a context built by the same set-up (lbm_create,
// lbm_upload_cells, lbm_upload_thermal, lbm_init): stepping it then gives the
// same results, bit for bit, as stepping the saved context.  A context with an
// immersed boundary is covered while its body is rigid and has never moved,
// between IBLBM calls: every IBM step then recomputes the markers' velocities
// and forces from the populations, and the restoring context repeats the
// same lbm_ibm_setup.  Moving bodies and filaments (their marker positions
// and nodes are state of their own) are not.

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "lbm_ctx.h"

bool ibm_checkpointable(const lbm_ctx *c);

namespace {

constexpr char MAGIC[8] = {'L', 'B', 'M', 'C', 'K', 'P', 'T', '1'};

struct CkptHeader {
  char magic[8];
  uint64_t total;  // bytes, header included
  int32_t nx, ny, nz, storage, thermal, planes;
  uint64_t ps, ncell;
  int32_t fnslot, has_stats, has_walls, has_force;
  int32_t state, fstar_ready, walls_from_setup, backeq_solid_pending, backeq_dynamic, walls_refresh;
};

// One device array of the checkpoint: its address and byte count.
struct Block {
  void *p;
  size_t bytes;
};

int planes_of(const lbm_ctx *c) { return c->thermal == 1 ? 38 : c->thermal == 2 ? 26 : 19; }

// The arrays in checkpoint order; `stats` whether the running statistics are
// part of it (allocated in the saved context).
std::vector<Block> blocks_of(const lbm_ctx *c, bool stats) {
  const size_t n = c->g.ncell, cs = c->csize;
  const size_t pop = c->esize * (size_t)c->g.ps * planes_of(c);
  std::vector<Block> b{{c->cur, pop}, {c->oth, pop}};
  for (int q = 0; q < 4; q++) b.push_back({c->mac[q], n * cs});
  b.push_back({c->status, sizeof(int)});
  b.push_back({c->first_bad, sizeof(unsigned long long)});
  // the kernel mask and the link words: the wall tables rewrite them (moving-
  // wall bits) when the solids' moments replace their set-up velocities
  b.push_back({c->mask, n});
  if (c->links) {
    b.push_back({c->links, n * 4});
    for (int q = 0; q < 3; q++) b.push_back({c->uw[q], n * cs});
  }
  if (c->force_mode == 1)
    for (int q = 0; q < 3; q++) b.push_back({c->F[q], n * cs});
  if (c->fnslot) {
    const size_t nrow = c->thermal == 1 ? 38 : 19, nmac = c->thermal == 1 ? 8 : 4;
    b.push_back({c->fof, (size_t)c->fnslot * nrow * sizeof(double)});
    b.push_back({c->fom, (size_t)c->fnslot * nmac * sizeof(double)});
  }
  if (c->thermal) {
    b.push_back({c->T, n * cs});
    for (int q = 0; q < 3; q++) b.push_back({c->qv[q], n * cs});
  }
  if (c->thermal == 1) {
    b.push_back({c->phi, n * cs});
    for (int q = 0; q < 3; q++) b.push_back({c->up[q], n * cs});
    b.push_back({c->rho_w, n * cs});
    b.push_back({c->Tw, n * cs});
    b.push_back({c->tlinks, n * 8});
  }
  if (stats) b.push_back({c->stats, (size_t)LBM_STAT_PLANES * n * sizeof(double)});
  return b;
}

CkptHeader header_of(const lbm_ctx *c) {
  CkptHeader h{};
  std::memcpy(h.magic, MAGIC, 8);
  h.nx = (int32_t)c->g.nx;
  h.ny = (int32_t)c->g.ny;
  h.nz = (int32_t)c->g.nz;
  h.storage = c->storage;
  h.thermal = c->thermal;
  h.planes = planes_of(c);
  h.ps = c->g.ps;
  h.ncell = c->g.ncell;
  h.fnslot = c->fnslot;
  h.has_stats = c->stats != nullptr;
  h.has_walls = c->links != nullptr;
  h.has_force = c->force_mode == 1;
  h.state = c->state;
  h.fstar_ready = c->fstar_ready;
  h.walls_from_setup = c->walls_from_setup;
  h.backeq_solid_pending = c->backeq_solid_pending;
  h.backeq_dynamic = c->backeq_dynamic;
  h.walls_refresh = c->walls_refresh;
  size_t total = sizeof(CkptHeader);
  for (const Block &b : blocks_of(c, h.has_stats)) total += b.bytes;
  h.total = total;
  return h;
}

}  // namespace

extern "C" {

int lbm_checkpoint_size(lbm_ctx *c, uint64_t *bytes) {
  if (!c || !bytes) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  if (!ibm_checkpointable(c)) return LBM_ERR_UNSUPPORTED;
  *bytes = header_of(c).total;
  return LBM_OK;
}

int lbm_checkpoint_save(lbm_ctx *c, void *buf, uint64_t bytes) {
  if (!c || !buf) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  if (!ibm_checkpointable(c)) return LBM_ERR_UNSUPPORTED;
  const CkptHeader h = header_of(c);
  if (bytes < h.total) return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  char *p = static_cast<char *>(buf);
  std::memcpy(p, &h, sizeof h);
  p += sizeof h;
  for (const Block &b : blocks_of(c, h.has_stats)) {
    HIPCHK(hipMemcpy(p, b.p, b.bytes, hipMemcpyDeviceToHost));
    p += b.bytes;
  }
  return LBM_OK;
}

int lbm_checkpoint_load(lbm_ctx *c, const void *buf, uint64_t bytes) {
  if (!c || !buf || bytes < sizeof(CkptHeader)) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;  // the set-up, lbm_init included, comes first
  if (!ibm_checkpointable(c)) return LBM_ERR_UNSUPPORTED;
  CkptHeader h;
  std::memcpy(&h, buf, sizeof h);
  const CkptHeader mine = header_of(c);
  // the same lattice, storage, thermal model and set-up
  if (std::memcmp(h.magic, MAGIC, 8) != 0 || h.total != bytes || h.nx != mine.nx || h.ny != mine.ny ||
      h.nz != mine.nz || h.storage != mine.storage || h.thermal != mine.thermal || h.planes != mine.planes ||
      h.ps != mine.ps || h.ncell != mine.ncell || h.fnslot != mine.fnslot || h.has_walls != mine.has_walls ||
      h.has_force != mine.has_force)
    return LBM_ERR_ARG;
  if (h.state != ST_F && h.state != ST_FSTAR) return LBM_ERR_ARG;
  size_t need = sizeof h;
  for (const Block &b : blocks_of(c, h.has_stats)) need += b.bytes;
  if (need != h.total) return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  // the statistics as the saved context had them: accumulated, or none yet
  if (h.has_stats && !c->stats)
    HIPCHK(hipMalloc((void **)&c->stats, (size_t)LBM_STAT_PLANES * c->g.ncell * sizeof(double)));
  if (!h.has_stats && c->stats) {
    HIPCHK(hipFree(c->stats));
    c->stats = nullptr;
  }
  const char *p = static_cast<const char *>(buf) + sizeof h;
  for (const Block &b : blocks_of(c, h.has_stats)) {
    HIPCHK(hipMemcpy(b.p, p, b.bytes, hipMemcpyHostToDevice));
    p += b.bytes;
  }
  c->state = h.state;
  c->fstar_ready = h.fstar_ready;
  c->walls_from_setup = h.walls_from_setup;
  c->backeq_solid_pending = h.backeq_solid_pending;
  c->backeq_dynamic = h.backeq_dynamic;
  c->walls_refresh = h.walls_refresh;
  return LBM_OK;
}

}  // extern "C"
