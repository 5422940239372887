// lbm_capi.hip -- the C-ABI of include/lbm_hip.h.
//
// A context owns one device lattice: two population buffers (A/B), the cell
// types, the kernel mask, the bounce-back link words, the wall velocities,
// the macroscopic arrays and the force.  Host arrays at the boundary are in
// the reference's [i][j][k] order (k fastest); device arrays are x fastest.
//
// State machine (mirrors where the reference's data is between its calls):
//   ST_F     the current buffer holds f (after Init, Streaming or a step)
//   ST_FSTAR the current buffer holds f* (after Collision)
// LBM::Collision needs ST_F, LBM::Streaming needs ST_FSTAR.  lbm_step from
// ST_F runs one collision, then nsteps-1 fused stream+macro+collide launches,
// then one stream+macro launch, and ends in ST_F with rho,u of the last step
// in the macro arrays -- the same data the reference holds after MacroProp.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_ctx.h"
#include "lbm_device.h"
#include "lbm_launch.h"

using namespace lbm;

LaunchCtx launch_ctx(const lbm_ctx *c) {
  LaunchCtx L;
  memset(&L, 0, sizeof L);
  L.g = c->g;
  L.storage = c->storage;
  L.force_mode = c->force_mode;
  L.block = c->block;
  L.cpt = c->cpt;
  L.omega = 1.0 / c->d.tau;  // omega = dt/tau, lbm.cpp:76
  L.Fu[0] = c->Fu[0];
  L.Fu[1] = c->Fu[1];
  L.Fu[2] = c->Fu[2];
  L.Fx = c->F[0];
  L.Fy = c->F[1];
  L.Fz = c->F[2];
  L.links = c->links;
  L.uwx = c->uw[0];
  L.uwy = c->uw[1];
  L.uwz = c->uw[2];
  L.rho = c->mac[0];
  L.ux = c->mac[1];
  L.uy = c->mac[2];
  L.uz = c->mac[3];
  L.mask = c->mask;
  L.macro_sel = 1 | 16;  // TYPE_F and TYPE_S cells
  L.status = c->status;
  L.first_bad = c->first_bad;
  L.stream = c->stream;
  L.box = c->box;
  // bounce-back through the solids' slots (the ENERGY kernels patch their
  // wall links themselves)
  L.wall_slots = c->thermal == 1 ? 0 : 1;
  L.slot_links = c->links != nullptr && (c->g.fold == 0 || c->slot_links);
  return L;
}

// Faces of the padded lattice whose cells are all TYPE_S (Geo::fold bits),
// and whether some fluid cell still borders a solid off those faces.
static uint32_t solid_faces(const lbm_ctx *c) {
  const Geo &g = c->g;
  const auto solid = [&](uint32_t x, uint32_t y, uint32_t z) {
    return c->h_type[((size_t)z * g.ny + y) * g.px + x] == LBM_TYPE_S;
  };
  uint32_t f = 63;
  for (uint32_t z = 0; z < g.nz; z++)
    for (uint32_t y = 0; y < g.ny; y++) {
      if (!solid(0, y, z)) f &= ~1u;
      if (!solid(g.nx - 1, y, z)) f &= ~2u;
    }
  for (uint32_t z = 0; z < g.nz; z++)
    for (uint32_t x = 0; x < g.nx; x++) {
      if (!solid(x, 0, z)) f &= ~4u;
      if (!solid(x, g.ny - 1, z)) f &= ~8u;
    }
  for (uint32_t y = 0; y < g.ny; y++)
    for (uint32_t x = 0; x < g.nx; x++) {
      if (!solid(x, y, 0)) f &= ~16u;
      if (!solid(x, y, g.nz - 1)) f &= ~32u;
    }
  return f;
}

// Keep faces (Geo::fold bits 8-13): faces, not folded, whose cells are all
// neither fluid nor solid, except solids of a folded face.
static uint32_t keep_faces(const lbm_ctx *c) {
  const Geo &g = c->g;
  const uint32_t f = g.fold;
  const auto ok = [&](uint32_t x, uint32_t y, uint32_t z) {
    const int16_t t = c->h_type[((size_t)z * g.ny + y) * g.px + x];
    if (t != LBM_TYPE_F && t != LBM_TYPE_S) return true;
    if (t == LBM_TYPE_F) return false;
    return ((f & 1u) && x == 0) || ((f & 2u) && x == g.nx - 1) || ((f & 4u) && y == 0) ||
           ((f & 8u) && y == g.ny - 1) || ((f & 16u) && z == 0) || ((f & 32u) && z == g.nz - 1);
  };
  uint32_t k = 63u & ~f;
  for (uint32_t z = 0; z < g.nz; z++)
    for (uint32_t y = 0; y < g.ny; y++) {
      if (!ok(0, y, z)) k &= ~1u;
      if (!ok(g.nx - 1, y, z)) k &= ~2u;
    }
  for (uint32_t z = 0; z < g.nz; z++)
    for (uint32_t x = 0; x < g.nx; x++) {
      if (!ok(x, 0, z)) k &= ~4u;
      if (!ok(x, g.ny - 1, z)) k &= ~8u;
    }
  for (uint32_t y = 0; y < g.ny; y++)
    for (uint32_t x = 0; x < g.nx; x++) {
      if (!ok(x, y, 0)) k &= ~16u;
      if (!ok(x, y, g.nz - 1)) k &= ~32u;
    }
  return k << 8;
}

static bool unfolded_wall_links(const lbm_ctx *c) {
  const Geo &g = c->g;
  const uint32_t f = g.fold;
  for (uint32_t z = 0; z < g.nz; z++)
    for (uint32_t y = 0; y < g.ny; y++)
      for (uint32_t x = 0; x < g.nx; x++) {
        if (c->h_type[((size_t)z * g.ny + y) * g.px + x] != LBM_TYPE_S) continue;
        if (((f & 1u) && x == 0) || ((f & 2u) && x == g.nx - 1) || ((f & 4u) && y == 0) ||
            ((f & 8u) && y == g.ny - 1) || ((f & 16u) && z == 0) || ((f & 32u) && z == g.nz - 1))
          continue;
        for (int m = 1; m < Q; m++) {  // the pushing target x + c_m, wrapped
          const uint32_t xs = (x + g.nx + CX[m]) % g.nx, ys = (y + g.ny + CY[m]) % g.ny,
                         zs = (z + g.nz + CZ[m]) % g.nz;
          if (c->h_type[((size_t)zs * g.ny + ys) * g.px + xs] == LBM_TYPE_F) return true;
        }
      }
  return false;
}

// The shell fold is used when the population offsets it forms fit 32 bits:
// lane offset (idx + ps + plane + px + 1) elements plus a fold or wrap
// correction of at most ps + plane + px + 1 or nx + plane + nz plane.
bool fold_fits(const lbm_ctx *c) {
  const Geo &g = c->g;
  const uint64_t bias = (uint64_t)g.ps + g.plane + g.px + 1;
  const uint64_t wrap = (uint64_t)g.nx + g.plane + (uint64_t)g.nz * g.plane;
  return ((uint64_t)g.ncell + bias + std::max(bias, wrap)) * c->esize < (1ull << 32);
}

// reference order r = (i*ny + j)*nz + k  <->  device order d = (k*ny + j)*px + i
template <class F>
static void for_cells(const lbm_ctx *c, F &&f) {
  const uint32_t nx = c->g.nx, ny = c->g.ny, nz = c->g.nz, px = c->g.px;
  for (uint32_t i = 0; i < nx; i++)
    for (uint32_t j = 0; j < ny; j++)
      for (uint32_t k = 0; k < nz; k++) {
        const size_t r = ((size_t)i * ny + j) * nz + k;
        const size_t d = ((size_t)k * ny + j) * px + i;
        f(r, d);
      }
}

// device index -> reference index (d must be a real cell, x < nx)
static int64_t ref_index(const lbm_ctx *c, uint64_t d) {
  const uint64_t i = d % c->g.px, t = d / c->g.px, j = t % c->g.ny, k = t / c->g.ny;
  return (int64_t)((i * c->g.ny + j) * c->g.nz + k);
}

static int upload_real(lbm_ctx *c, void *dst, const double *src, double dflt) {
  const size_t n = c->g.ncell;
  if (c->csize == 8) {
    std::vector<double> h(n, dflt);
    if (src) for_cells(c, [&](size_t r, size_t d) { h[d] = src[r]; });
    HIPCHK(hipMemcpy(dst, h.data(), n * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> h(n, (float)dflt);
    if (src) for_cells(c, [&](size_t r, size_t d) { h[d] = (float)src[r]; });
    HIPCHK(hipMemcpy(dst, h.data(), n * 4, hipMemcpyHostToDevice));
  }
  return LBM_OK;
}

static int download_real(lbm_ctx *c, double *dst, const void *src) {
  const size_t n = c->g.ncell;
  if (c->csize == 8) {
    std::vector<double> h(n);
    HIPCHK(hipMemcpy(h.data(), src, n * 8, hipMemcpyDeviceToHost));
    for_cells(c, [&](size_t r, size_t d) { dst[r] = h[d]; });
  } else {
    std::vector<float> h(n);
    HIPCHK(hipMemcpy(h.data(), src, n * 4, hipMemcpyDeviceToHost));
    for_cells(c, [&](size_t r, size_t d) { dst[r] = (double)h[d]; });
  }
  return LBM_OK;
}

static void free_all(lbm_ctx *c) {
  output_free(c);
  faces_free(c);
  ibm_free(c);
  thermal_free(c);
  void *ptrs[] = {c->cur, c->oth, c->type, c->mask, c->links, c->uw[0], c->uw[1], c->uw[2],
                  c->mac[0], c->mac[1], c->mac[2], c->mac[3], c->F[0], c->F[1], c->F[2],
                  c->status, c->first_bad, c->red_partial, c->red_out, c->stats};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

int after_stream(lbm_ctx *c, bool forced_inlets) {
  if (!c->backeq_solid_pending && !(forced_inlets && c->backeq_dynamic)) return LBM_OK;
  // the inlets' rho, u are still those of the last MacroProp, which the
  // reference's k = 0 loop reads (lbm.cpp:599-611)
  HIPCHK(launch_backeq_solid(launch_ctx(c), c->cur, c->oth));
  if (int e = thermal_backeq_solid(c)) return e;  // ENERGY: their g too
  c->backeq_solid_pending = 0;
  // the written solids' new moments enter the wall velocities from the next
  // MacroProp on, also when an earlier MacroProp already replaced the setup
  // velocities
  if (c->walls_from_setup < 1) c->walls_from_setup = 1;
  return LBM_OK;
}

extern "C" {

const char *lbm_error_string(int code) {
  switch (code) {
    case LBM_OK: return "ok";
    case LBM_ERR_FEQ_NEGATIVE: return "negative equilibrium in collision (lbm.cpp:114-117)";
    case LBM_ERR_DENSITY_RANGE: return "density outside [0,100] (lbm.cpp:772-775)";
    case LBM_ERR_TEMPERATURE_RANGE: return "temperature outside [0,100] (lbm.cpp:794-797)";
    case LBM_ERR_ARG: return "invalid argument";
    case LBM_ERR_STATE: return "call out of order for the lattice state";
    case LBM_ERR_HIP: return "HIP runtime error";
    case LBM_ERR_NOMEM: return "device allocation failed";
    case LBM_ERR_UNSUPPORTED: return "unsupported configuration";
  }
  return "unknown error";
}

int lbm_create(const lbm_desc *desc, lbm_ctx **out) {
  if (!desc || !out) return LBM_ERR_ARG;
  *out = nullptr;
  if (desc->nx < 1 || desc->ny < 1 || desc->nz < 1) return LBM_ERR_ARG;
  if (desc->lattice != 19) return LBM_ERR_UNSUPPORTED;
  if (desc->storage < LBM_STORE_F64 || desc->storage > LBM_STORE_F16C) return LBM_ERR_ARG;
  if (!(desc->tau > 0.5) || !std::isfinite(desc->tau)) return LBM_ERR_ARG;
  if (desc->bounce_back != LBM_BB_AUTO && desc->bounce_back != LBM_BB_SLOTS) return LBM_ERR_ARG;
  const uint64_t nx = desc->nx + 2ull, ny = desc->ny + 2ull, nz = desc->nz + 2ull;
  const uint64_t px = (nx + 3) / 4 * 4;  // row pitch: 4-cell chunks start aligned
  const uint64_t ncell = px * ny * nz;
  // plane stride: at least ncell + plane + px + 1, so that the speculative
  // bulk loads of the stream kernels (m*ps + idx - off_m for any idx <
  // ncell, |off_m| <= plane + px + 1) stay inside the 19-plane allocation
  const uint64_t ps = (ncell + px * ny + px + 1 + 255) / 256 * 256;
  // 32-bit element offsets: 19 planes must stay below 2^32 elements
  if (ps * 19ull >= (1ull << 32)) return LBM_ERR_UNSUPPORTED;

  lbm_ctx *c = new lbm_ctx();
  c->d = *desc;
  c->g.nx = (uint32_t)nx;
  c->g.ny = (uint32_t)ny;
  c->g.nz = (uint32_t)nz;
  c->g.px = (uint32_t)px;
  c->g.plane = (uint32_t)(px * ny);
  c->g.ncell = (uint32_t)ncell;
  c->g.ps = (uint32_t)ps;
  if (!set_divisors(c->g)) {
    delete c;
    return LBM_ERR_ARG;
  }
  c->storage = desc->storage;
  c->esize = desc->storage == LBM_STORE_F64 ? 8 : (desc->storage == LBM_STORE_F32 ? 4 : 2);
  c->csize = desc->storage == LBM_STORE_F64 ? 8 : 4;
  c->dev = desc->device;

  auto fail = [&](int code) {
    free_all(c);
    delete c;
    return code;
  };
  if (hipSetDevice(c->dev) != hipSuccess) return fail(LBM_ERR_HIP);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(LBM_ERR_HIP);
  const size_t popbytes = c->esize * ps * 19;
  if (hipMalloc(&c->cur, popbytes) != hipSuccess) return fail(LBM_ERR_NOMEM);
  if (hipMalloc(&c->oth, popbytes) != hipSuccess) return fail(LBM_ERR_NOMEM);
  if (hipMalloc((void **)&c->type, ncell * 2) != hipSuccess) return fail(LBM_ERR_NOMEM);
  if (hipMalloc((void **)&c->mask, ncell) != hipSuccess) return fail(LBM_ERR_NOMEM);
  for (int q = 0; q < 4; q++)
    if (hipMalloc(&c->mac[q], ncell * c->csize) != hipSuccess) return fail(LBM_ERR_NOMEM);
  if (hipMalloc((void **)&c->status, sizeof(int)) != hipSuccess) return fail(LBM_ERR_NOMEM);
  if (hipMalloc((void **)&c->first_bad, sizeof(unsigned long long)) != hipSuccess)
    return fail(LBM_ERR_NOMEM);
  const int nb = reduce_blocks((uint32_t)ncell);
  if (hipMalloc((void **)&c->red_partial, sizeof(double) * 3 * nb) != hipSuccess) return fail(LBM_ERR_NOMEM);
  if (hipMalloc((void **)&c->red_out, sizeof(double) * 4) != hipSuccess) return fail(LBM_ERR_NOMEM);
  int zero = 0;
  unsigned long long big = ~0ull;
  if (hipMemcpy(c->status, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) return fail(LBM_ERR_HIP);
  if (hipMemcpy(c->first_bad, &big, sizeof big, hipMemcpyHostToDevice) != hipSuccess) return fail(LBM_ERR_HIP);
  // defaults of LATTICE (LBM.h:560-566): TYPE_F, rho 1, u 0, F 0
  if (lbm_upload_cells(c, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) != LBM_OK)
    return fail(LBM_ERR_HIP);
  *out = c;
  return LBM_OK;
}

int lbm_destroy(lbm_ctx *c) {
  if (!c) return LBM_ERR_ARG;
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return LBM_OK;
}

int lbm_upload_cells(lbm_ctx *c, const int16_t *type, const double *rho, const double *ux,
                     const double *uy, const double *uz, const double *Fx, const double *Fy,
                     const double *Fz) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  // new cells: a previous IBM set-up no longer applies (lbm_init rebuilds
  // the mask it marked)
  ibm_free(c);
  c->box = BoxArgs{};
  thermal_free(c);  // lbm_upload_thermal follows when the case is thermal
  const size_t n = c->g.ncell;
  const size_t nref = (size_t)c->g.nx * c->g.ny * c->g.nz;
  c->h_type.assign(n, (int16_t)-1);  // row padding
  if (type)
    for (size_t r = 0; r < nref; r++)
      if (type[r] < 0 || type[r] > 7) return LBM_ERR_ARG;
  for_cells(c, [&](size_t r, size_t d) { c->h_type[d] = type ? type[r] : (int16_t)LBM_TYPE_F; });
  c->nfluid = 0;
  for (size_t d = 0; d < n; d++) c->nfluid += c->h_type[d] == LBM_TYPE_F;
  HIPCHK(hipMemcpy(c->type, c->h_type.data(), n * 2, hipMemcpyHostToDevice));
  int e;
  if ((e = upload_real(c, c->mac[0], rho, 1.0)) != LBM_OK) return e;
  if ((e = upload_real(c, c->mac[1], ux, 0.0)) != LBM_OK) return e;
  if ((e = upload_real(c, c->mac[2], uy, 0.0)) != LBM_OK) return e;
  if ((e = upload_real(c, c->mac[3], uz, 0.0)) != LBM_OK) return e;

  // force: none, uniform (same F in every cell) or a per-cell field
  const double *Fs[3] = {Fx, Fy, Fz};
  bool any = false, uniform = true;
  double F0[3] = {0, 0, 0};
  for (int q = 0; q < 3; q++) {
    if (!Fs[q]) continue;
    F0[q] = Fs[q][0];
    for (size_t r = 0; r < nref; r++) {
      if (Fs[q][r] != 0.0) any = true;
      if (Fs[q][r] != F0[q]) uniform = false;
    }
  }
  for (int q = 0; q < 3; q++)
    if (c->F[q]) {
      (void)hipFree(c->F[q]);
      c->F[q] = nullptr;
    }
  if (!any) {
    c->force_mode = 0;
    c->Fu[0] = c->Fu[1] = c->Fu[2] = 0;
  } else if (uniform) {
    c->force_mode = 2;
    for (int q = 0; q < 3; q++) c->Fu[q] = F0[q];
  } else {
    c->force_mode = 1;
    for (int q = 0; q < 3; q++) {
      if (hipMalloc(&c->F[q], n * c->csize) != hipSuccess) return LBM_ERR_NOMEM;
      if ((e = upload_real(c, c->F[q], Fs[q], 0.0)) != LBM_OK) return e;
    }
  }
  // a TYPE_IN cell at k = 0 under a TYPE_S cell at k = Nz+1 that carries a
  // force: the back-face write into that solid changes every step
  c->inlet_forced_over_solid = 0;
  if (c->force_mode != 0) {
    const uint32_t plane = c->g.plane, top = (c->g.nz - 1) * plane;
    for (uint32_t d = 0; d < plane && !c->inlet_forced_over_solid; d++) {
      if (c->h_type[d] != LBM_TYPE_IN || c->h_type[top + d] != LBM_TYPE_S) continue;
      const uint32_t i = d % c->g.px, j = d / c->g.px;
      const size_t r = ((size_t)i * c->g.ny + j) * c->g.nz;  // (i, j, 0)
      for (int q = 0; q < 3; q++)
        if (c->force_mode == 2 ? c->Fu[q] != 0.0 : Fs[q] && Fs[q][r] != 0.0) c->inlet_forced_over_solid = 1;
    }
  }
  c->uploaded = 1;
  c->state = ST_NONE;
  return LBM_OK;
}

int lbm_init(lbm_ctx *c) {
  if (!c || !c->uploaded) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = c->g.ncell;
  // wall tables: allocated when some fluid cell borders a solid
  bool any_solid = false;
  for (size_t d = 0; d < n; d++)
    if (c->h_type[d] == LBM_TYPE_S) {
      any_solid = true;
      break;
    }
  if (any_solid && !c->links) {
    HIPCHK(hipMalloc((void **)&c->links, n * 4));
    for (int q = 0; q < 3; q++) HIPCHK(hipMalloc(&c->uw[q], n * c->csize));
  }
  // wall links into all-solid shell faces resolved in the pull address (the
  // ENERGY and D3Q7 kernels keep their own wall handling)
  c->g.fold = 0;
  c->slot_links = any_solid;
  if (c->thermal == 0 && c->d.bounce_back == LBM_BB_AUTO && fold_fits(c)) {
    if (any_solid) {
      c->g.fold = solid_faces(c);
      c->slot_links = c->g.fold == 0 || unfolded_wall_links(c);
    }
    c->g.fold |= keep_faces(c);
  }
  LaunchCtx L = launch_ctx(c);
  // 1. kernel mask, bounce-back links and wall velocities from the setup
  //    velocities of the solids (what the first Streaming reads)
  HIPCHK(launch_tables(L, c->mask, c->links, c->uw[0], c->uw[1], c->uw[2], c->type));
  // 2. f = f* = feq (LBM::Init); with ENERGY also g = g* = geq and the
  //    thermal wall tables
  if (c->thermal == 1) {
    if (int e = thermal_init(c)) return e;
    // the inlet/outlet cells' own f, g (the face loops move both)
    if (int e = faces_setup(c)) return e;
  } else {
    HIPCHK(launch_init(L, c->cur, c->oth));
    if (c->thermal == 2) {
      if (int e = q7_init(c)) return e;
    }
    // the inlet/outlet cells' own populations (before the solids' macro
    // entries below change)
    if (int e = faces_setup(c)) return e;
    if (any_solid && L.wall_slots) {
      // the solids' slots will carry bounce-back values: their MacroProp
      // values (moments of their Init populations, which never change) are
      // taken now, for the wall tables after the first MacroProp
      LaunchCtx Ls = L;
      Ls.macro_sel = 16;  // TYPE_S only
      HIPCHK(launch_macro(Ls, c->cur));
    }
  }
  if (int e = ibm_remark(c)) return e;
  HIPCHK(hipStreamSynchronize(c->stream));
  // MacroProps after which the wall tables change: the first (solids' rho,
  // u, T become moments of their populations); with the Boussinesq force
  // also the second (the solids' F_y set by the first enters their u)
  c->walls_from_setup = any_solid ? (c->thermal == 1 && c->td.boussinesq ? 2 : 1) : 0;
  c->backeq_solid_pending = 0;
  {
    const uint32_t plane = c->g.plane, top = (c->g.nz - 1) * plane;
    for (uint32_t d = 0; d < plane; d++)
      if (c->h_type[d] == LBM_TYPE_IN && c->h_type[top + d] == LBM_TYPE_S) c->backeq_solid_pending = 1;
  }
  // the write changes every step with a force on the inlet (its u), and
  // under ENERGY (its T = sum g of the last write, i.e. rho T of the one
  // before)
  c->backeq_dynamic = c->backeq_solid_pending && (c->inlet_forced_over_solid || c->thermal == 1);
  c->state = ST_F;
  c->fstar_ready = 0;
  return LBM_OK;
}


// After the first MacroProp: the solids' rho,u become the moments of their
// (never changing) Init populations, and the wall velocities of every later
// Streaming are averaged from those (lbm.cpp:941-962).
// solids_done: this MacroProp already covered the solids (lbm_macro); the
// fused step's MacroProp covers fluid cells only.  (The thermal MacroProp
// with the Boussinesq force is not idempotent: it rewrites the solids' F.)
static int walls_from_moments(lbm_ctx *c, bool solids_done = false) {
  if (!c->walls_from_setup) return LBM_OK;
  LaunchCtx L = launch_ctx(c);
  L.macro_sel = 16;  // TYPE_S only
  if (!solids_done) {
    if (c->thermal == 1) {
      if (int e = thermal_launch(c, 4, 16)) return e;
    } else if (!L.wall_slots) {
      HIPCHK(launch_macro(L, c->cur));
    }
    // with wall slots the solids' values were taken at lbm_init (and by the
    // back-face write, after_stream)
  }
  HIPCHK(launch_tables(L, c->mask, c->links, c->uw[0], c->uw[1], c->uw[2], c->type));
  if (c->thermal == 1) {
    if (int e = thermal_tables(c)) return e;
  } else if (c->thermal == 2) {
    if (int e = q7_tables(c)) return e;
  }
  c->walls_from_setup--;
  return LBM_OK;
}

int lbm_collide(lbm_ctx *c) {
  if (!c) return LBM_ERR_ARG;
  if (c->state != ST_F) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  LaunchCtx L = launch_ctx(c);
  if (ibm_active(c)) {
    // an IBM run collides with the result of the preceding IBLBM
    // (lbm_ibm_apply): force mode 3
    if (!ibm_ready(c)) return LBM_ERR_STATE;
    L.force_mode = 3;
    ibm_consume(c);
  }
  if (c->thermal == 1) {
    if (int e = thermal_launch(c, 3)) return e;
  } else if (c->thermal == 2) {
    if (int e = q7_launch(c, Q7_COLLIDE, L.force_mode)) return e;
  } else {
    HIPCHK(launch_collide(L, c->cur));
  }
  c->state = ST_FSTAR;
  return LBM_OK;
}

int lbm_stream(lbm_ctx *c) {
  if (!c) return LBM_ERR_ARG;
  if (c->state != ST_FSTAR) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  if (c->thermal == 1) {
    if (int e = thermal_face_sources(c)) return e;
    if (int e = thermal_launch(c, 2)) return e;
  } else if (c->thermal == 2) {
    if (int e = q7_launch(c, Q7_STREAM)) return e;
  } else {
    HIPCHK(launch_stream(launch_ctx(c), c->cur, c->oth, STREAM_ONLY));
  }
  std::swap(c->cur, c->oth);
  c->fstar_ready = 1;
  c->state = ST_F;
  if (int e = after_stream(c, true)) return e;
  // the face loops; the other-type cells' MacroProp waits for lbm_macro
  return faces_step(c, false, 0);
}

int lbm_macro(lbm_ctx *c) {
  if (!c) return LBM_ERR_ARG;
  if (c->state != ST_F || ibm_active(c)) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  if (c->thermal == 1) {
    if (int e = thermal_launch(c, 4, 1 | 16)) return e;
  } else if (c->thermal == 2) {
    if (int e = q7_launch(c, Q7_MACRO)) return e;
  } else {
    LaunchCtx L = launch_ctx(c);
    // with wall slots the solids' populations are no longer in the buffer;
    // their (constant) MacroProp values are already stored
    if (L.wall_slots) L.macro_sel = 1;
    HIPCHK(launch_macro(L, c->cur));
  }
  if (int e = faces_macro(c, false)) return e;
  return walls_from_moments(c, true);
}

// one launch of the step sequence, plain or thermal
static int step_launch(lbm_ctx *c, const LaunchCtx &L, int mode) {
  if (c->thermal == 1) {
    const int tm = mode == STREAM_COLLIDE ? 0 : mode == STREAM_COLLIDE_MACRO ? 1 : mode == STREAM_ONLY ? 2 : 5;
    if (int e = thermal_face_sources(c)) return e;
    if (int e = thermal_launch(c, tm)) return e;
  } else if (c->thermal == 2) {
    const int tm = mode == STREAM_COLLIDE ? Q7_FUSED
                   : mode == STREAM_COLLIDE_MACRO ? Q7_FUSED_EMIT
                   : mode == STREAM_ONLY ? Q7_STREAM
                                         : Q7_STREAM_EMIT;
    if (int e = q7_launch(c, tm)) return e;
  } else {
    HIPCHK(launch_stream(L, c->cur, c->oth, mode));
  }
  std::swap(c->cur, c->oth);
  c->fstar_ready = 1;
  // the back-face write into solids, before the face loops' MacroProp
  // replaces the inlets' rho, u
  if (int e = after_stream(c, true)) return e;
  // MacroProp of the other-type cells: the inlets every step, all of them
  // on steps that emit rho, u
  return faces_step(c, false, mode == STREAM_ONLY ? 0 : mode == STREAM_COLLIDE ? 1 : 2);
}

static int collide_launch(lbm_ctx *c, const LaunchCtx &L) {
  if (c->thermal == 1) return thermal_launch(c, 3);
  if (c->thermal == 2) return q7_launch(c, Q7_COLLIDE);
  HIPCHK(launch_collide(L, c->cur));
  return LBM_OK;
}

int lbm_step(lbm_ctx *c, int32_t nsteps, int32_t emit_macro) {
  (void)emit_macro;  // the last launch always emits rho,u (the state after MacroProp)
  if (!c || nsteps < 0) return LBM_ERR_ARG;
  if (c->state == ST_NONE || ibm_active(c)) return LBM_ERR_STATE;
  if (nsteps == 0) return LBM_OK;
  HIPCHK(hipSetDevice(c->dev));
  LaunchCtx L = launch_ctx(c);
  int e;
  if (c->state == ST_F && (e = collide_launch(c, L)) != LBM_OK) return e;
  while (c->walls_from_setup || c->backeq_solid_pending || c->backeq_dynamic) {
    // steps whose MacroProp changes the wall tables run unfused
    if ((e = step_launch(c, L, STREAM_MACRO)) != LBM_OK) return e;
    c->state = ST_F;
    e = walls_from_moments(c);
    if (e != LBM_OK) return e;
    if (--nsteps == 0) return LBM_OK;
    L = launch_ctx(c);
    if ((e = collide_launch(c, L)) != LBM_OK) return e;
  }
  for (int s = 0; s < nsteps - 1; s++)
    if ((e = step_launch(c, L, STREAM_COLLIDE)) != LBM_OK) return e;
  if ((e = step_launch(c, L, STREAM_MACRO)) != LBM_OK) return e;
  c->state = ST_F;
  return LBM_OK;
}

int lbm_time_steps(lbm_ctx *c, int32_t nsteps, double *ms) {
  if (!c || nsteps < 1 || !ms) return LBM_ERR_ARG;
  if (c->state == ST_NONE || ibm_active(c)) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  if (c->backeq_dynamic) {
    // every step runs unfused (its wall tables change): time lbm_step
    EventPair ev;
    HIPCHK(ev.create());
    HIPCHK(hipEventRecord(ev.t0, c->stream));
    if (int e = lbm_step(c, nsteps, 1)) return e;
    HIPCHK(hipEventRecord(ev.t1, c->stream));
    return ev.elapsed(ms);
  }
  while (c->walls_from_setup || c->backeq_solid_pending) {
    int e = lbm_step(c, 1, 1);
    if (e != LBM_OK) return e;
  }
  const LaunchCtx L = launch_ctx(c);
  int er;
  if (c->state == ST_F && (er = collide_launch(c, L)) != LBM_OK) return er;
  EventPair ev;
  HIPCHK(ev.create());
  HIPCHK(hipEventRecord(ev.t0, c->stream));
  for (int s = 0; s < nsteps; s++)
    if ((er = step_launch(c, L, STREAM_COLLIDE)) != LBM_OK) return er;
  HIPCHK(hipEventRecord(ev.t1, c->stream));
  if ((er = step_launch(c, L, STREAM_MACRO)) != LBM_OK) return er;
  c->state = ST_F;
  return ev.elapsed(ms);
}

int lbm_sync(lbm_ctx *c) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  return lbm_status(c, nullptr);
}

int lbm_status(lbm_ctx *c, int64_t *first_bad_cell) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  int st = 0;
  unsigned long long fb = 0;
  HIPCHK(hipMemcpy(&st, c->status, sizeof st, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&fb, c->first_bad, sizeof fb, hipMemcpyDeviceToHost));
  if (first_bad_cell) {
    if (st == 0) {
      *first_bad_cell = -1;
    } else {
      *first_bad_cell = ref_index(c, fb);
    }
  }
  if (st & 2) return LBM_ERR_DENSITY_RANGE;
  if (st & 4) return LBM_ERR_TEMPERATURE_RANGE;
  if (st & 1) return LBM_ERR_FEQ_NEGATIVE;
  return LBM_OK;
}

int lbm_download_macro(lbm_ctx *c, double *rho, double *ux, double *uy, double *uz) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  double *outs[4] = {rho, ux, uy, uz};
  for (int q = 0; q < 4; q++)
    if (outs[q]) {
      int e = download_real(c, outs[q], c->mac[q]);
      if (e != LBM_OK) return e;
    }
  return LBM_OK;
}

int lbm_download_force(lbm_ctx *c, double *Fx, double *Fy, double *Fz) {
  if (!c) return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipStreamSynchronize(c->stream));
  double *outs[3] = {Fx, Fy, Fz};
  const size_t nref = (size_t)c->g.nx * c->g.ny * c->g.nz;
  for (int q = 0; q < 3; q++) {
    if (!outs[q]) continue;
    if (c->force_mode == 1) {
      int e = download_real(c, outs[q], c->F[q]);
      if (e != LBM_OK) return e;
    } else {
      for (size_t r = 0; r < nref; r++) outs[q][r] = c->force_mode == 2 ? c->Fu[q] : 0.0;
    }
  }
  return LBM_OK;
}

int lbm_download_f(lbm_ctx *c, double *f) {
  if (!c || !f) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = c->g.ncell;
  double *d = nullptr;
  HIPCHK(hipMalloc((void **)&d, n * 19 * sizeof(double)));
  hipError_t e = launch_f_out(launch_ctx(c), c->cur, d);
  if (e == hipSuccess) e = faces_f_out(c, d);
  std::vector<double> h;
  if (e == hipSuccess) {
    h.resize(n * 19);
    e = hipMemcpyAsync(h.data(), d, n * 19 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  (void)hipFree(d);
  if (e != hipSuccess) return LBM_ERR_HIP;
  for_cells(c, [&](size_t r, size_t dd) { memcpy(f + r * 19, h.data() + dd * 19, 19 * sizeof(double)); });
  return LBM_OK;
}

int lbm_export_links(lbm_ctx *c, int16_t *type, int32_t *src, uint8_t *kind) {
  if (!c) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  const size_t n = c->g.ncell;
  if (type) for_cells(c, [&](size_t r, size_t d) { type[r] = c->h_type[d]; });
  if (!src && !kind) return LBM_OK;
  int32_t *dsrc = nullptr;
  uint8_t *dkind = nullptr;
  HIPCHK(hipMalloc((void **)&dsrc, n * 19 * 4));
  if (hipMalloc((void **)&dkind, n * 19) != hipSuccess) {
    (void)hipFree(dsrc);
    return LBM_ERR_NOMEM;
  }
  std::vector<int32_t> hs(n * 19);
  std::vector<uint8_t> hk(n * 19);
  hipError_t e = launch_links(launch_ctx(c), dsrc, dkind);
  if (e == hipSuccess) e = hipMemcpyAsync(hs.data(), dsrc, n * 19 * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(hk.data(), dkind, n * 19, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dsrc);
  (void)hipFree(dkind);
  if (e != hipSuccess) return LBM_ERR_HIP;
  for_cells(c, [&](size_t r, size_t d) {
    for (int m = 0; m < 19; m++) {
      int32_t s = hs[d * 19 + m];
      if (s >= 0) s = (int32_t)ref_index(c, (uint64_t)s);
      if (src) src[r * 19 + m] = s;
      if (kind) kind[r * 19 + m] = hk[d * 19 + m];
    }
  });
  return LBM_OK;
}

int lbm_reduce_fp64(lbm_ctx *c, int32_t what, double *out) {
  if (!c || !out || what < 0 || what > 2) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(launch_reduce(launch_ctx(c), what, c->red_partial, c->red_out));
  const int nc = what == 1 ? 3 : 1;
  HIPCHK(hipMemcpyAsync(out, c->red_out, sizeof(double) * nc, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

int lbm_set_launch(lbm_ctx *c, int32_t block, int32_t cells_per_thread) {
  if (!c || block < 64 || block > 256 || block % 64 != 0) return LBM_ERR_ARG;
  if (cells_per_thread != 1 && cells_per_thread != 2 && cells_per_thread != 4) return LBM_ERR_ARG;
  c->block = block;
  c->cpt = cells_per_thread;
  return LBM_OK;
}

int lbm_get_launch(const lbm_ctx *c, int32_t *block, int32_t *cells_per_thread) {
  if (!c) return LBM_ERR_ARG;
  if (block) *block = c->block;
  if (cells_per_thread) *cells_per_thread = c->cpt;
  return LBM_OK;
}

int lbm_bounce_back_faces(const lbm_ctx *c, int32_t *faces, int32_t *slot_links) {
  if (!c) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  if (faces) *faces = (int32_t)c->g.fold;
  if (slot_links) *slot_links = c->slot_links;
  return LBM_OK;
}

int64_t lbm_ncells(const lbm_ctx *c) {
  return c ? (int64_t)c->g.nx * c->g.ny * c->g.nz : -1;
}
int64_t lbm_nfluid(const lbm_ctx *c) { return c ? c->nfluid : -1; }
int32_t lbm_storage(const lbm_ctx *c) { return c ? c->storage : -1; }

double lbm_bytes_per_cell(const lbm_ctx *c) {
  if (!c) return 0.0;
  // A/B pull: 19 populations read + 19 written + the 1-byte mask; with
  // ENERGY also 19 g read + written and the previous Phi, u read + written
  if (c->thermal == 1) return 4.0 * 19.0 * (double)c->esize + 8.0 * (double)c->csize + 1.0;
  if (c->thermal == 2) return 2.0 * 26.0 * (double)c->esize + 1.0;
  return 2.0 * 19.0 * (double)c->esize + 1.0;
}

}  // extern "C"
