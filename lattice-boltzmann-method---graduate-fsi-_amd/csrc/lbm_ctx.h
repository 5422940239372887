// lbm_ctx.h -- the context behind the opaque lbm_ctx* of the C-ABI, shared
// by the C-ABI translation units (lbm_capi.hip, lbm_ibm.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_launch.h"

enum { ST_NONE = 0, ST_F = 1, ST_FSTAR = 2 };

// FP64 planes of the running statistics (lbm_stats.hip)
constexpr int LBM_STAT_PLANES = 26;

struct lbm_ibm;  // immersed-boundary state (lbm_ibm.hip)

struct lbm_ctx {
  lbm_desc d{};
  lbm::Geo g{};
  int storage = 0;
  size_t esize = 8;  // bytes per stored population
  size_t csize = 8;  // bytes per macro/force value
  int dev = 0;
  hipStream_t stream = nullptr;
  void *cur = nullptr, *oth = nullptr;
  int16_t *type = nullptr;
  uint8_t *mask = nullptr;
  uint32_t *links = nullptr;
  void *uw[3] = {nullptr, nullptr, nullptr};
  void *mac[4] = {nullptr, nullptr, nullptr, nullptr};  // rho, ux, uy, uz
  void *F[3] = {nullptr, nullptr, nullptr};
  int force_mode = 0;
  double Fu[3] = {0, 0, 0};
  int *status = nullptr;
  unsigned long long *first_bad = nullptr;
  double *red_partial = nullptr, *red_out = nullptr;
  int state = ST_NONE;
  // the last buffer operation was a streaming that pulled from oth: oth holds
  // the f* that produced cur's f (what the momentum exchange reads)
  int fstar_ready = 0;
  // running statistics of Sup_MacroProp (lbm_stats_update): LBM_STAT_PLANES
  // FP64 planes of ncell values, allocated on the first update
  double *stats = nullptr;
  int uploaded = 0;
  // The reference's first Streaming reads the solids' setup velocities; from
  // the first MacroProp on (lbm.cpp:672-675 runs over solids too) it reads
  // their moments.  Set while the wall tables still hold the setup values.
  int walls_from_setup = 0;
  // TYPE_S cells at k = Nz+1 under a TYPE_IN cell at k = 0 take feq2 of that
  // cell at the first Streaming (the back-face inflow loop, lbm.cpp:599-611)
  int backeq_solid_pending = 0;
  // ... and, when such an inlet carries a force, at every Streaming: its u
  // gains F/(2 rho) at every MacroProp (lbm.cpp:672-675 over its feq2
  // populations), so the solid's moments and the wall velocities next to it
  // change every step
  int backeq_dynamic = 0;
  int inlet_forced_over_solid = 0;  // set by lbm_upload_cells
  // bounce-back of the solid shell faces (g.fold, set by lbm_init): some
  // fluid cell has a wall link that is not folded, so the stream kernel
  // reads link words and writes solids' slots
  int slot_links = 0;
  int64_t nfluid = 0;
  int block = 256;
  int cpt = 1;  // cells per thread of the stream kernel
  std::vector<int16_t> h_type;  // device order
  lbm_ibm *ibm = nullptr;       // set by lbm_ibm_setup
  lbm::BoxArgs box{};           // the IBM box as force mode 3 reads it
  // ---- thermal branch (lbm_thermal.hip); the population buffers then hold
  // 38 planes: f in 0..18, g in 19..37
  int thermal = 0;  // 1: the reference's ENERGY branch, 2: D3Q7 (26 planes)
  lbm_thermal_desc td{};
  double T0 = 0;                     // storage shift of g (FP32)
  void *T = nullptr, *qv[3] = {nullptr, nullptr, nullptr};  // macro T, q
  void *phi = nullptr, *up[3] = {nullptr, nullptr, nullptr};  // previous collision's Phi, u
  void *rho_w = nullptr;             // rho of the last collision (wall cells)
  void *Tw = nullptr;                // wall temperature of wall cells
  uint64_t *tlinks = nullptr;        // bit m: Dirichlet link m; bit 32+m: Neumann
  int walls_refresh = 0;             // MacroProps left that change the wall tables
  std::vector<int16_t> h_typeE;      // device order
  int16_t *typeE = nullptr;          // device copy of type_Energy
  struct lbm_output *out = nullptr;  // D3Q7 temperature (lbm_q7.hip), thermal == 2
  // ---- cells of the other types (IN, OUT, P, L, T, Q) and the solids that an
  // outlet copies from (lbm_faces.hip): FP64 populations and rho, u per slot.
  // ENERGY (thermal == 1): g too, and every cell an outlet copies from has a
  // slot; the fluid ones hold their pulled f, g (thermal_face_sources)
  int32_t *fslot = nullptr;   // per cell: its slot or -1
  double *fof = nullptr;      // 19 f per slot (ENERGY: then 19 g), direction-major
  double *fom = nullptr;      // rho, ux, uy, uz (ENERGY: then T, qx, qy, qz) per slot, component-major
  int32_t *fcells = nullptr;  // cell of each slot
  int32_t *fsrc = nullptr;    // ENERGY: (cell, slot) of the fluid cells an outlet copies from
  int fnsrc = 0;
  void *fxe = nullptr, *fye = nullptr, *fze = nullptr, *fme = nullptr;  // work lists
  int fnslot = 0;
  int fnx = 0, fnx0 = 0, fny = 0, fny0 = 0;  // i-face and j-face entries
  int fnz = 0, fnz_macro = 0;                 // columns with loop work; with MacroProp only too
  int fnm = 0, fnm_in = 0, fnm_xy = 0;        // MacroProp entries: all, inlets, inlets the i/j launches take
};

#define HIPCHK(x)                              \
  do {                                         \
    hipError_t e_ = (x);                       \
    if (e_ != hipSuccess) return LBM_ERR_HIP;  \
  } while (0)

// Two timing events, destroyed on every return path.
struct EventPair {
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipError_t create() {
    hipError_t e = hipEventCreate(&t0);
    return e == hipSuccess ? hipEventCreate(&t1) : e;
  }
  // waits for t1; *ms = t1 - t0
  int elapsed(double *ms) const {
    float t = 0.f;
    if (hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(&t, t0, t1) != hipSuccess)
      return LBM_ERR_HIP;
    *ms = (double)t;
    return LBM_OK;
  }
  ~EventPair() {
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
  }
};

// thermal branch (lbm_thermal.hip)
void thermal_free(lbm_ctx *c);
int thermal_init(lbm_ctx *c);          // tables, f and g = equilibrium
int thermal_tables(lbm_ctx *c);        // Dirichlet/Neumann links, wall temperature
// mode: 0 fused stream+collide, 1 as 0 + emit macros, 2 stream only,
// 3 collide in place, 4 MacroProp of the cells with mask bits `sel`
int thermal_launch(lbm_ctx *c, int mode, uint8_t sel = 1);
int thermal_download_g(lbm_ctx *c, double *g_dev);

// D3Q7 temperature (lbm_q7.hip), thermal == 2
enum Q7Mode {
  Q7_FUSED = 0,       // Streaming + MacroProp + next Collision
  Q7_FUSED_EMIT = 1,  // as Q7_FUSED, and write rho, u, T, q
  Q7_STREAM = 2,      // Streaming alone
  Q7_STREAM_EMIT = 3, // Streaming + MacroProp
  Q7_COLLIDE = 4,     // Collision in place (rho, u, T as stored)
  Q7_MACRO = 5        // MacroProp of the fluid cells
};
int q7_upload(lbm_ctx *c, const lbm_thermal_desc *d, const double *T, const int16_t *type_energy);
int q7_init(lbm_ctx *c);    // Dirichlet link bits, g = geq
int q7_tables(lbm_ctx *c);  // Dirichlet link bits (after every k_tables)
// force_mode: the context's unless given (3: the IBM step)
int q7_launch(lbm_ctx *c, int mode, int force_mode = -1);
int q7_download_g(lbm_ctx *c, double *g_dev);

// VTK output staging (lbm_output.hip)
void output_free(lbm_ctx *c);

// face loops of Streaming and MacroProp of the cells of the other types
// (lbm_faces.hip): slots and Init populations (at lbm_init); the loops after
// a streaming launch (`oth` its source) with MacroProp of no cell (macro 0),
// the inlets (1) or every cell of the other types (2); MacroProp alone; the
// slotted cells' populations into a download of f.  ibm: the moments of an
// IBM run (F = 0, TYPE_IN/TYPE_OUT only, no range check).
void faces_free(lbm_ctx *c);
int faces_setup(lbm_ctx *c);
// the folded pull's 32-bit lane offsets fit this lattice (lbm_capi.hip)
bool fold_fits(const lbm_ctx *c);
int faces_step(lbm_ctx *c, bool ibm, int macro);
int faces_macro(lbm_ctx *c, bool ibm);
hipError_t faces_f_out(lbm_ctx *c, double *out);
hipError_t faces_g_out(lbm_ctx *c, double *out);  // ENERGY: g of the slotted cells
int thermal_face_sources(lbm_ctx *c);              // before a pulling ENERGY launch
int thermal_backeq_solid(lbm_ctx *c);              // g of the back-face write into solids

// launch parameters of the context's current state (lbm_capi.hip)
lbm::LaunchCtx launch_ctx(const lbm_ctx *c);
// the back-face inflow write into solids that the reference's Streaming
// makes (lbm_capi.hip): the first time, and with forced_inlets at every
// Streaming of a context whose forced inlets change it (IBM runs take the
// inlets' moments without force, so theirs stays as first written)
int after_stream(lbm_ctx *c, bool forced_inlets = false);
// the IBM state (lbm_ibm.hip): release; set up at all; IBLBM done and
// waiting for its collision; mark that collision done
void ibm_free(lbm_ctx *c);
int ibm_remark(lbm_ctx *c);
bool ibm_active(const lbm_ctx *c);
bool ibm_ready(const lbm_ctx *c);
void ibm_consume(lbm_ctx *c);


