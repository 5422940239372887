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
// outlet copies from, has a slot: its 19 populations in `of` and its rho, u
// in `om`, both FP64.  Their plane entries keep carrying the frozen values of
// their fluid neighbours.  So does a fluid cell at k = Nz+1 under an inlet
// at k = 0 that an outlet copies from: the k = 0 loop overwrites all its 19
// populations, so its frozen links (from cells of the other types) hold that
// write, which its slot keeps, not their Init values.
//
// Per step, after the streaming launch that pulled from buffer `src`:
//   launch 1  faces i = 0 and i = Nx+1 (they write disjoint cells and read
//             none that the other writes);
//   launch 2  faces j = Ny+1 and j = 0 (likewise);
//   launch 3  per (i,j) column the k = Nz+1 loop, then the k = 0 loop, then
//             MacroProp of the column's two end cells; MacroProp of the
//             other slotted cells.
// A loop thus reads the values the earlier loops of this step wrote and the
// previous step's values of cells that later loops write, as the
// reference's order gives.  A fluid neighbour's population is its pulled
// value, read again from `src` where the streaming read it (plane l at
// x - c_l: a fluid cell's f*, a solid's bounce-back slot or a frozen value).
//
// In IBM runs (main.cpp:197-211) MacroProp is not called: the IBM step takes
// rho = sum f, u = sum f c / rho of the TYPE_IN/TYPE_OUT cells with F = 0 and
// no range check (lbm.cpp:1262-1285, 1173-1193), which mode 3 does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_step.h"

using namespace lbm;

namespace {

constexpr int FACE_THREADS = 256;

template <class S>
struct FaceArgs {
  using C = typename Store<S>::C;
  Geo g;
  const S *src;          // the source buffer of the last streaming launch
  const int16_t *type;   // cell types, padded order
  const int32_t *slot;   // slot of each cell, -1 for cells without one
  double *of;            // 19 populations per slot
  double *om;            // rho, ux, uy, uz per slot
  C *rho, *ux, *uy, *uz;       // the context's macro arrays
  const C *Fx, *Fy, *Fz;       // force mode 1
  double Fu[3];                // force mode 2
  int mode;                    // force mode 0/1/2; 3: the IBM step's moments
  int *status;
  unsigned long long *first_bad;
};

__device__ __forceinline__ uint32_t cell(const Geo &g, uint32_t x, uint32_t y, uint32_t z) {
  return (z * g.ny + y) * g.px + x;
}

// Population L of cell s as the face loops see it: a slotted cell's own
// value; for a fluid cell the value the streaming pulled into it, or, on a
// frozen link of a slotted fluid cell, the last k = 0 inflow write.
template <class S, int L>
__device__ __forceinline__ double pop(const FaceArgs<S> &a, uint32_t s) {
  const int32_t k = a.slot[s];
  if (k >= 0 && a.type[s] != LBM_TYPE_F) return a.of[(size_t)k * Q + L];
  uint32_t x, y, z;
  coords(a.g, s, x, y, z);
  const uint32_t q = wrap_src(a.g, x, y, z, CX[L], CY[L], CZ[L]);
  if (k >= 0) {
    const int16_t tq = a.type[q];
    if (tq != LBM_TYPE_F && tq != LBM_TYPE_S) return a.of[(size_t)k * Q + L];
  }
  double v = (double)Store<S>::ld(a.src + (size_t)L * a.g.ps + q);
  if constexpr (Store<S>::shifted) v = v / (double)Store<S>::K + W[L];
  return v;
}

// TYPE_OUT: f[D] of cell d = f[D] of its inner neighbour s
template <class S, int D0, int D1, int D2>
__device__ __forceinline__ void out_copy(const FaceArgs<S> &a, uint32_t d, uint32_t s) {
  double *o = a.of + (size_t)a.slot[d] * Q;
  const double v0 = pop<S, D0>(a, s), v1 = pop<S, D1>(a, s), v2 = pop<S, D2>(a, s);
  o[D0] = v0;
  o[D1] = v1;
  o[D2] = v2;
}

// TYPE_IN: the second-order equilibrium of cell s's rho, u into slot k
// (lbm.cpp:325-337), the reference's expression with cs2 = 1/3
template <class S>
__device__ __forceinline__ void in_feq2(const FaceArgs<S> &a, uint32_t s, int32_t k) {
#pragma clang fp contract(off)
  const double *m = a.om + (size_t)a.slot[s] * 4;
  const double rho = m[0], ux = m[1], uy = m[2], uz = m[3];
  const double cs2 = 1.0 / 3.0;
  const double u_dot_u = ux * ux + uy * uy + uz * uz;
  double *o = a.of + (size_t)k * Q;
  for_dirs([&](auto M) {
    constexpr int l = decltype(M)::value;
    const double c_dot_u = ux * CX[l] + uy * CY[l] + uz * CZ[l];
    o[l] = rho * W[l] * (1 + c_dot_u / cs2 + (c_dot_u * c_dot_u) / (2 * cs2 * cs2) - u_dot_u / (2 * cs2));
  });
}

template <class S>
__device__ __forceinline__ void face_cell(const FaceArgs<S> &a, uint32_t d, uint32_t s, int face) {
  const int16_t t = a.type[d];
  if (t == LBM_TYPE_IN) {
    in_feq2(a, d, a.slot[d]);
  } else if (t == LBM_TYPE_OUT) {
    switch (face) {
      case 0: out_copy<S, 1, 5, 8>(a, d, s); break;  // i = 0, lbm.cpp:303-306
      case 1: out_copy<S, 3, 6, 7>(a, d, s); break;  // i = Nx+1, lbm.cpp:357-360
      case 2: out_copy<S, 4, 7, 8>(a, d, s); break;  // j = Ny+1, lbm.cpp:412-415
      case 3: out_copy<S, 2, 5, 6>(a, d, s); break;  // j = 0, lbm.cpp:467-470
      default: out_copy<S, 4, 7, 8>(a, d, s); break; // k = Nz+1, k = 0
    }
  }
}

// MacroProp of one slotted cell of the other types (lbm.cpp:668-675, 772-775):
// the sums in the reference's l order.
template <class S>
__device__ __forceinline__ void other_macro(const FaceArgs<S> &a, uint32_t idx) {
#pragma clang fp contract(off)
  using C = typename Store<S>::C;
  const int16_t t = a.type[idx];
  if (t == LBM_TYPE_F || t == LBM_TYPE_S) return;
  if (a.mode == 3 && t != LBM_TYPE_IN && t != LBM_TYPE_OUT) return;
  const int32_t k = a.slot[idx];
  const double *f = a.of + (size_t)k * Q;
  double rho = 0, jx = 0, jy = 0, jz = 0;
  for_dirs([&](auto M) {
    constexpr int l = decltype(M)::value;
    rho += f[l];
    jx += f[l] * CX[l];
    jy += f[l] * CY[l];
    jz += f[l] * CZ[l];
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
  double *m = a.om + (size_t)k * 4;
  m[0] = rho;
  m[1] = ux;
  m[2] = uy;
  m[3] = uz;
  a.rho[idx] = (C)rho;
  a.ux[idx] = (C)ux;
  a.uy[idx] = (C)uy;
  a.uz[idx] = (C)uz;
  if (a.mode != 3 && (rho < 0 || rho > 100)) flag(a.status, a.first_bad, ST_RHO_RANGE, idx);
}

// launch 1 (x faces, AXIS 0) and launch 2 (y faces, AXIS 1): thread t
// handles one cell of face 2*AXIS + (t >= n), n = cells per face
template <class S, int AXIS>
__global__ __launch_bounds__(FACE_THREADS) void k_faces_xy(FaceArgs<S> a) {
  const Geo &g = a.g;
  const uint32_t n = AXIS == 0 ? g.ny * g.nz : g.nx * g.nz;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * n) return;
  const uint32_t side = t >= n, r = t - side * n;
  uint32_t d, s;
  if constexpr (AXIS == 0) {
    const uint32_t y = r % g.ny, z = r / g.ny;
    d = cell(g, side ? g.nx - 1 : 0, y, z);
    s = cell(g, side ? g.nx - 2 : 1, y, z);
  } else {
    const uint32_t x = r % g.nx, z = r / g.nx;
    // face 2 is j = Ny+1, face 3 is j = 0
    d = cell(g, x, side ? 0 : g.ny - 1, z);
    s = cell(g, x, side ? 1 : g.ny - 2, z);
  }
  if (a.slot[d] < 0) return;
  face_cell(a, d, s, 2 * AXIS + (int)side);
}

// launch 3: threads [0, nx*ny) take the columns, the rest the listed
// slotted cells with 0 < k < Nz+1 (MacroProp only); macro = 0: the loops
// alone (a Streaming call without the MacroProp of a fused step)
template <class S>
__global__ __launch_bounds__(FACE_THREADS) void k_faces_z(FaceArgs<S> a, const int32_t *__restrict__ mid,
                                                          int nmid, int macro) {
  const Geo &g = a.g;
  const uint32_t ncol = g.nx * g.ny;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol) {
    if (macro && t - ncol < (uint32_t)nmid) other_macro(a, (uint32_t)mid[t - ncol]);
    return;
  }
  const uint32_t x = t % g.nx, y = t / g.nx;
  const uint32_t top = cell(g, x, y, g.nz - 1), bot = cell(g, x, y, 0);
  const bool st = a.slot[top] >= 0, sb = a.slot[bot] >= 0;
  if (!st && !sb) return;
  if (st) face_cell(a, top, cell(g, x, y, g.nz - 2), 4);  // k = Nz+1, lbm.cpp:518-570
  if (sb) {                                                 // k = 0, lbm.cpp:572-625
    const int16_t tb = a.type[bot];
    if (tb == LBM_TYPE_OUT) out_copy<S, 4, 7, 8>(a, bot, cell(g, x, y, 1));
    else if (tb == LBM_TYPE_IN && st) in_feq2(a, bot, a.slot[top]);
  }
  if (!macro) return;
  if (st) other_macro(a, top);
  if (sb) other_macro(a, bot);
}

// MacroProp of every slotted cell of the other types (lbm_macro; the moments
// at the start of an IBM run)
template <class S>
__global__ __launch_bounds__(FACE_THREADS) void k_faces_macro(FaceArgs<S> a, const int32_t *__restrict__ cells,
                                                              int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) other_macro(a, (uint32_t)cells[t]);
}

// Init (lbm.cpp:31-57): f = the cubic equilibrium of the set-up rho, u
template <class S>
__global__ __launch_bounds__(FACE_THREADS) void k_faces_init(FaceArgs<S> a, const int32_t *__restrict__ cells,
                                                             int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t idx = (uint32_t)cells[k];
  const double rho = (double)a.rho[idx], ux = (double)a.ux[idx], uy = (double)a.uy[idx], uz = (double)a.uz[idx];
  const double uu = ux * ux + uy * uy + uz * uz;
  double *o = a.of + (size_t)k * Q;
  for_dirs([&](auto M) {
    constexpr int l = decltype(M)::value;
    const double cu = ux * CX[l] + uy * CY[l] + uz * CZ[l];
    o[l] = rho * W[l] * (1.0 + eq_A(cu, uu));
  });
  double *m = a.om + (size_t)k * 4;
  m[0] = rho;
  m[1] = ux;
  m[2] = uy;
  m[3] = uz;
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
  out[(size_t)cells[k] * Q + l] = of[t];
}

template <class S>
FaceArgs<S> face_args(const lbm_ctx *c, int mode) {
  using C = typename Store<S>::C;
  FaceArgs<S> a;
  a.g = c->g;
  a.src = (const S *)c->oth;
  a.type = c->type;
  a.slot = c->fslot;
  a.of = c->fof;
  a.om = c->fom;
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
  return a;
}

unsigned blocks(uint64_t n) { return (unsigned)((n + FACE_THREADS - 1) / FACE_THREADS); }

template <class S>
hipError_t faces_step_s(lbm_ctx *c, int mode, int macro) {
  const FaceArgs<S> a = face_args<S>(c, mode);
  const Geo &g = c->g;
  if (c->face_x) hipLaunchKernelGGL((k_faces_xy<S, 0>), dim3(blocks(2ull * g.ny * g.nz)), dim3(FACE_THREADS), 0, c->stream, a);
  if (c->face_y) hipLaunchKernelGGL((k_faces_xy<S, 1>), dim3(blocks(2ull * g.nx * g.nz)), dim3(FACE_THREADS), 0, c->stream, a);
  hipLaunchKernelGGL(k_faces_z<S>, dim3(blocks((uint64_t)g.nx * g.ny + c->fnmid)), dim3(FACE_THREADS), 0, c->stream,
                     a, c->fmid, c->fnmid, macro);
  return hipGetLastError();
}

template <class S>
hipError_t faces_macro_s(lbm_ctx *c, int mode) {
  const FaceArgs<S> a = face_args<S>(c, mode);
  hipLaunchKernelGGL(k_faces_macro<S>, dim3(blocks((uint64_t)c->fnslot)), dim3(FACE_THREADS), 0, c->stream, a,
                     c->fcells, c->fnslot);
  return hipGetLastError();
}

template <class S>
hipError_t faces_init_s(lbm_ctx *c) {
  const FaceArgs<S> a = face_args<S>(c, 0);
  hipLaunchKernelGGL(k_faces_init<S>, dim3(blocks((uint64_t)c->fnslot)), dim3(FACE_THREADS), 0, c->stream, a,
                     c->fcells, c->fnslot);
  return hipGetLastError();
}

hipError_t dispatch_step(lbm_ctx *c, int mode, int macro) {
  switch (c->storage) {
    case LBM_STORE_F64: return faces_step_s<double>(c, mode, macro);
    case LBM_STORE_F32: return faces_step_s<float>(c, mode, macro);
    case LBM_STORE_F16S: return faces_step_s<__half>(c, mode, macro);
    default: return faces_step_s<f16c_t>(c, mode, macro);
  }
}
hipError_t dispatch_macro(lbm_ctx *c, int mode) {
  switch (c->storage) {
    case LBM_STORE_F64: return faces_macro_s<double>(c, mode);
    case LBM_STORE_F32: return faces_macro_s<float>(c, mode);
    case LBM_STORE_F16S: return faces_macro_s<__half>(c, mode);
    default: return faces_macro_s<f16c_t>(c, mode);
  }
}
hipError_t dispatch_init(lbm_ctx *c) {
  switch (c->storage) {
    case LBM_STORE_F64: return faces_init_s<double>(c);
    case LBM_STORE_F32: return faces_init_s<float>(c);
    case LBM_STORE_F16S: return faces_init_s<__half>(c);
    default: return faces_init_s<f16c_t>(c);
  }
}

}  // namespace

void faces_free(lbm_ctx *c) {
  void *ptrs[] = {c->fslot, c->fof, c->fom, c->fcells, c->fmid};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->fslot = nullptr;
  c->fof = nullptr;
  c->fom = nullptr;
  c->fcells = nullptr;
  c->fmid = nullptr;
  c->fnslot = c->fnmid = 0;
  c->face_x = c->face_y = 0;
}

// Slots for the cells of the other types and for the solids that an outlet
// copies from, their Init populations, and the work lists.  Call after the
// set-up rho, u are in the macro arrays and before anything overwrites the
// solids' entries there.  The ENERGY branch keeps these cells at their set-up
// values (no slots).
int faces_setup(lbm_ctx *c) {
  faces_free(c);
  if (c->thermal == 1) return LBM_OK;
  const lbm::Geo &g = c->g;
  const std::vector<int16_t> &t = c->h_type;
  auto at = [&](uint32_t x, uint32_t y, uint32_t z) { return (size_t)((z * g.ny + y) * g.px + x); };
  std::vector<int32_t> slot(g.ncell, -1);
  std::vector<int32_t> cells, mid;
  auto give = [&](size_t d) {
    if (slot[d] < 0) {
      slot[d] = (int32_t)cells.size();
      cells.push_back((int32_t)d);
    }
  };
  for (uint32_t z = 0; z < g.nz; z++)
    for (uint32_t y = 0; y < g.ny; y++)
      for (uint32_t x = 0; x < g.nx; x++) {
        const size_t d = at(x, y, z);
        if (t[d] != LBM_TYPE_F && t[d] != LBM_TYPE_S) {
          give(d);
          if (z > 0 && z < g.nz - 1) mid.push_back((int32_t)d);
        }
      }
  if (cells.empty()) return LBM_OK;
  // solids that some outlet copies from (each face's inner neighbour), and
  // fluid cells at k = Nz+1 under an inlet at k = 0 that some outlet copies
  // from
  const size_t top = (size_t)(g.nz - 1) * g.plane;
  auto src_solid = [&](size_t d, size_t s) {
    if (t[d] != LBM_TYPE_OUT) return;
    if (t[s] == LBM_TYPE_S || (t[s] == LBM_TYPE_F && s >= top && t[s - top] == LBM_TYPE_IN)) give(s);
  };
  int fx = 0, fy = 0;
  for (uint32_t z = 0; z < g.nz; z++)
    for (uint32_t y = 0; y < g.ny; y++) {
      src_solid(at(0, y, z), at(1, y, z));
      src_solid(at(g.nx - 1, y, z), at(g.nx - 2, y, z));
      for (size_t d : {at(0, y, z), at(g.nx - 1, y, z)})
        if (t[d] == LBM_TYPE_IN || t[d] == LBM_TYPE_OUT) fx = 1;
    }
  for (uint32_t z = 0; z < g.nz; z++)
    for (uint32_t x = 0; x < g.nx; x++) {
      src_solid(at(x, g.ny - 1, z), at(x, g.ny - 2, z));
      src_solid(at(x, 0, z), at(x, 1, z));
      for (size_t d : {at(x, 0, z), at(x, g.ny - 1, z)})
        if (t[d] == LBM_TYPE_IN || t[d] == LBM_TYPE_OUT) fy = 1;
    }
  for (uint32_t y = 0; y < g.ny; y++)
    for (uint32_t x = 0; x < g.nx; x++) {
      src_solid(at(x, y, g.nz - 1), at(x, y, g.nz - 2));
      src_solid(at(x, y, 0), at(x, y, 1));
    }
  const size_t ns = cells.size();
  if (ns > (size_t)INT32_MAX / Q) return LBM_ERR_UNSUPPORTED;
  if (hipMalloc((void **)&c->fslot, g.ncell * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void **)&c->fof, ns * Q * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&c->fom, ns * 4 * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&c->fcells, ns * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void **)&c->fmid, std::max<size_t>(mid.size(), 1) * sizeof(int32_t)) != hipSuccess) {
    faces_free(c);
    return LBM_ERR_NOMEM;
  }
  c->fnslot = (int)ns;
  c->fnmid = (int)mid.size();
  c->face_x = fx;
  c->face_y = fy;
  HIPCHK(hipMemcpyAsync(c->fslot, slot.data(), g.ncell * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->fcells, cells.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  if (!mid.empty())
    HIPCHK(hipMemcpyAsync(c->fmid, mid.data(), mid.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(dispatch_init(c));
  // the host vectors go out of scope: the copies must have landed
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

// The face loops and (macro) MacroProp of the slotted cells after a
// streaming launch (the buffers already swapped: `oth` is the launch's
// source).  ibm: the moments of an IBM run instead of MacroProp.
int faces_step(lbm_ctx *c, bool ibm, bool macro) {
  if (!c->fnslot) return LBM_OK;
  HIPCHK(dispatch_step(c, ibm ? 3 : c->force_mode, macro ? 1 : 0));
  return LBM_OK;
}

int faces_macro(lbm_ctx *c, bool ibm) {
  if (!c->fnslot) return LBM_OK;
  HIPCHK(dispatch_macro(c, ibm ? 3 : c->force_mode));
  return LBM_OK;
}

hipError_t faces_f_out(lbm_ctx *c, double *out) {
  if (!c->fnslot) return hipSuccess;
  hipLaunchKernelGGL(k_faces_f_out, dim3(blocks((uint64_t)c->fnslot * Q)), dim3(FACE_THREADS), 0, c->stream,
                     c->fof, c->fcells, c->type, c->fnslot, out);
  return hipGetLastError();
}
