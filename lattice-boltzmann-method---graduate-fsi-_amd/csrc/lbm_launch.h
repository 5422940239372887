// lbm_launch.h -- host-side launch interface between the C-ABI layer and the
// kernel translation unit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

// Device lattice geometry.  Cell (x,y,z) of the padded (Nx+2)(Ny+2)(Nz+2)
// lattice sits at idx = (z*ny + y)*px + x.  The row pitch px >= nx is a
// multiple of 4 (rows start 16-byte aligned for FP32); cells with x >= nx
// are dead padding (mask 0, never read by a live cell).
struct Geo {
  uint32_t nx, ny, nz;  // padded extents Nx+2, Ny+2, Nz+2
  uint32_t px;          // row pitch (cells), multiple of 4
  uint32_t plane;       // px*ny
  uint32_t ncell;       // px*ny*nz (includes row padding)
  uint32_t ps;          // population plane stride (elements), >= ncell
  // n / px and n / ny for n < 2^31 as (mulhi(n, m) >> s) (set_divisors)
  uint32_t px_m, px_s, ny_m, ny_s;
  // faces of the padded lattice whose cells are all TYPE_S and whose wall
  // links the pull resolves in the address (bit 0: x = 0, 1: x = nx-1,
  // 2: y = 0, 3: y = ny-1, 4: z = 0, 5: z = nz-1); bits 8-13, the same
  // faces: every cell neither fluid nor solid (inlets, outlets) except
  // solids of a folded face, so the stream kernel rewrites those cells'
  // frozen-link values unchanged instead of leaving holes in its stores.
  // 0 in the thermal branches and when the lattice is too large for 32-bit
  // offsets
  uint32_t fold;
};

// Division by an invariant d >= 2 for 0 <= n < 2^31 (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994):
// with l = ceil(log2 d) and m = ceil(2^(31+l) / d) < 2^32,
// floor(n / d) = floor(n m / 2^(31+l)), because n (m d - 2^(31+l)) <
// 2^31 d.  Returns false when d is out of range.
inline bool magic_divisor(uint32_t d, uint32_t &m, uint32_t &s) {
  if (d < 2) return false;
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  const unsigned __int128 num = (unsigned __int128)1 << (31 + l);
  m = (uint32_t)((num + d - 1) / d);
  s = l - 1;
  return true;
}
inline bool set_divisors(Geo &g) {
  return magic_divisor(g.px, g.px_m, g.px_s) && magic_divisor(g.ny, g.ny_m, g.ny_s);
}

// stream kernel variants
enum StreamMode {
  STREAM_ONLY = 0,          // LBM::Streaming alone (store f)
  STREAM_MACRO = 1,         // Streaming + MacroProp (store f, emit rho,u)
  STREAM_COLLIDE = 2,       // Streaming + MacroProp + next Collision (store f*)
  STREAM_COLLIDE_MACRO = 3  // as STREAM_COLLIDE, and emit rho,u
};

// IBM box data of force mode 3 (the immersed-boundary step): cells inside
// [x0, x0+nx) x [y0, y0+ny) x [z0, z0+nz) (mask bit M_BOX) take rho, u and
// F from these FP64 arrays, indexed (x-x0) + nx*((y-y0) + ny*(z-z0)).
struct BoxArgs {
  int32_t x0, y0, z0, nx, ny, nz;
  const double *rho, *ux, *uy, *uz, *Fx, *Fy, *Fz;
};

struct LaunchCtx {
  Geo g;
  int storage;     // LBM_STORE_*
  int force_mode;  // 0 none, 1 per-cell field, 2 uniform
  int block;       // threads per block
  int cpt;         // cells per thread of the fused k_stream launch (grid-stride passes)
  double omega;
  double Fu[3];
  const void *Fx, *Fy, *Fz;
  const uint32_t *links;
  const void *uwx, *uwy, *uwz;
  void *rho, *ux, *uy, *uz;
  const uint8_t *mask;
  BoxArgs box;        // force mode 3
  int wall_slots;     // bounce-back through the solids' slots
  int slot_links;     // ... of which some are not folded (Geo::fold): links are read
  uint8_t macro_sel;  // k_macro: cells whose mask has any of these bits
  int *status;
  unsigned long long *first_bad;
  hipStream_t stream;
};

// one launch over an ensemble of lattices (k_stream_batch)
struct BatchLaunch {
  int storage, force_mode, mode;
  const void *la;  // LatArgs<S>[2 * nlat]: both buffer parities per lattice
  const uint32_t *tile_lat, *tile_first;
  uint32_t ntile, per;
  int parity;
  int fold;  // some lattice has folded or kept faces (Geo::fold)
  unsigned grid, block;
  hipStream_t stream;
};
size_t batch_args_size(int storage);
void batch_fill_args(const LaunchCtx &L, const void *in, void *out, void *dst);
hipError_t launch_stream_batch(const BatchLaunch &B);

hipError_t launch_stream(const LaunchCtx &L, const void *fin, void *fout, int mode);
// the same launches for lattices with folded or kept faces (lbm_stream_fold.hip)
hipError_t launch_stream_fold(const LaunchCtx &L, const void *fin, void *fout, int mode);
hipError_t launch_stream_batch_fold(const BatchLaunch &B);
hipError_t launch_collide(const LaunchCtx &L, void *buf);
hipError_t launch_macro(const LaunchCtx &L, const void *buf);
hipError_t launch_init(const LaunchCtx &L, void *A, void *B);
hipError_t launch_tables(const LaunchCtx &L, uint8_t *mask, uint32_t *links, void *uwx,
                         void *uwy, void *uwz, const int16_t *type);
hipError_t launch_backeq_solid(const LaunchCtx &L, void *A, void *B);
hipError_t launch_links(const LaunchCtx &L, int32_t *src, uint8_t *kind);
hipError_t launch_f_out(const LaunchCtx &L, const void *buf, double *out);
int reduce_blocks(uint32_t n);
hipError_t launch_reduce(const LaunchCtx &L, int what, double *partial, double *out);

}  // namespace lbm
