// lbm_device.h -- device-side definitions shared by the gfx950 kernels.
//
// Lattice: D3Q19 in the reference's direction order (LBM.h:546-552).
// Device layout: structure of arrays, one plane of `ps` elements per
// direction, cell index idx = (z*ny + y)*px + x with x fastest over the
// padded (Nx+2)(Ny+2)(Nz+2) lattice and a row pitch px (multiple of 4).
// Two such population buffers (A/B) alternate every step.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "lbm_launch.h"

namespace lbm {

constexpr int Q = 19;
// LBM.h:546-552
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr double W[Q] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                         1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                         1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

// per-cell kernel mask bits (derived from the cell types at init); 0 marks
// a row-padding cell
constexpr uint8_t M_FLUID = 1;   // TYPE_F: collides and pulls
constexpr uint8_t M_WALL = 2;    // some x - c_m is TYPE_S: read links[idx]
constexpr uint8_t M_WRAP = 4;    // on the padded shell: neighbours wrap
constexpr uint8_t M_MOVING = 8;  // nonzero wall velocity: read uw*[idx]
constexpr uint8_t M_SOLID = 16;  // TYPE_S
constexpr uint8_t M_OTHER = 32;  // any other type (IN, OUT, P, L, T, Q)
// cell (i,j,Nz+1) whose (i,j,0) is TYPE_IN: the back-face inflow loop of the
// reference writes its 2nd-order feq there after every Streaming
// (lbm.cpp:599-611 reads k = 0 and writes k = Nz+1)
constexpr uint8_t M_BACKEQ = 64;
// inside the immersed-boundary box (force mode 3 reads the box arrays)
constexpr uint8_t M_BOX = 128;

// sticky status bits
constexpr int ST_FEQ_NEG = 1;
constexpr int ST_RHO_RANGE = 2;
constexpr int ST_T_RANGE = 4;  // temperature outside [0,100] (thermal branches)

template <int... Ms, class F>
__device__ __forceinline__ void for_dirs_impl(std::integer_sequence<int, Ms...>, F &&f) {
  (f(std::integral_constant<int, Ms>{}), ...);
}
// Calls f(std::integral_constant<int, m>) for m = 0..18, fully unrolled, so
// that c_m and w_m are compile-time constants inside f.
template <class F>
__device__ __forceinline__ void for_dirs(F &&f) {
  for_dirs_impl(std::make_integer_sequence<int, Q>{}, f);
}

// Linear index of x - c (the pulling source of direction c), each axis
// wrapped modulo its padded extent as the reference does
// ((i + cx + (Nx+2)) % (Nx+2), lbm.cpp:197-199).
__device__ __forceinline__ uint32_t wrap_src(const Geo &g, uint32_t x, uint32_t y, uint32_t z,
                                             int cx, int cy, int cz) {
  uint32_t xs = cx > 0 ? (x == 0 ? g.nx - 1 : x - 1) : (cx < 0 ? (x == g.nx - 1 ? 0 : x + 1) : x);
  uint32_t ys = cy > 0 ? (y == 0 ? g.ny - 1 : y - 1) : (cy < 0 ? (y == g.ny - 1 ? 0 : y + 1) : y);
  uint32_t zs = cz > 0 ? (z == 0 ? g.nz - 1 : z - 1) : (cz < 0 ? (z == g.nz - 1 ? 0 : z + 1) : z);
  return (zs * g.ny + ys) * g.px + xs;
}
// x + c, wrapped (the pushing target of direction c)
__device__ __forceinline__ uint32_t wrap_dst(const Geo &g, uint32_t x, uint32_t y, uint32_t z,
                                             int cx, int cy, int cz) {
  return wrap_src(g, x, y, z, -cx, -cy, -cz);
}
__device__ __forceinline__ void coords(const Geo &g, uint32_t idx, uint32_t &x, uint32_t &y,
                                       uint32_t &z) {
  const uint32_t t = __umulhi(idx, g.px_m) >> g.px_s;  // idx / px
  x = idx - t * g.px;
  z = __umulhi(t, g.ny_m) >> g.ny_s;  // t / ny
  y = t - z * g.ny;
}
// the pull offset of direction m for a cell off the shell: src = idx - off
template <int m>
__device__ __forceinline__ uint32_t pull_off(const Geo &g) {
  return (uint32_t)(CX[m] + CY[m] * (int)g.px + CZ[m] * (int)g.plane);
}

// ---- shell fold ------------------------------------------------------------
// A fluid cell x whose pull source x - c_m (wrapped) lies on a face of the
// padded lattice listed in g.fold (every cell of that face TYPE_S) takes the
// bounce-back value f*(x, opp m) + 6 w_m (u_w . c_m) from its own cell: plane
// opp m, which nothing else reads.  Source on face x = 0 for c_x > 0 means
// x == 1 (or x == 0 wrapping onto x = nx-1 when that face is the solid one);
// likewise for the other five faces.  Bit m of the result: direction m
// folds.  Cells off the fluid interior get meaningless bits.
template <int A, int SGN>
constexpr uint32_t axis_dirs() {
  uint32_t b = 0;
  for (int m = 0; m < Q; m++) {
    const int c = A == 0 ? CX[m] : A == 1 ? CY[m] : CZ[m];
    if (SGN > 0 ? c > 0 : c < 0) b |= 1u << m;
  }
  return b;
}
struct FoldLane {
  bool xp, xm, yp, ym, zp, zm;  // directions with c > 0 / c < 0 on each axis fold
};
__device__ __forceinline__ FoldLane fold_lane(const Geo &g, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t f = g.fold;
  FoldLane l;
  l.xp = ((f & 1u) && x == 1) || ((f & 2u) && x == 0);
  l.xm = ((f & 2u) && x == g.nx - 2) || ((f & 1u) && x == g.nx - 1);
  l.yp = ((f & 4u) && y == 1) || ((f & 8u) && y == 0);
  l.ym = ((f & 8u) && y == g.ny - 2) || ((f & 4u) && y == g.ny - 1);
  l.zp = ((f & 16u) && z == 1) || ((f & 32u) && z == 0);
  l.zm = ((f & 32u) && z == g.nz - 2) || ((f & 16u) && z == g.nz - 1);
  return l;
}
template <int m>
__device__ __forceinline__ bool folds(const FoldLane &l) {
  bool r = false;
  if constexpr (CX[m] > 0) r = r || l.xp;
  if constexpr (CX[m] < 0) r = r || l.xm;
  if constexpr (CY[m] > 0) r = r || l.yp;
  if constexpr (CY[m] < 0) r = r || l.ym;
  if constexpr (CZ[m] > 0) r = r || l.zp;
  if constexpr (CZ[m] < 0) r = r || l.zm;
  return r;
}
__device__ __forceinline__ uint32_t fold_mask(const Geo &g, uint32_t x, uint32_t y, uint32_t z) {
  if (!(g.fold & 63u)) return 0u;
  const FoldLane l = fold_lane(g, x, y, z);
  return (l.xp ? axis_dirs<0, 1>() : 0u) | (l.xm ? axis_dirs<0, -1>() : 0u) |
         (l.yp ? axis_dirs<1, 1>() : 0u) | (l.ym ? axis_dirs<1, -1>() : 0u) |
         (l.zp ? axis_dirs<2, 1>() : 0u) | (l.zm ? axis_dirs<2, -1>() : 0u);
}
// cell on a face listed in bits 0-5 (fold) or, with shift 8, bits 8-13
// (keep) of g.fold
template <int SHIFT = 0>
__device__ __forceinline__ bool on_fold_face(const Geo &g, uint32_t x, uint32_t y, uint32_t z) {
  const uint32_t f = g.fold >> SHIFT;
  return ((f & 1u) && x == 0) || ((f & 2u) && x == g.nx - 1) || ((f & 4u) && y == 0) ||
         ((f & 8u) && y == g.ny - 1) || ((f & 16u) && z == 0) || ((f & 32u) && z == g.nz - 1);
}

// ---- population storage -------------------------------------------------
// Stored value s, compute ("raw") value r and population f:
//   double: r = s = f                          K = 1  (unshifted, FP64)
//   float : r = s = f - w                      K = 1  (shifted, FP32)
//   half  : r = float(s) = (f - w) * 2^12      K = 4096 (shifted, scaled)
// Kernels work on raw values: the scale K is folded into the constants that
// produce equilibria and sources, and 1/K into the moments.
template <class S>
struct Store;


template <>
struct Store<double> {
  using C = double;
  static constexpr bool shifted = false;
  static constexpr double K = 1.0;
  static __device__ __forceinline__ double ld(const double *p) { return *p; }
  static __device__ __forceinline__ void st(double *p, double r) { __builtin_nontemporal_store(r, p); }
  static __device__ __forceinline__ void st_raw(double *p, double v) { __builtin_nontemporal_store(v, p); }
  using E = double;  // encoded value in a register (encw / st_e)
  static __device__ __forceinline__ double encw(double r) { return r; }
  static __device__ __forceinline__ void st_e(double *p, double v) { __builtin_nontemporal_store(v, p); }
  static __device__ __forceinline__ double enc(double r) { return r; }
};

template <>
struct Store<float> {
  using C = float;
  static constexpr bool shifted = true;
  static constexpr float K = 1.0f;
  static __device__ __forceinline__ float ld(const float *p) { return *p; }
  static __device__ __forceinline__ void st(float *p, float r) { __builtin_nontemporal_store(r, p); }
  static __device__ __forceinline__ void st_raw(float *p, float v) { __builtin_nontemporal_store(v, p); }
  using E = float;
  static __device__ __forceinline__ float encw(float r) { return r; }
  static __device__ __forceinline__ void st_e(float *p, float v) { __builtin_nontemporal_store(v, p); }
  static __device__ __forceinline__ float enc(float r) { return r; }
};

constexpr float F16S_SCALE = 4096.0f;  // 2^LBM_F16S_SCALE_LOG2

template <>
struct Store<__half> {
  using C = float;
  static constexpr bool shifted = true;
  static constexpr float K = F16S_SCALE;
  static __device__ __forceinline__ float ld(const __half *p) {
    return (float)*reinterpret_cast<const _Float16 *>(p);
  }
  static __device__ __forceinline__ void st(__half *p, float r) {
    __builtin_nontemporal_store((_Float16)r, reinterpret_cast<_Float16 *>(p));
  }
  static __device__ __forceinline__ void st_raw(__half *p, __half v) {
    __builtin_nontemporal_store(*reinterpret_cast<const _Float16 *>(&v), reinterpret_cast<_Float16 *>(p));
  }
  using E = _Float16;
  static __device__ __forceinline__ _Float16 encw(float r) { return (_Float16)r; }
  static __device__ __forceinline__ void st_e(__half *p, _Float16 v) {
    __builtin_nontemporal_store(v, reinterpret_cast<_Float16 *>(p));
  }
  static __device__ __forceinline__ __half enc(float r) {
    _Float16 h = (_Float16)r;
    return *reinterpret_cast<__half *>(&h);
  }
};

// LBM_STORE_F16C: f - w_i in a 16-bit float of this library's own layout,
// sign | 4 exponent bits (bias 17) | 11 fraction bits, with subnormals: one
// more significant bit than IEEE half over |f - w| < 0.5, the range shifted
// populations occupy (the largest, f_0 - w_0 = w_0 (rho - 1 - 1.5 rho u^2),
// stays below 0.5 for |rho - 1| < 1.5).  Values beyond it saturate.
//
// The kernels work on these populations in units of 2^-110 (K): the FP32
// value whose bit pattern is the 16-bit code shifted left by 12 (sign moved
// to bit 31) is then the population itself, the format's exponent field
// being the low 4 bits of FP32's, and its subnormals FP32 denormals (kept,
// not flushed).  Decoding is a shift and a mask; encoding rounds the FP32
// value to nearest even on the 12 low bits, after a clamp to the largest
// code, and moves the sign back.
struct f16c_t {
  uint16_t bits;
};

constexpr float F16C_MAX_RAW = 0x1.ffep-112f;  // code 0x7fff in units of 2^-110

__device__ __forceinline__ float f16c_decode(int32_t h) {  // h sign-extended
  return __uint_as_float(((uint32_t)h << 12) & 0x87fff000u);
}
// the code in the high half of the result (the low half is not zero)
__device__ __forceinline__ uint32_t f16c_encode_hi(float r) {
  uint32_t b = __float_as_uint(__builtin_amdgcn_fmed3f(r, -F16C_MAX_RAW, F16C_MAX_RAW));
  b += 0x7ffu + ((b >> 12) & 1u);  // magnitude bits only: no carry reaches the sign
  // bits 30..16 <- magnitude bits 26..12, bit 31 the sign
  // v_bfi_b32: (mask & b) | (~mask & (b << 4)), one instruction where the
  // compiler's own lowering of the select takes three
  uint32_t d;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(d) : "s"(0x80000000u), "v"(b), "v"(b << 4));
  return d;
}
__device__ __forceinline__ uint32_t f16c_encode(float r) { return f16c_encode_hi(r) >> 16; }

template <>
struct Store<f16c_t> {
  using C = float;
  static constexpr bool shifted = true;
  static constexpr float K = 0x1p-110f;
  static __device__ __forceinline__ float ld(const f16c_t *p) {
    return f16c_decode(*reinterpret_cast<const int16_t *>(p));
  }
  static __device__ __forceinline__ void st(f16c_t *p, float r) {
    __builtin_nontemporal_store((uint16_t)f16c_encode(r), reinterpret_cast<uint16_t *>(p));
  }
  static __device__ __forceinline__ void st_raw(f16c_t *p, f16c_t v) {
    __builtin_nontemporal_store(v.bits, reinterpret_cast<uint16_t *>(p));
  }
  static __device__ __forceinline__ f16c_t enc(float r) { return f16c_t{(uint16_t)f16c_encode(r)}; }
  using E = uint32_t;  // the code in the high half: stored by a high-half store
  static __device__ __forceinline__ uint32_t encw(float r) { return f16c_encode_hi(r); }
  static __device__ __forceinline__ void st_e(f16c_t *p, uint32_t v) {
    __builtin_nontemporal_store((uint16_t)(v >> 16), reinterpret_cast<uint16_t *>(p));
  }
};

// Cubic equilibrium of LBM::Calc_feq (lbm.cpp:628-633) with cs2 = 1/3:
//   feq = rho w (1 + A),  A = 3cu + 4.5cu^2 - 1.5uu + 4.5cu^3 - 4.5 cu uu.
template <class C>
__device__ __forceinline__ C eq_A(C cu, C uu) {
  return cu * (C(3) + C(4.5) * cu * cu - C(4.5) * uu) + C(4.5) * cu * cu - C(1.5) * uu;
}

// Second-order inflow equilibrium of the face loops (lbm.cpp:337 etc.):
//   feq2 = rho w (1 + 3cu + 4.5cu^2 - 1.5uu); returned as a raw value of the
//   storage (shifted by w and scaled by K where the storage is).
template <class C, bool SHIFTED>
__device__ __forceinline__ C feq2_raw(C rho, C cu, C uu, C w, C K) {
  const C B = C(3) * cu + C(4.5) * cu * cu - C(1.5) * uu;
  if constexpr (SHIFTED) return K * w * ((rho - C(1)) + rho * B);
  else return rho * w * (C(1) + B);
}

}  // namespace lbm
