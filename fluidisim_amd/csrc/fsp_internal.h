// Internal device structures of the MI355X projection library.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fsp {

// Per-cell row code of the fine operator, rebuilt once per projection.
// Bits 0-5: the face couples to a wet neighbour, in the reference's row order
// x-, x+, y-, y+, z-, z+ (delta3d_mobile.rs:149-188). Bit 6: the cell is wet.
// Bit 7: the row carries at least one ghost (free-surface) face; such rows
// are evaluated from the geometry itself.
enum : uint8_t {
  M_XM = 1, M_XP = 2, M_YM = 4, M_YP = 8, M_ZM = 16, M_ZP = 32, M_WET = 64, M_GHOST = 128
};

constexpr int BLOCK = 256;        // threads per workgroup (4 wavefronts of 64)
constexpr int MAX_BLOCKS = 2048;  // grid cap of grid-stride kernels (8 per CU)
constexpr int NUM_SLOTS = 16;     // reduction slots
constexpr int MAX_PARTIALS = 65536;  // blocks a reducing kernel may have
constexpr int G_SLOTS = 8;        // ghost rows tabulated per column

// Reduction slots / finish operations.
enum Op : int {
  OP_B2 = 0,        // |b|^2 (and finiteness)
  OP_RR = 1,        // |r|^2 of a recomputed true residual
  OP_DQ = 2,        // <d, A d>  -> alpha, or stop
  OP_RN = 3,        // |r|^2 after the update
  OP_RZ = 4,        // <r, z> at the end of an iteration -> beta, it++, loop test
  OP_RZ_PRIME = 5,  // <r, z> of a prime (beta = 0)
  OP_MEAN_B = 6,    // sum of b and count of active rows (gauge)
  OP_MEAN_P = 7,    // sum of p and count of active rows (gauge)
  OP_DIVMAX = 8,    // max |div| over wet rows, over plain rows
  OP_UMAX = 9,      // max |u| over all faces
  OP_BACKWARD = 10, // backward error max
  OP_COUNT = 11,    // wet cells, ghost faces
  OP_GENERIC = 12,  // plain sum (tests)
  OP_RN_TRUE = 13,  // |r|^2 of the true residual replacing the recurrence inside the loop
};

// Device-resident control block of the conjugate gradient.
struct Ctl {
  double b2, rr, rn, rz, dq, alpha, beta, tol_b2, physical_target;
  double actual;         // last true residual |b - Ap|^2
  double red[NUM_SLOTS][2];  // raw totals of every reduction slot
  double mean_b, mean_p;
  uint32_t it, max_iters, pass_limit;
  uint32_t replace_every;  // every replace_every-th iteration r becomes the true residual b - A p (0: never)
  int32_t active;        // 1 while the CG loop runs; every iteration kernel returns early at 0
  int32_t first_dir;     // 1: the next direction is d = z
  int32_t dsel;          // which direction buffer holds the current d
  int32_t nonfinite;     // a reduction met a non-finite value
  int32_t stop_dq;       // the loop stopped on <d, A d> <= 0
  int32_t pure_neumann;  // no Dirichlet face on the fine level (and gauge on)
  int32_t gauge;
  int32_t flexible;      // 1: beta = -alpha <z', q> / <z, r> (flexible CG), else <z', r'> / <z, r>
  int32_t stalled;       // a replaced true residual did not fall below the previous one: the pass ends
  double rr_replaced;    // |r|^2 of the last replaced true residual of the pass
  uint32_t wet_cells, ghost_faces;
  int32_t table_miss;    // a ghost row lies past its column's table: the general fine kernels run
  uint32_t counters[NUM_SLOTS];
};

// Fine-level geometry, passed by value to kernels.
struct Fine {
  int nx, ny, nz;
  int cut;               // aperture-weighted divergence (reference cut geometry)
  size_t n;              // cells
  float dx, inv;         // dx, 1/dx^2
  float rho, g, rest;
  int homogeneous;
  const uint8_t* solid;  // cells
  const uint8_t* ou;     // u-face apertures (0/1)
  const uint8_t* ov;
  const uint8_t* ow;
  const float* eta;      // columns
  const float* eta_rt;   // columns
  const uint8_t* mask;   // cells (row codes)
  const uint16_t* gk0;   // columns: lowest ghost row (0xFFFF: none)
  const float* gtab;     // columns x G_SLOTS x 8: 1/theta of the ghost faces
  const float* precv;    // cells: row_prec of every row (k_prec)
  // Fractional apertures (the reference's cut geometry): the u8 face arrays
  // hold 0 closed, 1 open, 2 fractional, and a code-2 face's aperture is in
  // au/av/aw (null and frac == 0 when every aperture is 0 or 1).
  const float* au;
  const float* av;
  const float* aw;
  int frac;
  // Level-set geometry (the APIC-3D projection): cells labelled water, air
  // or solid, wetness from the labels and the ghost fraction of an air face
  // theta = max(phi_c / (phi_c - phi_a), theta_min) from the level set phi at
  // the cell centres (p = 0 on the zero level). phi == null: the mobile
  // geometry (wetness and ghost fractions from the column heights).
  const uint8_t* water;  // cells: 1 water
  const float* phi;      // cells
  float theta_min;
  // Coupled surface (the perturbation step over a background wave,
  // delta3d_coupling.rs:97-189, delta3d_mobile.rs:92-147): eta above holds
  // the total surface, whose heights set wetness and theta; eta_val is the
  // perturbation's own surface (the up ghost's value) and bg_up (columns),
  // bg_x (u faces), bg_y (v faces) the background's terms added to the ghost
  // values. bg_up == null: not coupled, eta_val == eta.
  const float* eta_val;
  const float* bg_up;
  const float* bg_x;
  const float* bg_y;
  float prec_tab[7];     // 1/(n * inv) for a plain row with n faces
};

// One coarse level (delta3d_multigrid.rs:37-52).
struct Level {
  int nx, ny, nz;
  size_t n;
  float inv;
  int8_t* kind;          // 1 active, -1 air, 0 solid
  float* ou;
  float* ov;
  float* ow;
  float* diag;
  float* x;
  float* r;
  float* t;
  float* dd;             // Galerkin levels: Dirichlet part of the diagonal (units of inv)
  int galerkin;          // 1: P^T A P / 16 coarse operator, 0: re-discretised
  int compact;           // 1: smoothed by the compact quad kernels (Galerkin, binary16 couplings)
  uint16_t* cx;          // compact levels: binary16 coupling of each cell's x- face
  uint16_t* cy;          // ... y- face
  uint16_t* cz;          // ... z- face
  float* dg;             // compact levels: diagonal of each row (gc_diag of the binary16 couplings)
};

// ---------------------------------------------------------------- V-cycle storage
// V-cycle vectors (the fine iterate z and every level's x and r) may be held
// in 16-bit storage: arithmetic stays f32, and each value is rounded to the
// storage type (round to nearest even) when it is stored. Store<ST>::rnd is
// that rounding as an f32 value, so code that keeps a stored value in a
// register uses exactly what a reload would give.
enum : int { ST_F32 = 0, ST_F16 = 1, ST_BF16 = 2 };

template <int ST> struct Store;

template <> struct Store<ST_F32> {
  using T = float;
  static __device__ __forceinline__ float rnd(float v) { return v; }
  static __device__ __forceinline__ float ld(const T* p, size_t c) { return p[c]; }
  static __device__ __forceinline__ void st(T* p, size_t c, float v) { p[c] = v; }
  static __device__ __forceinline__ float2 ld2(const T* p, size_t c) { return *reinterpret_cast<const float2*>(p + c); }
  static __device__ __forceinline__ void st2(T* p, size_t c, float2 v) { *reinterpret_cast<float2*>(p + c) = v; }
  static __device__ __forceinline__ float4 ld4(const T* p, size_t c) { return *reinterpret_cast<const float4*>(p + c); }
  static __device__ __forceinline__ void st4(T* p, size_t c, float4 v) { *reinterpret_cast<float4*>(p + c) = v; }
};

// bfloat16: the upper half of the f32 bit pattern.
template <> struct Store<ST_BF16> {
  using T = uint16_t;
  static __device__ __forceinline__ float rnd(float v) {
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7F800000u) == 0x7F800000u) return v;  // inf / nan unchanged
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
  }
  static __device__ __forceinline__ float up(uint32_t h) { return __uint_as_float(h << 16); }
  static __device__ __forceinline__ uint32_t down(float v) { return __float_as_uint(rnd(v)) >> 16; }
  static __device__ __forceinline__ float ld(const T* p, size_t c) { return up(p[c]); }
  static __device__ __forceinline__ void st(T* p, size_t c, float v) { p[c] = (T)down(v); }
  static __device__ __forceinline__ float2 ld2(const T* p, size_t c) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(p + c);
    return make_float2(up(w & 0xFFFFu), up(w >> 16));
  }
  static __device__ __forceinline__ void st2(T* p, size_t c, float2 v) {
    *reinterpret_cast<uint32_t*>(p + c) = down(v.x) | (down(v.y) << 16);
  }
  static __device__ __forceinline__ float4 ld4(const T* p, size_t c) {
    const uint2 w = *reinterpret_cast<const uint2*>(p + c);
    return make_float4(up(w.x & 0xFFFFu), up(w.x >> 16), up(w.y & 0xFFFFu), up(w.y >> 16));
  }
  static __device__ __forceinline__ void st4(T* p, size_t c, float4 v) {
    *reinterpret_cast<uint2*>(p + c) = make_uint2(down(v.x) | (down(v.y) << 16), down(v.z) | (down(v.w) << 16));
  }
};

// IEEE binary16, converted in software (round to nearest even, subnormals
// kept, overflow to infinity) so that no packed or truncating hardware
// conversion can be substituted for it.
__device__ __forceinline__ uint32_t f32_to_f16_bits(float f) {
  const uint32_t x = __float_as_uint(f);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return sign | (ax > 0x7F800000u ? 0x7E00u : 0x7C00u);
  if (ax >= 0x477FF000u) return sign | 0x7C00u;  // >= 65520 rounds to infinity
  if (ax < 0x38800000u) {                         // below 2^-14: step 2^-24
    const uint32_t m = (uint32_t)rintf(__uint_as_float(ax) * 16777216.0f);
    return sign | m;
  }
  const uint32_t r = ax + 0xFFFu + ((ax >> 13) & 1u);
  return sign | ((r >> 13) - (112u << 10));
}
__device__ __forceinline__ float f16_bits_to_f32(uint32_t h) {
  const uint32_t sign = (h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  if (e == 0) {
    const float v = (float)m * 5.9604644775390625e-08f;  // exact
    return sign ? -v : v;
  }
  if (e == 31) return __uint_as_float(sign | 0x7F800000u | (m << 13));
  return __uint_as_float(sign | ((e + 112u) << 23) | (m << 13));
}

template <> struct Store<ST_F16> {
  using T = uint16_t;
  static __device__ __forceinline__ float rnd(float v) { return f16_bits_to_f32(f32_to_f16_bits(v)); }
  static __device__ __forceinline__ float up(uint32_t h) { return f16_bits_to_f32(h); }
  static __device__ __forceinline__ uint32_t down(float v) { return f32_to_f16_bits(v); }
  static __device__ __forceinline__ float ld(const T* p, size_t c) { return up(p[c]); }
  static __device__ __forceinline__ void st(T* p, size_t c, float v) { p[c] = (T)down(v); }
  static __device__ __forceinline__ float2 ld2(const T* p, size_t c) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(p + c);
    return make_float2(up(w & 0xFFFFu), up(w >> 16));
  }
  static __device__ __forceinline__ void st2(T* p, size_t c, float2 v) {
    *reinterpret_cast<uint32_t*>(p + c) = down(v.x) | (down(v.y) << 16);
  }
  static __device__ __forceinline__ float4 ld4(const T* p, size_t c) {
    const uint2 w = *reinterpret_cast<const uint2*>(p + c);
    return make_float4(up(w.x & 0xFFFFu), up(w.x >> 16), up(w.y & 0xFFFFu), up(w.y >> 16));
  }
  static __device__ __forceinline__ void st4(T* p, size_t c, float4 v) {
    *reinterpret_cast<uint2*>(p + c) = make_uint2(down(v.x) | (down(v.y) << 16), down(v.z) | (down(v.w) << 16));
  }
};

}  // namespace fsp
