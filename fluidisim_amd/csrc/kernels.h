// Device code of the projection: fine operator, assembly, CG steps, V-cycle,
// correction and metrics. Included once, by fsp.hip.
//
// Fine-level arithmetic follows the reference's face order x-, x+, y-, y+,
// z-, z+ with FP contraction off, so the operator action, the right-hand
// side, the Jacobi diagonal, the correction and the divergence are the
// reference's bit for bit (code/water-core/src/delta3d_mobile.rs:126-343,
// delta3d.rs:747-823). Reductions accumulate in f64.
#pragma once
#include "fsp_internal.h"

namespace fsp {

// ---------------------------------------------------------------- indexing
__device__ __forceinline__ size_t cidx(const Fine& G, int i, int j, int k) {
  return ((size_t)k * G.ny + j) * G.nx + i;
}
__device__ __forceinline__ size_t fuidx(const Fine& G, int i, int j, int k) {
  return ((size_t)k * G.ny + j) * (G.nx + 1) + i;
}
__device__ __forceinline__ size_t fvidx(const Fine& G, int i, int j, int k) {
  return ((size_t)k * (G.ny + 1) + j) * G.nx + i;
}
__device__ __forceinline__ size_t fwidx(const Fine& G, int i, int j, int k) {
  return ((size_t)k * G.ny + j) * G.nx + i;
}

// Blocks are dispatched round-robin over the 8 XCDs. FSP_XCD_REMAP = 1
// remaps them so that each XCD walks a contiguous range of tiles (stencil
// neighbours in its own L2); measured on the 256x128x256 tank, the plain
// round-robin order is faster (the red-black pass runs at 5.2 TB/s instead
// of 3.1 TB/s), so the remap is off by default.
#ifndef FSP_XCD_REMAP
#define FSP_XCD_REMAP 0
#endif
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G) {
  if (FSP_XCD_REMAP && (G & 7u) == 0u) return (b & 7u) * (G >> 3) + (b >> 3);
  return b;
}
__device__ __forceinline__ unsigned logical_block() { return xcd_block(blockIdx.x, gridDim.x); }
#define FSP_FOR(idx, count)                                                         \
  for (size_t idx = (size_t)logical_block() * BLOCK + threadIdx.x; idx < (count); \
       idx += (size_t)gridDim.x * BLOCK)

__device__ __forceinline__ void decode(const Fine& G, size_t c, int& i, int& j, int& k) {
  const uint32_t cc = (uint32_t)c;
  const uint32_t nx = (uint32_t)G.nx, ny = (uint32_t)G.ny;
  const uint32_t t = cc / nx;
  i = (int)(cc - t * nx);
  k = (int)(t / ny);
  j = (int)(t - (uint32_t)k * ny);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Block-level reduction of two values into this block's slot of `partials`.
// The grid total is folded by k_finish, launched right after the producing
// kernel: no atomics and no device-wide fences on the hot path, and the
// totals do not depend on scheduling.
template <bool MAX>
__device__ __forceinline__ void block_partial(double a, double b, int slot, double* partials) {
  __shared__ double sa[BLOCK / 64], sb[BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  a = MAX ? wave_max(a) : wave_sum(a);
  b = MAX ? wave_max(b) : wave_sum(b);
  if (lane == 0) { sa[wid] = a; sb[wid] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = sa[0], tb = sb[0];
    for (int w = 1; w < nw; ++w) {
      ta = MAX ? fmax(ta, sa[w]) : ta + sa[w];
      tb = MAX ? fmax(tb, sb[w]) : tb + sb[w];
    }
    double* part = partials + (size_t)slot * MAX_PARTIALS * 2;
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tb;
  }
}

// Scalar updates of the conjugate gradient, done on the device by the last
// block of the kernel that produced the reduction (delta3d_mobile.rs:513-545).
__device__ void finish(int op, Ctl* ctl, double a, double b) {
  switch (op) {
    case OP_B2:
      ctl->b2 = a;
      if (!isfinite(a)) { ctl->nonfinite = 1; ctl->active = 0; }
      break;
    case OP_RR:
      ctl->actual = a;
      if (!isfinite(a)) { ctl->nonfinite = 1; ctl->active = 0; }
      break;
    case OP_DQ:
      ctl->dq = a;
      ctl->first_dir = 0;
      if (!(a > 0.0)) {
        ctl->active = 0;
        ctl->stop_dq = 1;
      } else {
        ctl->alpha = ctl->rz / a;
        ctl->dsel ^= 1;
      }
      break;
    case OP_RN:
      ctl->rn = a;
      if (!isfinite(a)) { ctl->nonfinite = 1; ctl->active = 0; }
      break;
    case OP_RN_TRUE:
      // the true residual replaces the recurrence; no progress since the
      // previous replacement ends the pass, so that the driver's round-off
      // floor checks run on it
      ctl->rn = a;
      if (!isfinite(a)) { ctl->nonfinite = 1; ctl->active = 0; break; }
      if (!(a < ctl->rr_replaced)) ctl->stalled = 1;
      ctl->rr_replaced = a;
      break;
    case OP_RZ:
      if (!isfinite(a) || (ctl->flexible && !isfinite(b))) { ctl->nonfinite = 1; ctl->active = 0; break; }
      // flexible CG (Polak-Ribiere): <z', r' - r> = -alpha <z', q> since r' = r - alpha q
      ctl->beta = ctl->flexible ? -(ctl->alpha * b) / ctl->rz : a / ctl->rz;
      ctl->rz = a;
      ctl->rr = ctl->rn;
      ctl->it += 1;
      ctl->active = ((ctl->rr > ctl->tol_b2 || ctl->rr > ctl->physical_target) && ctl->it < ctl->max_iters &&
                     ctl->it < ctl->pass_limit && !ctl->stalled) ? 1 : 0;
      break;
    case OP_RZ_PRIME:
      if (!isfinite(a)) { ctl->nonfinite = 1; ctl->active = 0; break; }
      ctl->rz = a;
      ctl->first_dir = 1;
      break;
    case OP_MEAN_B:
      ctl->mean_b = (ctl->pure_neumann && b > 0.0) ? a / b : 0.0;
      break;
    case OP_MEAN_P:
      ctl->mean_p = (ctl->pure_neumann && b > 0.0) ? a / b : 0.0;
      break;
    case OP_COUNT:
      ctl->wet_cells = (uint32_t)a;
      ctl->ghost_faces = (uint32_t)b;
      ctl->pure_neumann = (ctl->gauge && b == 0.0 && a > 0.0) ? 1 : 0;
      break;
    default:
      break;
  }
}

// The iteration in progress replaces the recurrence residual by the true
// one: every replace_every-th iteration of the CG loop (the direction and
// <r, z> recurrences then continue from the replaced r).
__device__ __forceinline__ bool replacing(const Ctl* ctl) {
  return ctl->active && ctl->replace_every > 0u && (ctl->it + 1u) % ctl->replace_every == 0u;
}

template <bool MAX>
__device__ __forceinline__ void reduce_finish(double a, double b, int op, Ctl* ctl, double* partials) {
  (void)ctl;
  block_partial<MAX>(a, b, op, partials);
}

// Folds the partials of `nblocks` blocks in block order and applies the
// finish operation. `gated`: 1, skipped like its producer once the loop
// stopped; 2, skipped like the gauge kernels unless the problem is pure
// Neumann.
// FIN_THREADS threads with eight loads each fold 8192 partials (the fine
// kernels' grid) in one round of memory latency.
constexpr int FIN_THREADS = 1024;

template <bool MAX>
__global__ void __launch_bounds__(FIN_THREADS) k_finish(int op, unsigned nblocks, int gated, Ctl* ctl,
                                                        double* partials) {
  if ((gated == 1 && !ctl->active) || (gated == 2 && !ctl->pure_neumann) || (gated == 3 && !replacing(ctl))) return;
  __shared__ double sa[FIN_THREADS / 64], sb[FIN_THREADS / 64];
  const double2* part = reinterpret_cast<const double2*>(partials + (size_t)op * MAX_PARTIALS * 2);
  // eight independent loads in flight per step, folded in a fixed order
  double a = 0.0, b = 0.0;
  for (unsigned base = threadIdx.x; base < nblocks; base += 8 * FIN_THREADS) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const unsigned q = base + (unsigned)u * FIN_THREADS;
      v[u] = q < nblocks ? part[q] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a = MAX ? fmax(a, v[u].x) : a + v[u].x;
      b = MAX ? fmax(b, v[u].y) : b + v[u].y;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  a = MAX ? wave_max(a) : wave_sum(a);
  b = MAX ? wave_max(b) : wave_sum(b);
  if (lane == 0) { sa[wid] = a; sb[wid] = b; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  a = sa[0];
  b = sb[0];
  for (int w = 1; w < FIN_THREADS / 64; ++w) {
    a = MAX ? fmax(a, sa[w]) : a + sa[w];
    b = MAX ? fmax(b, sb[w]) : b + sb[w];
  }
  ctl->red[op][0] = a;
  ctl->red[op][1] = b;
  finish(op, ctl, a, b);
}

// ---------------------------------------------------------------- fine geometry
__device__ __forceinline__ bool solid_at(const Fine& G, int i, int j, int k) { return G.solid[cidx(G, i, j, k)] != 0; }
__device__ __forceinline__ float zc_of(const Fine& G, int k) {
#pragma clang fp contract(off)
  return ((float)k + 0.5f) * G.dx;
}
// delta3d_mobile.rs:97-103; with a level set, the water label
// (apic3d.rs:688-692).
__device__ __forceinline__ bool wet_at(const Fine& G, int i, int j, int k) {
  if (G.phi) return G.water[cidx(G, i, j, k)] != 0;
  if (solid_at(G, i, j, k)) return false;
  return zc_of(G, k) < G.eta[(size_t)j * G.nx + i];
}
// The ghost fraction between water cell (i, j, k) and air cell (x, y, z)
// from the level set (apic3d.rs:902-907): 1/theta, value 0.
__device__ __forceinline__ void ghost_phi(const Fine& G, int i, int j, int k, int x, int y, int z, float& it,
                                          float& val) {
#pragma clang fp contract(off)
  const float fc = G.phi[cidx(G, i, j, k)], fa = G.phi[cidx(G, x, y, z)];
  it = 1.0f / fmaxf(fc / (fc - fa), G.theta_min);
  val = 0.f;
}
// delta3d_mobile.rs:126-134; coupled, theta from the total surface and the
// background's column term added to the value (:130-132)
__device__ __forceinline__ void ghost_up(const Fine& G, int i, int j, int k, float& it, float& val) {
#pragma clang fp contract(off)
  if (G.phi) { ghost_phi(G, i, j, k, i, j, k + 1, it, val); return; }
  const size_t col = (size_t)j * G.nx + i;
  const float h = G.eta[col];
  const float theta = fmaxf((h - zc_of(G, k)) / G.dx, 1e-3f);
  it = 1.0f / theta;
  if (G.homogeneous) { val = 0.f; return; }
  if (!G.bg_up) { val = G.rho * G.g * ((h - G.rest) - G.eta_rt[col]); return; }
  val = G.rho * G.g * ((G.eta_val[col] - G.rest) - G.eta_rt[col]);
  val += G.bg_up[col];
}
// delta3d_mobile.rs:136-147: wet cell (i, j, k), dry neighbour (x, y, k);
// coupled, the background's term of the face between them (:142-145)
__device__ __forceinline__ void ghost_side(const Fine& G, int i, int j, int k, int x, int y, float& it, float& val) {
#pragma clang fp contract(off)
  if (G.phi) { ghost_phi(G, i, j, k, x, y, k, it, val); return; }
  const float zc = zc_of(G, k);
  const float h = G.eta[(size_t)j * G.nx + i];
  const float theta = fmaxf((h - zc) / (h - G.eta[(size_t)y * G.nx + x]), 1e-3f);
  it = 1.0f / theta;
  if (G.homogeneous) { val = 0.f; return; }
  val = G.rho * G.g * (zc - G.rest);
  if (G.bg_up)
    val += x != i ? G.bg_x[fuidx(G, i > x ? i : x, j, k)] : G.bg_y[fvidx(G, i, j > y ? j : y, k)];
}

// Aperture of a face from its code in the u8 face arrays (0 closed, 1 open,
// 2 the fraction held in au/av/aw); axis 0 u, 1 v, 2 w.
__device__ __forceinline__ float aperture(const Fine& G, int axis, size_t f, uint8_t code) {
  if (code != 2) return (float)code;
  return axis == 0 ? G.au[f] : axis == 1 ? G.av[f] : G.aw[f];
}

// The reference's mobile_row (delta3d_mobile.rs:152-188): type 1 = wet
// neighbour, 2 = ghost face, 0 = nothing; a = the face's aperture.
struct RowG {
  uint8_t type[6];
  float it[6];
  float val[6];
  float a[6];
};

__device__ void full_row(const Fine& G, int i, int j, int k, RowG& R) {
#pragma unroll
  for (int f = 0; f < 6; ++f) { R.type[f] = 0; R.it[f] = 0.f; R.val[f] = 0.f; R.a[f] = 0.f; }
  const bool has[4] = {i > 0, i + 1 < G.nx, j > 0, j + 1 < G.ny};
  const int xs[4] = {i - 1, i + 1, i, i};
  const int ys[4] = {j, j, j - 1, j + 1};
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (!has[f]) continue;
    const size_t face = f == 0 ? fuidx(G, i, j, k) : f == 1 ? fuidx(G, i + 1, j, k)
                      : f == 2 ? fvidx(G, i, j, k) : fvidx(G, i, j + 1, k);
    const uint8_t a = f < 2 ? G.ou[face] : G.ov[face];
    const int x = xs[f], y = ys[f];
    if (a == 0 || solid_at(G, x, y, k)) continue;
    R.a[f] = aperture(G, f < 2 ? 0 : 1, face, a);
    if (wet_at(G, x, y, k)) {
      R.type[f] = 1;
    } else {
      R.type[f] = 2;
      ghost_side(G, i, j, k, x, y, R.it[f], R.val[f]);
    }
  }
  const size_t fd = fwidx(G, i, j, k), fu = fwidx(G, i, j, k + 1);
  if (k > 0 && G.ow[fd] && !solid_at(G, i, j, k - 1)) {
    R.a[4] = aperture(G, 2, fd, G.ow[fd]);
    if (!G.phi || wet_at(G, i, j, k - 1)) {
      R.type[4] = 1;
    } else {
      // a level set can put air below water
      R.type[4] = 2;
      ghost_phi(G, i, j, k, i, j, k - 1, R.it[4], R.val[4]);
    }
  }
  if (G.ow[fu] && !(G.phi && k + 1 < G.nz && solid_at(G, i, j, k + 1))) {
    R.a[5] = aperture(G, 2, fu, G.ow[fu]);
    if (k + 1 < G.nz && wet_at(G, i, j, k + 1)) {
      R.type[5] = 1;
    } else {
      R.type[5] = 2;
      ghost_up(G, i, j, k, R.it[5], R.val[5]);
    }
  }
}

// ---------------------------------------------------------------- fine operator
template <class T>
__device__ __forceinline__ float ldv(const T* p, size_t i) { return (float)p[i]; }

// 1/theta of the six faces of a ghost row, from the per-column table built
// once per projection (k_ghost_table); false if the row is outside the
// column's table, in which case the caller derives it from the geometry.
__device__ __forceinline__ bool ghost_its(const Fine& G, int i, int j, int k, float it[6]) {
  const size_t col = (size_t)j * G.nx + i;
  const int slot = k - (int)G.gk0[col];
  if (slot < 0 || slot >= G_SLOTS) return false;
  const float* e = G.gtab + (col * G_SLOTS + slot) * 8;
  const float4 a = *reinterpret_cast<const float4*>(e);
  const float2 b = *reinterpret_cast<const float2*>(e + 4);
  it[0] = a.x; it[1] = a.y; it[2] = a.z; it[3] = a.w; it[4] = b.x; it[5] = b.y;
  return true;
}

// Per-face coefficients of a row with ghost faces or fractional apertures:
// type 1 neighbour, 2 ghost, with the face's aperture a (1 with binary
// apertures). With fractional apertures every such row comes from the
// geometry (the ghost table holds binary rows only).
// GEN false: the table only (the host runs these forms when every ghost row
// of a binary column geometry lies in its column's table; the geometry's
// arithmetic, which sets a kernel's register count, is then not compiled).
template <bool GEN = true>
__device__ __forceinline__ void ghost_row(const Fine& G, int i, int j, int k, uint8_t m, uint8_t type[6], float it[6],
                                          float a[6]) {
  if (!G.frac && !G.phi && ghost_its(G, i, j, k, it)) {
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      type[f] = (m & (1u << f)) ? 1 : (it[f] > 0.f ? 2 : 0);
      a[f] = 1.0f;
    }
    return;
  }
  if (!GEN) {
#pragma unroll
    for (int f = 0; f < 6; ++f) { type[f] = 0; it[f] = 0.f; a[f] = 0.f; }
    return;
  }
  RowG R;
  full_row(G, i, j, k, R);
#pragma unroll
  for (int f = 0; f < 6; ++f) { type[f] = R.type[f]; it[f] = R.it[f]; a[f] = R.a[f]; }
}

// (A x)_c in accumulator precision A from storage T, in the reference's face
// order (apply_mobile3, delta3d_mobile.rs:190-213), and prec as rhs_mobile3
// computes it (:227-237).
template <class A, class T>
__device__ __forceinline__ A apply_row_t(const Fine& G, int i, int j, int k, size_t c, uint8_t m, const T* x,
                                         float* prec_out = nullptr) {
#pragma clang fp contract(off)
  if (prec_out) *prec_out = 0.f;
  if (!(m & M_WET)) return (A)0;
  const size_t plane = (size_t)G.nx * G.ny;
  const A pc = (A)x[c];
  A acc = (A)0;
  if (!(m & M_GHOST)) {
    if (m & M_XM) acc += pc - (A)x[c - 1];
    if (m & M_XP) acc += pc - (A)x[c + 1];
    if (m & M_YM) acc += pc - (A)x[c - G.nx];
    if (m & M_YP) acc += pc - (A)x[c + G.nx];
    if (m & M_ZM) acc += pc - (A)x[c - plane];
    if (m & M_ZP) acc += pc - (A)x[c + plane];
    if (prec_out) *prec_out = G.prec_tab[__popc(m & 0x3Fu)];
  } else {
    uint8_t type[6];
    float it[6], a[6];
    ghost_row(G, i, j, k, m, type, it, a);
    const size_t nb[6] = {c - 1, c + 1, c - G.nx, c + G.nx, c - plane, c + plane};
    float diag = 0.f;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      if (type[f] == 1) { acc += (A)a[f] * (pc - (A)x[nb[f]]); diag += a[f]; }
      else if (type[f] == 2) { acc += (A)a[f] * pc * (A)it[f]; diag += a[f] * it[f]; }
    }
    if (prec_out) *prec_out = diag > 0.f ? 1.0f / (diag * G.inv) : 0.f;
  }
  return acc * (A)G.inv;
}

template <class T>
__device__ __forceinline__ float apply_prec(const Fine& G, int i, int j, int k, size_t c, uint8_t m, const T* x,
                                            float& prec) {
  return apply_row_t<float>(G, i, j, k, c, m, x, &prec);
}

template <class T>
__device__ __forceinline__ float apply_row(const Fine& G, int i, int j, int k, size_t c, uint8_t m, const T* x) {
  return apply_row_t<float>(G, i, j, k, c, m, x);
}

// The Jacobi preconditioner alone (rhs_mobile3's prec, delta3d_mobile.rs:227-237).
__device__ __forceinline__ float row_prec(const Fine& G, int i, int j, int k, uint8_t m) {
#pragma clang fp contract(off)
  if (!(m & M_WET)) return 0.f;
  if (!(m & M_GHOST)) return G.prec_tab[__popc(m & 0x3Fu)];
  uint8_t type[6];
  float it[6], a[6];
  ghost_row(G, i, j, k, m, type, it, a);
  float diag = 0.f;
  for (int f = 0; f < 6; ++f) {
    if (type[f] == 1) diag += a[f];
    else if (type[f] == 2) diag += a[f] * it[f];
  }
  return diag > 0.f ? 1.0f / (diag * G.inv) : 0.f;
}

// prec == 0 exactly on dry cells and on wet cells without any face.
__device__ __forceinline__ bool prec_positive(uint8_t m) { return (m & M_WET) && (m & (0x3F | M_GHOST)); }

// row_prec of every cell, for kernels that need the factor of neighbouring
// rows (the first red-black pass recomputed beside the second).
__global__ void __launch_bounds__(BLOCK) k_prec(Fine G, float* out) {
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    out[c] = row_prec(G, i, j, k, G.mask[c]);
  }
}

// delta3d.rs:747-763 (plain) and :806-823 (cut with fixed solids).
__device__ __forceinline__ float divergence_at(const Fine& G, const float* u, const float* v, const float* w, int i,
                                               int j, int k, size_t c) {
#pragma clang fp contract(off)
  if (G.cut && G.solid[c]) return 0.f;
  const size_t a0 = fuidx(G, i, j, k), a1 = fuidx(G, i + 1, j, k);
  const size_t b0 = fvidx(G, i, j, k), b1 = fvidx(G, i, j + 1, k);
  const size_t c0 = fwidx(G, i, j, k), c1 = fwidx(G, i, j, k + 1);
  float fl = u[a0], fr = u[a1], ff = v[b0], fk = v[b1], fb = w[c0], ft = w[c1];
  if (G.cut) {
    fl = aperture(G, 0, a0, G.ou[a0]) * fl; fr = aperture(G, 0, a1, G.ou[a1]) * fr;
    ff = aperture(G, 1, b0, G.ov[b0]) * ff; fk = aperture(G, 1, b1, G.ov[b1]) * fk;
    fb = aperture(G, 2, c0, G.ow[c0]) * fb; ft = aperture(G, 2, c1, G.ow[c1]) * ft;
  }
  return ((fr - fl + ft - fb) + (fk - ff)) / G.dx;
}

// ---------------------------------------------------------------- geometry kernels
// Row codes (the face types of full_row, packed): bit f set for a wet
// neighbour across face f, M_WET for a wet cell, M_GHOST when a face is a
// ghost face. Every operand of a cell (its six apertures, seven solid flags,
// five column heights) is loaded at once, indices clamped into the grid, and
// the types are decided after; full_row gives the same types.
// Row code of cell (i, j, k) from its solid flag s_c, its apertures ap
// (x-, x+, y-, y+, z-, z+), the solid flags sn of its neighbours (same order,
// clamped into the grid) and the surface heights of its column (h) and of the
// four side columns (hn, clamped); counts wet rows and ghost faces.
__device__ __forceinline__ uint8_t row_code(const Fine& G, int i, int j, int k, uint8_t s_c, const uint8_t sn[6],
                                            const uint8_t ap[6], float h, const float hn[4], double& wet,
                                            double& ghosts) {
  const float zc = zc_of(G, k);
  uint8_t m = 0;
  if (!s_c && zc < h) {
    m = M_WET;
    bool fractional = false;  // a face of the row with a fractional aperture
    const bool has[4] = {i > 0, i + 1 < G.nx, j > 0, j + 1 < G.ny};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (!has[f] || ap[f] == 0 || sn[f]) continue;
      fractional |= ap[f] == 2;
      if (zc < hn[f]) m |= (uint8_t)(1u << f);  // wet neighbour (not solid, below its surface)
      else { m |= M_GHOST; ghosts += 1.0; }
    }
    if (k > 0 && ap[4] && !sn[4]) {
      m |= M_ZM;
      fractional |= ap[4] == 2;
    }
    if (ap[5]) {
      fractional |= ap[5] == 2;
      if (k + 1 < G.nz && !sn[5] && zc_of(G, k + 1) < h) m |= M_ZP;
      else { m |= M_GHOST; ghosts += 1.0; }
    }
    // rows with a fractional aperture take the general row (full_row) as
    // free-surface rows do
    if (fractional) m |= M_GHOST;
    wet += 1.0;
  }
  return m;
}

__global__ void __launch_bounds__(BLOCK) k_mask(Fine G, uint8_t* mask, Ctl* ctl, double* partials) {
  double wet = 0.0, ghosts = 0.0;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    const int im = i > 0 ? i - 1 : i, ip = i + 1 < nx ? i + 1 : i;
    const int jm = j > 0 ? j - 1 : j, jp = j + 1 < ny ? j + 1 : j;
    const int km = k > 0 ? k - 1 : k, kp = k + 1 < nz ? k + 1 : k;
    // apertures: x-, x+, y-, y+, z- (face k), z+ (face k + 1)
    const uint8_t a0 = G.ou[fuidx(G, i, j, k)], a1 = G.ou[fuidx(G, i + 1, j, k)];
    const uint8_t a2 = G.ov[fvidx(G, i, j, k)], a3 = G.ov[fvidx(G, i, j + 1, k)];
    const uint8_t a4 = G.ow[fwidx(G, i, j, k)], a5 = G.ow[fwidx(G, i, j, k + 1)];
    const uint8_t s_c = G.solid[cidx(G, i, j, k)];
    const uint8_t s0 = G.solid[cidx(G, im, j, k)], s1 = G.solid[cidx(G, ip, j, k)];
    const uint8_t s2 = G.solid[cidx(G, i, jm, k)], s3 = G.solid[cidx(G, i, jp, k)];
    const uint8_t s4 = G.solid[cidx(G, i, j, km)], s5 = G.solid[cidx(G, i, j, kp)];
    const float h = G.eta[(size_t)j * nx + i];
    const float h0 = G.eta[(size_t)j * nx + im], h1 = G.eta[(size_t)j * nx + ip];
    const float h2 = G.eta[(size_t)jm * nx + i], h3 = G.eta[(size_t)jp * nx + i];
    const uint8_t ap[6] = {a0, a1, a2, a3, a4, a5}, sn[6] = {s0, s1, s2, s3, s4, s5};
    const float hn[4] = {h0, h1, h2, h3};
    mask[c] = row_code(G, i, j, k, s_c, sn, ap, h, hn, wet, ghosts);
  }
  reduce_finish<false>(wet, ghosts, OP_COUNT, ctl, partials);
}

// Row codes of a level-set geometry (the APIC-3D projection's labels,
// apic3d.rs:912-935): bit f for a water neighbour across an open face f,
// M_GHOST for an air neighbour (either side, below or above); solid
// neighbours and closed faces add nothing. Counts wet rows and ghost faces.
__global__ void __launch_bounds__(BLOCK) k_mask_ls(Fine G, uint8_t* mask, Ctl* ctl, double* partials) {
  double wet = 0.0, ghosts = 0.0;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    uint8_t m = 0;
    if (G.water[c]) {
      m = M_WET;
      const bool has[6] = {i > 0, i + 1 < nx, j > 0, j + 1 < ny, k > 0, k + 1 < nz};
      const size_t nb[6] = {c - 1, c + 1, c - nx, c + nx, c - (size_t)nx * ny, c + (size_t)nx * ny};
      const uint8_t ap[6] = {G.ou[fuidx(G, i, j, k)], G.ou[fuidx(G, i + 1, j, k)], G.ov[fvidx(G, i, j, k)],
                             G.ov[fvidx(G, i, j + 1, k)], G.ow[fwidx(G, i, j, k)], G.ow[fwidx(G, i, j, k + 1)]};
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        if (!has[f] || ap[f] == 0 || G.solid[nb[f]]) continue;
        if (G.water[nb[f]]) m |= (uint8_t)(1u << f);
        else { m |= M_GHOST; ghosts += 1.0; }
      }
      wet += 1.0;
    }
    mask[c] = m;
  }
  reduce_finish<false>(wet, ghosts, OP_COUNT, ctl, partials);
}

// Per-column table of the ghost rows' 1/theta (six faces, 0 where the face
// is not a ghost), for the G_SLOTS rows from the column's lowest ghost row:
// the operator reads it instead of re-deriving the surface geometry. Rows
// past the table fall back to full_row.
__global__ void __launch_bounds__(BLOCK) k_ghost_table(Fine G, uint16_t* gk0, float* gtab, Ctl* ctl) {
  const size_t cols = (size_t)G.nx * G.ny;
  const size_t plane = cols;
  constexpr int SCAN = 16;  // rows per round of independent loads
  FSP_FOR(col, cols) {
    const int i = (int)(col % (size_t)G.nx), j = (int)(col / (size_t)G.nx);
    // Ghost rows lie between the lowest surface around the column and its
    // own top wet row: an up ghost needs the row above dry (zc(k+1) >= h),
    // a side ghost a dry neighbour at the same height (zc(k) >= h_n). Rows
    // outside, less a margin of two rows for rounding, are not scanned.
    const float h = G.eta[col];
    float hn = h;
    if (i > 0) hn = fminf(hn, G.eta[col - 1]);
    if (i + 1 < G.nx) hn = fminf(hn, G.eta[col + 1]);
    if (j > 0) hn = fminf(hn, G.eta[col - G.nx]);
    if (j + 1 < G.ny) hn = fminf(hn, G.eta[col + G.nx]);
    const float lo = hn / G.dx - 3.0f, hi = h / G.dx + 2.0f;
    // (the top row stays in range: the top of the domain is a ghost face)
    const int kstart = lo > 0.f ? (lo < (float)(G.nz - 1) ? (int)lo : G.nz - 1) : 0;
    const int kend = hi > 0.f ? (hi < (float)(G.nz - 1) ? (int)hi + 1 : G.nz) : 0;  // exclusive
    // the lowest ghost row, SCAN rows per round
    int k0 = -1;
    for (int kb = kstart; kb < kend && k0 < 0; kb += SCAN) {
      uint8_t mm[SCAN];
#pragma unroll
      for (int u = 0; u < SCAN; ++u) mm[u] = kb + u < kend ? G.mask[(size_t)(kb + u) * plane + col] : 0;
#pragma unroll
      for (int u = 0; u < SCAN; ++u)
        if (k0 < 0 && (mm[u] & M_GHOST)) k0 = kb + u;
    }
    gk0[col] = (uint16_t)(k0 < 0 ? 0xFFFF : k0);
    if (k0 < 0) continue;
    uint8_t ms[G_SLOTS];
#pragma unroll
    for (int s = 0; s < G_SLOTS; ++s) ms[s] = k0 + s < G.nz ? G.mask[(size_t)(k0 + s) * plane + col] : 0;
    for (int s = 0; s < G_SLOTS && k0 + s < G.nz; ++s) {
      const int k = k0 + s;
      float* e = gtab + (col * G_SLOTS + s) * 8;
      float it[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (ms[s] & M_GHOST) {
        RowG R;
        full_row(G, i, j, k, R);
        for (int f = 0; f < 6; ++f) it[f] = R.type[f] == 2 ? R.it[f] : 0.f;
      }
      for (int f = 0; f < 6; ++f) e[f] = it[f];
      e[6] = 0.f;
      e[7] = 0.f;
    }
    // a ghost row above the table: the fine kernels take their general form
    bool miss = false;
    for (int kb = k0 + G_SLOTS; kb < kend && !miss; kb += SCAN) {
      uint8_t mm[SCAN];
#pragma unroll
      for (int u = 0; u < SCAN; ++u) mm[u] = kb + u < kend ? G.mask[(size_t)(kb + u) * plane + col] : 0;
#pragma unroll
      for (int u = 0; u < SCAN; ++u) miss = miss || (mm[u] & M_GHOST);
    }
    if (miss) ctl->table_miss = 1;
  }
}

// b = scale*div + ghost terms; prec (optional). div from faces or given.
__global__ void __launch_bounds__(BLOCK) k_assemble(Fine G, const float* us, const float* vs, const float* ws,
                                                    const float* div_in, float scale, float* b, float* prec_out,
                                                    Ctl* ctl, double* partials) {
#pragma clang fp contract(off)
  double s2 = 0.0, unused = 0.0;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    const uint8_t m = G.mask[c];
    float bb = 0.f, pr = 0.f;
    if (m & M_WET) {
      const float div = div_in ? div_in[c] : divergence_at(G, us, vs, ws, i, j, k, c);
      bb = scale * div;
      float diag = 0.f;
      if (!(m & M_GHOST)) {
        for (int f = 0; f < 6; ++f) if (m & (1u << f)) diag += 1.0f;
      } else {
        RowG R;
        full_row(G, i, j, k, R);
        for (int f = 0; f < 6; ++f) {
          if (R.type[f] == 1) diag += R.a[f];
          else if (R.type[f] == 2) {
            diag += R.a[f] * R.it[f];
            bb += R.a[f] * R.val[f] * R.it[f] * G.inv;
          }
        }
      }
      pr = diag > 0.f ? 1.0f / (diag * G.inv) : 0.f;
    }
    b[c] = bb;
    if (prec_out) prec_out[c] = pr;
    s2 += (double)bb * (double)bb;
  }
  reduce_finish<false>(s2, unused, OP_B2, ctl, partials);
}

// Sum of x and count over rows with prec > 0 (gauge of a pure-Neumann problem).
// Skipped (with its fold, k_finish gated == 2) unless the problem is pure Neumann.
template <class T>
__global__ void __launch_bounds__(BLOCK) k_sum_active(Fine G, const T* x, int op, Ctl* ctl, double* partials) {
  if (!ctl->pure_neumann) return;
  double s = 0.0, n = 0.0;
  FSP_FOR(c, G.n) {
    if (prec_positive(G.mask[c])) { s += (double)x[c]; n += 1.0; }
  }
  reduce_finish<false>(s, n, op, ctl, partials);
}

// x -= mean on active rows when the problem is pure Neumann; for b also
// recompute |b|^2 (otherwise x and |b|^2 stand: skipped, with its fold).
template <class T>
__global__ void __launch_bounds__(BLOCK) k_sub_mean(Fine G, T* x, int which, int op, Ctl* ctl, double* partials) {
  const int pure = ctl->pure_neumann;
  if (!pure) return;
  const T mean = (T)(which == 0 ? ctl->mean_b : ctl->mean_p);
  double s2 = 0.0, unused = 0.0;
  FSP_FOR(c, G.n) {
    T val = x[c];
    if (pure && prec_positive(G.mask[c])) { val = val - mean; x[c] = val; }
    s2 += (double)val * (double)val;
  }
  if (op >= 0) reduce_finish<false>(s2, unused, op, ctl, partials);
}

// Warm start (delta3d_mobile.rs:486-496): p from the caller, 0 on dry cells;
// cold start: p = 0.
__global__ void __launch_bounds__(BLOCK) k_load_p(Fine G, const float* pin, double* p, int warm) {
  FSP_FOR(c, G.n) p[c] = (warm && (G.mask[c] & M_WET)) ? (double)pin[c] : 0.0;
}
__global__ void __launch_bounds__(BLOCK) k_store_p(size_t n, const double* p, float* pout) {
  FSP_FOR(c, n) pout[c] = (float)p[c];
}

__global__ void __launch_bounds__(BLOCK) k_fill(float* x, size_t n, float value) {
  FSP_FOR(c, n) x[c] = value;
}
// p += base in f32 (refine_mobile3's sum, delta3d_coupling.rs:244-246)
__global__ void __launch_bounds__(BLOCK) k_add_base(float* p, const float* base, size_t n) {
#pragma clang fp contract(off)
  FSP_FOR(c, n) p[c] = p[c] + base[c];
}

// r = b - A p, |r|^2 (delta3d_mobile.rs:497-503, :547-553), with the
// pressure held and the residual formed in f64; r is stored in f32.
// gated: only in an iteration that replaces the residual.
__global__ void __launch_bounds__(BLOCK) k_residual(Fine G, const double* p, const float* b, float* r, int op, Ctl* ctl,
                                                    double* partials, int gated) {
#pragma clang fp contract(off)
  if (gated && !replacing(ctl)) return;
  double s = 0.0, unused = 0.0;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    const double rc = (double)b[c] - apply_row_t<double>(G, i, j, k, c, G.mask[c], p);
    r[c] = (float)rc;
    s += rc * rc;
  }
  if (op >= 0) reduce_finish<false>(s, unused, op, ctl, partials);
}

// out = A p (the trial hook, delta3d_mobile.rs:936-949)
__global__ void __launch_bounds__(BLOCK) k_apply(Fine G, const float* p, float* out) {
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    out[c] = apply_row(G, i, j, k, c, G.mask[c], p);
  }
}

// ---------------------------------------------------------------- CG steps
// d = z (first) or z + beta d; q = A d; <d, q>.
__global__ void __launch_bounds__(BLOCK) k_spmv_dir(Fine G, const float* z, float* d0, float* d1, float* q, Ctl* ctl,
                                                    double* partials) {
  if (!ctl->active) return;
  const int sel = ctl->dsel, first = ctl->first_dir;
  const float beta = (float)ctl->beta;
  const float* dold = sel ? d1 : d0;
  float* dnew = sel ? d0 : d1;
  const size_t plane = (size_t)G.nx * G.ny;
  double s = 0.0, unused = 0.0;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    const uint8_t m = G.mask[c];
    auto dv = [&](size_t n) { return first ? z[n] : z[n] + beta * dold[n]; };
    const float dc = dv(c);
    dnew[c] = dc;
    float qc = 0.f;
    if (m & M_WET) {
#pragma clang fp contract(off)
      float acc = 0.f;
      if (!(m & M_GHOST)) {
        if (m & M_XM) acc += dc - dv(c - 1);
        if (m & M_XP) acc += dc - dv(c + 1);
        if (m & M_YM) acc += dc - dv(c - G.nx);
        if (m & M_YP) acc += dc - dv(c + G.nx);
        if (m & M_ZM) acc += dc - dv(c - plane);
        if (m & M_ZP) acc += dc - dv(c + plane);
      } else {
        RowG R;
        full_row(G, i, j, k, R);
        const size_t nb[6] = {c - 1, c + 1, c - G.nx, c + G.nx, c - plane, c + plane};
        for (int f = 0; f < 6; ++f) {
          if (R.type[f] == 1) acc += R.a[f] * (dc - dv(nb[f]));
          else if (R.type[f] == 2) acc += R.a[f] * dc * R.it[f];
        }
      }
      qc = acc * G.inv;
    }
    q[c] = qc;
    s += (double)dc * (double)qc;
  }
  reduce_finish<false>(s, unused, OP_DQ, ctl, partials);
}

// p += alpha d; r -= alpha q; |r|^2 (delta3d_mobile.rs:527-532)
__global__ void __launch_bounds__(BLOCK) k_update(size_t n, double* p, float* r, const float* d0, const float* d1,
                                                  const float* q, Ctl* ctl, double* partials) {
  if (!ctl->active) return;
  const double alpha64 = ctl->alpha;
  const float alpha = (float)alpha64;
  const float* d = ctl->dsel ? d1 : d0;
  double s = 0.0, unused = 0.0;
  FSP_FOR(c, n) {
    p[c] += alpha64 * (double)d[c];
    const float rc = r[c] - alpha * q[c];
    r[c] = rc;
    s += (double)rc * (double)rc;
  }
  reduce_finish<false>(s, unused, OP_RN, ctl, partials);
}

// <a, b> into a slot, gated on the loop flag when `gated`.
__global__ void __launch_bounds__(BLOCK) k_dot(size_t n, const float* a, const float* b, int op, int gated, Ctl* ctl,
                                               double* partials) {
  if (gated && !ctl->active) return;
  double s = 0.0, unused = 0.0;
  FSP_FOR(c, n) s += (double)a[c] * (double)b[c];
  reduce_finish<false>(s, unused, op, ctl, partials);
}

// <r, z> and <q, z> in one pass (flexible CG).
__global__ void __launch_bounds__(BLOCK) k_dot_rz_qz(size_t n, const float* r, const float* z, const float* q, int op,
                                                     Ctl* ctl, double* partials) {
  if (!ctl->active) return;
  double a = 0.0, b = 0.0;
  FSP_FOR(c, n) {
    const double zc = (double)z[c];
    a += (double)r[c] * zc;
    b += (double)q[c] * zc;
  }
  reduce_finish<false>(a, b, op, ctl, partials);
}

// Jacobi preconditioner z = prec * r with <r, z> (delta3d_mobile.rs:435-461).
__global__ void __launch_bounds__(BLOCK) k_jacobi_prec(Fine G, const float* r, float* z, int op, Ctl* ctl,
                                                       double* partials) {
  if (!ctl->active) return;
  double s = 0.0, unused = 0.0;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    const float prec = row_prec(G, i, j, k, G.mask[c]);
    const float zc = prec * r[c];
    z[c] = zc;
    s += (double)r[c] * (double)zc;
  }
  reduce_finish<false>(s, unused, op, ctl, partials);
}

// ---------------------------------------------------------------- fine smoothing
// Damped Jacobi (delta3d_multigrid.rs:398-405): zout = zin + w prec (r - A zin);
// zin == nullptr means zin = 0.
__global__ void __launch_bounds__(BLOCK) k_fjacobi(Fine G, const float* zin, float* zout, const float* r, Ctl* ctl) {
#pragma clang fp contract(off)
  if (!ctl->active) return;
  const float omega = 6.0f / 7.0f;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    const uint8_t m = G.mask[c];
    float prec, az = 0.f, zc = 0.f;
    if (zin) {
      az = apply_prec(G, i, j, k, c, m, zin, prec);
      zc = zin[c];
    } else {
      prec = row_prec(G, i, j, k, m);
    }
    if (prec > 0.f) zc += omega * prec * (r[c] - az);
    zout[c] = zc;
  }
}

// One colour of a red-black Gauss-Seidel sweep, in place. `first`: z starts
// at zero (colour 0 writes prec*r, the other colour 0).
__global__ void __launch_bounds__(BLOCK) k_fgs(Fine G, float* z, const float* r, int colour, int first, Ctl* ctl) {
#pragma clang fp contract(off)
  if (!ctl->active) return;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    const int col = (i + j + k) & 1;
    if (first) {
      if (col != colour) { z[c] = 0.f; continue; }
      const float prec = row_prec(G, i, j, k, G.mask[c]);
      z[c] = prec > 0.f ? prec * r[c] : 0.f;
      continue;
    }
    if (col != colour) continue;
    float prec;
    const float az = apply_prec(G, i, j, k, c, G.mask[c], z, prec);
    if (prec > 0.f) z[c] = z[c] + prec * (r[c] - az);
  }
}

// ---------------------------------------------------------------- coarse levels
__device__ __forceinline__ size_t lidx(const Level& L, int i, int j, int k) { return ((size_t)k * L.ny + j) * L.nx + i; }
__device__ __forceinline__ void ldecode(const Level& L, size_t c, int& i, int& j, int& k) {
  const uint32_t cc = (uint32_t)c, nx = (uint32_t)L.nx, ny = (uint32_t)L.ny;
  const uint32_t t = cc / nx;
  i = (int)(cc - t * nx);
  k = (int)(t / ny);
  j = (int)(t - (uint32_t)k * ny);
}

// row_sums (delta3d_multigrid.rs:216-255); x == nullptr gives the diagonal.
// x is held in storage ST.
template <int ST = ST_F32>
__device__ __forceinline__ float coarse_row(const Level& L, int i, int j, int k, size_t c,
                                            const typename Store<ST>::T* x) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (L.kind[c] <= 0) return 0.f;
  const float pc = x ? SV::ld(x, c) : 0.f;
  float acc = 0.f;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t plane = (size_t)nx * ny;
  const bool gal = L.galerkin != 0;
  auto face = [&](float a, bool has, size_t n, bool top) {
    if (a == 0.f) return;
    if (has) {
      const int8_t kn = L.kind[n];
      if (kn > 0) acc += x ? a * (pc - SV::ld(x, n)) : a;
      else if (!gal && kn < 0) acc += 2.0f * a * (x ? pc : 1.0f);
    } else if (!gal && top) {
      acc += 2.0f * a * (x ? pc : 1.0f);
    }
  };
  face(L.ou[((size_t)k * ny + j) * (nx + 1) + i], i > 0, c - 1, false);
  face(L.ou[((size_t)k * ny + j) * (nx + 1) + i + 1], i + 1 < nx, c + 1, false);
  face(L.ov[((size_t)k * (ny + 1) + j) * nx + i], j > 0, c - nx, false);
  face(L.ov[((size_t)k * (ny + 1) + j + 1) * nx + i], j + 1 < ny, c + nx, false);
  face(L.ow[((size_t)k * ny + j) * nx + i], k > 0, c - plane, false);
  face(L.ow[((size_t)(k + 1) * ny + j) * nx + i], k + 1 < nz, c + plane, true);
  if (gal) acc += x ? L.dd[c] * pc : L.dd[c];
  return acc * L.inv;
}

// kind of the first coarse level from the fine cells (delta3d_multigrid.rs:153-169, :357-367)
__global__ void __launch_bounds__(BLOCK) k_ckind_fine(Fine G, Level L) {
  FSP_FOR(c, L.n) {
    int i, j, k;
    ldecode(L, c, i, j, k);
    bool active = false, air = false;
    for (int ch = 0; ch < 8; ++ch) {
      const int di = ch & 1, dj = (ch >> 1) & 1, dk = (ch >> 2) & 1;
      const size_t f = cidx(G, 2 * i + di, 2 * j + dj, 2 * k + dk);
      const uint8_t m = G.mask[f];
      if (m & M_WET) active = true;
      else if (!G.solid[f]) air = true;
    }
    L.kind[c] = active ? 1 : (air ? -1 : 0);
  }
}
__global__ void __launch_bounds__(BLOCK) k_ckind(Level F, Level L) {
  FSP_FOR(c, L.n) {
    int i, j, k;
    ldecode(L, c, i, j, k);
    bool active = false, air = false;
    for (int ch = 0; ch < 8; ++ch) {
      const int di = ch & 1, dj = (ch >> 1) & 1, dk = (ch >> 2) & 1;
      const int8_t v = F.kind[lidx(F, 2 * i + di, 2 * j + dj, 2 * k + dk)];
      if (v > 0) active = true;
      else if (v < 0) air = true;
    }
    L.kind[c] = active ? 1 : (air ? -1 : 0);
  }
}

// Coarse apertures: a quarter of the four fine faces (delta3d_multigrid.rs:170-202).
// QUARTER order (0,0), (1,0), (0,1), (1,1).
template <class OpenU, class OpenV, class OpenW>
__device__ __forceinline__ void coarse_faces(const Level& L, size_t f, OpenU ou, OpenV ov, OpenW ow) {
#pragma clang fp contract(off)
  const size_t cx = L.nx, cy = L.ny, cz = L.nz;
  const size_t ncu = (cx + 1) * cy * cz, ncv = cx * (cy + 1) * cz;
  if (f < ncu) {
    const size_t i = f % (cx + 1), t = f / (cx + 1), j = t % cy, k = t / cy;
    float s = 0.f;
    s += ou(2 * i, 2 * j + 0, 2 * k + 0);
    s += ou(2 * i, 2 * j + 1, 2 * k + 0);
    s += ou(2 * i, 2 * j + 0, 2 * k + 1);
    s += ou(2 * i, 2 * j + 1, 2 * k + 1);
    L.ou[f] = 0.25f * s;
  } else if (f < ncu + ncv) {
    const size_t g = f - ncu;
    const size_t i = g % cx, t = g / cx, j = t % (cy + 1), k = t / (cy + 1);
    float s = 0.f;
    s += ov(2 * i + 0, 2 * j, 2 * k + 0);
    s += ov(2 * i + 1, 2 * j, 2 * k + 0);
    s += ov(2 * i + 0, 2 * j, 2 * k + 1);
    s += ov(2 * i + 1, 2 * j, 2 * k + 1);
    L.ov[g] = 0.25f * s;
  } else {
    const size_t g = f - ncu - ncv;
    const size_t i = g % cx, t = g / cx, j = t % cy, k = t / cy;
    float s = 0.f;
    s += ow(2 * i + 0, 2 * j + 0, 2 * k);
    s += ow(2 * i + 1, 2 * j + 0, 2 * k);
    s += ow(2 * i + 0, 2 * j + 1, 2 * k);
    s += ow(2 * i + 1, 2 * j + 1, 2 * k);
    L.ow[g] = 0.25f * s;
  }
}
__global__ void __launch_bounds__(BLOCK) k_cfaces_fine(Fine G, Level L) {
  const size_t cx = L.nx, cy = L.ny, cz = L.nz;
  const size_t total = (cx + 1) * cy * cz + cx * (cy + 1) * cz + cx * cy * (cz + 1);
  FSP_FOR(f, total) {
    coarse_faces(
        L, f,
        [&](size_t i, size_t j, size_t k) {
          const size_t q = fuidx(G, (int)i, (int)j, (int)k);
          return aperture(G, 0, q, G.ou[q]);
        },
        [&](size_t i, size_t j, size_t k) {
          const size_t q = fvidx(G, (int)i, (int)j, (int)k);
          return aperture(G, 1, q, G.ov[q]);
        },
        [&](size_t i, size_t j, size_t k) {
          const size_t q = fwidx(G, (int)i, (int)j, (int)k);
          return aperture(G, 2, q, G.ow[q]);
        });
  }
}
// Coarse face f of L from the faces of the finer level F.
__device__ __forceinline__ void level_face(const Level& F, const Level& L, size_t f) {
  const size_t fnx = F.nx, fny = F.ny;
  coarse_faces(
      L, f, [&](size_t i, size_t j, size_t k) { return F.ou[(k * fny + j) * (fnx + 1) + i]; },
      [&](size_t i, size_t j, size_t k) { return F.ov[(k * (fny + 1) + j) * fnx + i]; },
      [&](size_t i, size_t j, size_t k) { return F.ow[(k * fny + j) * fnx + i]; });
}
__global__ void __launch_bounds__(BLOCK) k_cfaces(Level F, Level L) {
  const size_t cx = L.nx, cy = L.ny, cz = L.nz;
  const size_t total = (cx + 1) * cy * cz + cx * (cy + 1) * cz + cx * cy * (cz + 1);
  FSP_FOR(f, total) level_face(F, L, f);
}
// Galerkin coarse levels (P^T A P / 16, P piecewise constant on wet cells):
// a coarse face carries a quarter of the couplings of its four fine faces
// (a fine face couples when both its cells are wet and it is open), and the
// Dirichlet part of the diagonal is a quarter of the children's ghost terms.
// Coarse coupling face f of L from the fine row codes (a fine face couples
// when both its cells are wet and it is open).
__device__ __forceinline__ void gal_fine_face(const Fine& G, const Level& L, size_t f) {
  coarse_faces(
      L, f,
      [&](size_t i, size_t j, size_t k) {
        return (i > 0 && i < (size_t)G.nx && (G.mask[cidx(G, (int)i, (int)j, (int)k)] & M_XM)) ? 1.0f : 0.0f;
      },
      [&](size_t i, size_t j, size_t k) {
        return (j > 0 && j < (size_t)G.ny && (G.mask[cidx(G, (int)i, (int)j, (int)k)] & M_YM)) ? 1.0f : 0.0f;
      },
      [&](size_t i, size_t j, size_t k) {
        return (k > 0 && k < (size_t)G.nz && (G.mask[cidx(G, (int)i, (int)j, (int)k)] & M_ZM)) ? 1.0f : 0.0f;
      });
}

// Dirichlet part of coarse cell c from the fine ghost terms of its children.
__device__ __forceinline__ void cdd_fine_cell(const Fine& G, const Level& L, size_t c) {
#pragma clang fp contract(off)
  int i, j, k;
  ldecode(L, c, i, j, k);
  float s = 0.f;
  for (int ch = 0; ch < 8; ++ch) {
    const int fi = 2 * i + (ch & 1), fj = 2 * j + ((ch >> 1) & 1), fk = 2 * k + ((ch >> 2) & 1);
    const uint8_t m = G.mask[cidx(G, fi, fj, fk)];
    float g = 0.f;
    if ((m & M_WET) && (m & M_GHOST)) {
      uint8_t type[6];
      float it[6], a[6];
      ghost_row(G, fi, fj, fk, m, type, it, a);
      for (int f = 0; f < 6; ++f)
        if (type[f] == 2) g += it[f];
    }
    s += g;
  }
  L.dd[c] = 0.25f * s;
}
// ... and from the Dirichlet parts of the children on the finer level F.
__device__ __forceinline__ void cdd_cell(const Level& F, const Level& L, size_t c) {
#pragma clang fp contract(off)
  int i, j, k;
  ldecode(L, c, i, j, k);
  float s = 0.f;
  for (int ch = 0; ch < 8; ++ch) {
    const int fi = 2 * i + (ch & 1), fj = 2 * j + ((ch >> 1) & 1), fk = 2 * k + ((ch >> 2) & 1);
    s += F.dd[lidx(F, fi, fj, fk)];
  }
  L.dd[c] = 0.25f * s;
}
// Galerkin level from the fine level: coupling faces and Dirichlet parts in
// two launches (one thread per face, one per cell: measured faster than one
// launch doing both, 33 against 39 us at 256x256x128).
__global__ void __launch_bounds__(BLOCK) k_cfaces_fine_gal(Fine G, Level L) {
  const size_t cx = L.nx, cy = L.ny, cz = L.nz;
  const size_t total = (cx + 1) * cy * cz + cx * (cy + 1) * cz + cx * cy * (cz + 1);
  FSP_FOR(f, total) gal_fine_face(G, L, f);
}
__global__ void __launch_bounds__(BLOCK) k_cdd_fine(Fine G, Level L) {
  FSP_FOR(c, L.n) cdd_fine_cell(G, L, c);
}
// Galerkin level from the finer Galerkin level F: faces (level_face) and
// Dirichlet parts (cdd_cell) in one launch, thread t doing face t and cell t
// (the coarse levels are launch-bound).
__global__ void __launch_bounds__(BLOCK) k_cgal(Level F, Level L) {
  const size_t cx = L.nx, cy = L.ny, cz = L.nz;
  const size_t nf = (cx + 1) * cy * cz + cx * (cy + 1) * cz + cx * cy * (cz + 1);
  const size_t total = nf > L.n ? nf : L.n;
  FSP_FOR(t, total) {
    if (t < nf) level_face(F, L, t);
    if (t < L.n) cdd_cell(F, L, t);
  }
}

__global__ void __launch_bounds__(BLOCK) k_cdiag(Level L) {
  FSP_FOR(c, L.n) {
    int i, j, k;
    ldecode(L, c, i, j, k);
    L.diag[c] = coarse_row<ST_F32>(L, i, j, k, c, nullptr);
  }
}

// Damped Jacobi on a coarse level (delta3d_multigrid.rs:290-299): xout = xin +
// w (r - L xin)/diag; xin == nullptr means xin = 0.
__global__ void __launch_bounds__(BLOCK) k_cjacobi(Level L, const float* xin, float* xout, Ctl* ctl) {
#pragma clang fp contract(off)
  if (!ctl->active) return;
  const float omega = 6.0f / 7.0f;
  FSP_FOR(c, L.n) {
    int i, j, k;
    ldecode(L, c, i, j, k);
    const float t = xin ? coarse_row(L, i, j, k, c, xin) : 0.f;
    float xc = xin ? xin[c] : 0.f;
    const float dg = L.diag[c];
    if (dg > 0.f) xc += omega * (L.r[c] - t) / dg;
    xout[c] = xc;
  }
}
// One colour of red-black Gauss-Seidel on a coarse level, in place; x and
// L.r are held in storage ST.
template <int ST>
__global__ void __launch_bounds__(BLOCK) k_cgs(Level L, typename Store<ST>::T* x, int colour, int first, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (!ctl->active) return;
  const typename SV::T* lr = reinterpret_cast<const typename SV::T*>(L.r);
  FSP_FOR(c, L.n) {
    int i, j, k;
    ldecode(L, c, i, j, k);
    const int col = (i + j + k) & 1;
    const float dg = L.diag[c];
    if (first) {
      if (col != colour) { SV::st(x, c, 0.f); continue; }
      SV::st(x, c, dg > 0.f ? SV::ld(lr, c) / dg : 0.f);
      continue;
    }
    if (col != colour) continue;
    if (dg > 0.f) {
      const float t = coarse_row<ST>(L, i, j, k, c, x);
      SV::st(x, c, SV::ld(x, c) + (SV::ld(lr, c) - t) / dg);
    }
  }
}

// Residual of the fine level restricted to the first coarse level
// (delta3d_multigrid.rs:406-411, restrict3 :258-272).
__global__ void __launch_bounds__(BLOCK) k_restrict_fine(Fine G, Level L, const float* z, const float* r, Ctl* ctl) {
#pragma clang fp contract(off)
  if (!ctl->active) return;
  FSP_FOR(c, L.n) {
    int i, j, k;
    ldecode(L, c, i, j, k);
    float s = 0.f;
    for (int ch = 0; ch < 8; ++ch) {
      const int fi = 2 * i + (ch & 1), fj = 2 * j + ((ch >> 1) & 1), fk = 2 * k + ((ch >> 2) & 1);
      const size_t f = cidx(G, fi, fj, fk);
      s += r[f] - apply_row(G, fi, fj, fk, f, G.mask[f], z);
    }
    L.r[c] = 0.125f * s;
  }
}
// Coarse residual restricted one level down (delta3d_multigrid.rs:312-321).
template <int ST>
__global__ void __launch_bounds__(BLOCK) k_crestrict(Level F, Level L, const typename Store<ST>::T* xf, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (!ctl->active) return;
  const typename SV::T* fr = reinterpret_cast<const typename SV::T*>(F.r);
  typename SV::T* lr = reinterpret_cast<typename SV::T*>(L.r);
  FSP_FOR(c, L.n) {
    int i, j, k;
    ldecode(L, c, i, j, k);
    float s = 0.f;
    for (int ch = 0; ch < 8; ++ch) {
      const int fi = 2 * i + (ch & 1), fj = 2 * j + ((ch >> 1) & 1), fk = 2 * k + ((ch >> 2) & 1);
      const size_t f = lidx(F, fi, fj, fk);
      s += SV::ld(fr, f) - coarse_row<ST>(F, fi, fj, fk, f, xf);
    }
    SV::st(lr, c, 0.125f * s);
  }
}
// Prolongation by injection (delta3d_multigrid.rs:275-287); on the fine
// level dry rows are forced back to 0 (:414-418).
__global__ void __launch_bounds__(BLOCK) k_prolong_fine(Fine G, Level L, const float* xc, float* z, Ctl* ctl) {
  if (!ctl->active) return;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    const uint8_t m = G.mask[c];
    if (!prec_positive(m)) { z[c] = 0.f; continue; }
    z[c] += xc[lidx(L, i >> 1, j >> 1, k >> 1)];
  }
}
template <int ST>
__global__ void __launch_bounds__(BLOCK) k_cprolong(Level F, Level L, const typename Store<ST>::T* xc,
                                                    typename Store<ST>::T* xf, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (!ctl->active) return;
  FSP_FOR(c, F.n) {
    int i, j, k;
    ldecode(F, c, i, j, k);
    SV::st(xf, c, SV::ld(xf, c) + SV::ld(xc, lidx(L, i >> 1, j >> 1, k >> 1)));
  }
}

// ---------------------------------------------------------------- correction and metrics
// Gradient subtract (delta3d_mobile.rs:284-343), one thread per face, from
// the f64 pressure (the gradient is formed in f64, the face rounded to f32),
// and max |u| over all faces (:422-427).
__device__ __forceinline__ float corrected_side(const Fine& G, float us, double pl, double pr, bool wl, bool wr,
                                                float k1, int x, int y, int i, int j, int k) {
#pragma clang fp contract(off)
  const double kk = (double)k1, dx = (double)G.dx;
  double ch;
  if (wl && wr) ch = kk * (pr - pl) / dx;
  else if (wl) { float a, b; ghost_side(G, x, y, k, i, j, a, b); ch = kk * ((double)b - pl) * (double)a / dx; }
  else if (wr) { float a, b; ghost_side(G, i, j, k, x, y, a, b); ch = kk * (pr - (double)b) * (double)a / dx; }
  else return us;
  return (float)((double)us - ch);
}

// One thread per cell: the cell's x-, y- and z- faces, and the closing face
// of its row / column / pillar at the +x, +y, +z boundary (32-bit index
// arithmetic only). Every operand of a face (aperture, both cells' solid
// flags and row codes, both pressures) is loaded at once, with indices
// clamped into the grid on the walls, and the face's case is selected after:
// one memory round trip per face instead of a chain of four.
__global__ void __launch_bounds__(BLOCK) k_correct(Fine G, const float* us, const float* vs, const float* ws,
                                                   const double* p, float k1, float* u, float* v, float* w, Ctl* ctl,
                                                   double* partials) {
#pragma clang fp contract(off)
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  double mx = 0.0, unused = 0.0;
  // horizontal face between cells l (x, y) and r (i, j) of plane k; inner = not on a wall
  auto side = [&](float val, uint8_t open, bool inner, size_t l, size_t r, int x, int y, int i, int j, int k) {
    const uint8_t sl = G.solid[l], sr = G.solid[r], ml = G.mask[l], mr = G.mask[r];
    const double pl = p[l], pr = p[r];
    if (!inner || !open || sl || sr) return val;
    return corrected_side(G, val, pl, pr, (ml & M_WET) != 0, (mr & M_WET) != 0, k1, x, y, i, j, k);
  };
  auto face_u = [&](int i, int j, int k) {
    const size_t f = fuidx(G, i, j, k);
    const int x = i > 0 ? i - 1 : 0, ii = i < nx ? i : nx - 1;
    const float val = side(us[f], G.ou[f], i > 0 && i < nx, cidx(G, x, j, k), cidx(G, ii, j, k), i - 1, j, i, j, k);
    u[f] = val;
    mx = fmax(mx, (double)fabsf(val));
  };
  auto face_v = [&](int i, int j, int k) {
    const size_t g = fvidx(G, i, j, k);
    const int y = j > 0 ? j - 1 : 0, jj = j < ny ? j : ny - 1;
    const float val = side(vs[g], G.ov[g], j > 0 && j < ny, cidx(G, i, y, k), cidx(G, i, jj, k), i, j - 1, i, j, k);
    v[g] = val;
    mx = fmax(mx, (double)fabsf(val));
  };
  auto face_w = [&](int i, int j, int k) {
    const size_t g = fwidx(G, i, j, k);
    const size_t below = cidx(G, i, j, k >= 1 ? k - 1 : 0), above = cidx(G, i, j, k < nz ? k : nz - 1);
    const uint8_t open = G.ow[g], mb = G.mask[below], ma = G.mask[above];
    const double pb = p[below], pa = p[above];
    float val = ws[g];
    if (G.phi) {
      // level set (apic3d.rs:1016-1048): a face next to a solid is not
      // corrected, and air may lie below water
      if (k >= 1 && k < nz && open && !G.solid[below] && !G.solid[above]) {
        const double kk = (double)k1, dx = (double)G.dx;
        const bool wb = (mb & M_WET) != 0, wa = (ma & M_WET) != 0;
        float a, b;
        if (wb && wa) val = (float)((double)val - kk * (pa - pb) / dx);
        else if (wb) { ghost_up(G, i, j, k - 1, a, b); val = (float)((double)val - kk * ((double)b - pb) * (double)a / dx); }
        else if (wa) { ghost_phi(G, i, j, k, i, j, k - 1, a, b); val = (float)((double)val - kk * (pa - (double)b) * (double)a / dx); }
      }
    } else if (k >= 1 && open && (mb & M_WET)) {
      const double kk = (double)k1, dx = (double)G.dx;
      if (k < nz && (ma & M_WET)) {
        val = (float)((double)val - kk * (pa - pb) / dx);
      } else {
        float a, b;
        ghost_up(G, i, j, k - 1, a, b);
        val = (float)((double)val - kk * ((double)b - pb) * (double)a / dx);
      }
    }
    w[g] = val;
    mx = fmax(mx, (double)fabsf(val));
  };
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    face_u(i, j, k);
    if (i == nx - 1) face_u(nx, j, k);
    face_v(i, j, k);
    if (j == ny - 1) face_v(i, ny, k);
    face_w(i, j, k);
    if (k == nz - 1) face_w(i, j, nz);
  }
  reduce_finish<true>(mx, unused, OP_UMAX, ctl, partials);
}

// max |div u| over wet rows and over plain rows (delta3d_mobile.rs:398-421).
__global__ void __launch_bounds__(BLOCK) k_divmetric(Fine G, const float* u, const float* v, const float* w, Ctl* ctl,
                                                     double* partials) {
  double all = 0.0, plain = 0.0;
  FSP_FOR(c, G.n) {
    const uint8_t m = G.mask[c];
    if (!(m & M_WET)) continue;
    int i, j, k;
    decode(G, c, i, j, k);
    const double d = (double)fabsf(divergence_at(G, u, v, w, i, j, k, c));
    all = fmax(all, d);
    // plain: no ghost face (a row flagged for a fractional aperture alone is plain)
    bool plain_row = !(m & M_GHOST);
    if (!plain_row && G.frac) {
      RowG R;
      full_row(G, i, j, k, R);
      plain_row = true;
      for (int f = 0; f < 6; ++f) plain_row = plain_row && R.type[f] != 2;
    }
    if (plain_row) plain = fmax(plain, d);
  }
  reduce_finish<true>(all, plain, OP_DIVMAX, ctl, partials);
}

__global__ void __launch_bounds__(BLOCK) k_divergence(Fine G, const float* u, const float* v, const float* w,
                                                      float* out) {
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    out[c] = divergence_at(G, u, v, w, i, j, k, c);
  }
}

// Oettli-Prager backward error of the true residual (delta3d_mobile.rs:243-271).
__global__ void __launch_bounds__(BLOCK) k_backward(Fine G, const double* p, const float* b, const float* r, Ctl* ctl,
                                                    double* partials) {
#pragma clang fp contract(off)
  const size_t plane = (size_t)G.nx * G.ny;
  double worst = 0.0, unused = 0.0;
  FSP_FOR(c, G.n) {
    int i, j, k;
    decode(G, c, i, j, k);
    // every operand loaded at once (neighbour indices clamped into the grid),
    // the row's case decided after
    const size_t nb[6] = {i > 0 ? c - 1 : c, i + 1 < G.nx ? c + 1 : c, j > 0 ? c - G.nx : c,
                          j + 1 < G.ny ? c + G.nx : c, k > 0 ? c - plane : c, k + 1 < G.nz ? c + plane : c};
    const uint8_t m = G.mask[c];
    const double pc = fabs(p[c]);
    double pn[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) pn[f] = p[nb[f]];
    const float bc = b[c], rc = r[c];
    if (!(m & M_WET)) continue;
    double acc = 0.0;
    if (!(m & M_GHOST)) {
      for (int f = 0; f < 6; ++f)
        if (m & (1u << f)) acc += pc + fabs(pn[f]);
    } else {
      uint8_t type[6];
      float it[6], a[6];
      ghost_row(G, i, j, k, m, type, it, a);
      for (int f = 0; f < 6; ++f) {
        if (type[f] == 1) acc += (double)a[f] * (pc + fabs(pn[f]));
        else if (type[f] == 2) acc += (double)a[f] * pc * (double)it[f];
      }
    }
    const double scale = fabs((double)bc) + acc * (double)G.inv;
    if (scale > 0.0) worst = fmax(worst, fabs((double)rc) / scale);
  }
  reduce_finish<true>(worst, unused, OP_BACKWARD, ctl, partials);
}

// State fingerprint of the pressure for the stall test of the pass loop (a
// repeated state ends it; the reference hashes p with FNV-1a,
// delta3d_mobile.rs:575): every value's bits, salted with its index, go
// through the splitmix64 finaliser (Steele, Lea and Flood 2014; Vigna's
// constants) and the results are summed mod 2^64 -- order-independent, so
// blocks reduce in parallel. Equal states give equal fingerprints. One
// partial per block into hashes[blockIdx.x].
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__global__ void __launch_bounds__(BLOCK) k_hash_chunks(const double* p, size_t n, unsigned long long* hashes) {
  __shared__ unsigned long long part[BLOCK];
  unsigned long long h = 0;
  for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x)
    h += mix64((unsigned long long)__double_as_longlong(p[s]) ^ ((unsigned long long)s * 0x9e3779b97f4a7c15ull));
  part[threadIdx.x] = h;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) hashes[blockIdx.x] = part[0];
}
// The chunk hashes' wrapping sum (exact in any order) by one workgroup of
// HASH_THREADS threads, then the final mix.
constexpr int HASH_THREADS = 1024;
__global__ void __launch_bounds__(HASH_THREADS) k_hash_combine(const unsigned long long* hashes, size_t chunks,
                                                               unsigned long long* out) {
  __shared__ unsigned long long part[HASH_THREADS];
  unsigned long long h = 0;
  for (size_t q = threadIdx.x; q < chunks; q += HASH_THREADS) h += hashes[q];
  part[threadIdx.x] = h;
  __syncthreads();
  for (int w = HASH_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = mix64(part[0]);
}

// Host-side parameters of a CG pass, written on the stream.
__global__ void k_set_pass(Ctl* ctl, int active, double rr, double tol_b2, double physical_target, uint32_t it,
                           uint32_t max_iters, uint32_t pass_limit, int flexible, uint32_t replace_every) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctl->active = active;
  ctl->flexible = flexible;
  ctl->replace_every = replace_every;
  ctl->stalled = 0;
  ctl->rr_replaced = rr;
  ctl->rr = rr;
  ctl->tol_b2 = tol_b2;
  ctl->physical_target = physical_target;
  ctl->it = it;
  ctl->max_iters = max_iters;
  ctl->pass_limit = pass_limit;
  ctl->stop_dq = 0;
  ctl->nonfinite = 0;
}

}  // namespace fsp
