// Kernels of the steps either side of the projection in the uncoupled
// surface step (Volume3::step_surface_mobile, delta3d_mobile.rs:819-927):
// advection of the MAC faces, extrapolation of the solved velocities up the
// columns, transport of the column heights, and the step's scalar checks.
//
// Arithmetic follows the reference expression by expression in f32 with FMA
// contraction off, so every face and height is bit-identical to the
// reference's on the same inputs. Scope: no sparse column set
// (sparse_outside3 is false, delta3d_sparse.rs:306-309), the second-order
// advection term off (its default, delta3d_advection.rs:31-36), binary
// apertures.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fsp_internal.h"

namespace fsp {

__device__ __forceinline__ size_t st_fu(const Fine& G, int i, int j, int k) {
  return ((size_t)k * G.ny + j) * (G.nx + 1) + i;
}
__device__ __forceinline__ size_t st_fv(const Fine& G, int i, int j, int k) {
  return ((size_t)k * (G.ny + 1) + j) * G.nx + i;
}
__device__ __forceinline__ size_t st_fw(const Fine& G, int i, int j, int k) { return ((size_t)k * G.ny + j) * G.nx + i; }
__device__ __forceinline__ size_t st_c(const Fine& G, int i, int j, int k) { return ((size_t)k * G.ny + j) * G.nx + i; }

// wet3 (delta3d_mobile.rs:97-103)
__device__ __forceinline__ bool st_wet(const Fine& G, int i, int j, int k) {
#pragma clang fp contract(off)
  if (G.solid[st_c(G, i, j, k)]) return false;
  return ((float)k + 0.5f) * G.dx < G.eta[(size_t)j * G.nx + i];
}

// Horizontal faces (AXIS 0: u, 1: v) of advect_mobile3 (delta3d_mobile.rs:
// 621-684): one thread per face; faces on the walls and closed faces keep
// their value.
template <int AXIS>
__global__ void __launch_bounds__(256) k_advect_h(Fine G, const float* u, const float* v, const float* w, float dt,
                                                  float h, float* out) {
#pragma clang fp contract(off)
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const size_t nf = AXIS == 0 ? (size_t)(nx + 1) * ny * nz : (size_t)nx * (ny + 1) * nz;
  const float* own = AXIS == 0 ? u : v;
  const uint8_t* open = AXIS == 0 ? G.ou : G.ov;
  for (size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    if (AXIS == 0) {
      i = (int)(f % (nx + 1));
      const size_t t = f / (nx + 1);
      j = (int)(t % ny);
      k = (int)(t / ny);
    } else {
      i = (int)(f % nx);
      const size_t t = f / nx;
      j = (int)(t % (ny + 1));
      k = (int)(t / (ny + 1));
    }
    const int a = AXIS == 0 ? i : j, na = AXIS == 0 ? nx : ny;
    const int b = AXIS == 0 ? j : i, nb = AXIS == 0 ? ny : nx;
    const float uc = own[f];
    if (a == 0 || a == na || open[f] == 0) {
      out[f] = uc;
      continue;
    }
    // (a, b) along / across the face's own axis, as in the reference
    auto val = [&](int aa, int bb, int kk) { return AXIS == 0 ? u[st_fu(G, aa, bb, kk)] : v[st_fv(G, bb, aa, kk)]; };
    auto vertical = [&](int aa, int bb, int kk) { return AXIS == 0 ? w[st_fw(G, aa, bb, kk)] : w[st_fw(G, bb, aa, kk)]; };
    auto transverse = [&](int aa, int bb, int kk) {
      return AXIS == 0 ? v[st_fv(G, aa, bb, kk)] : u[st_fu(G, bb, aa, kk)];
    };
    const float ux = (val(a + 1, b, k) - val(a - 1, b, k)) * h;
    const float up = k + 1 < nz ? val(a, b, k + 1) : uc;
    const float dn = k > 0 ? val(a, b, k - 1) : uc;
    const float uz = (up - dn) * h;
    const float wc =
        0.25f * (vertical(a - 1, b, k) + vertical(a, b, k) + vertical(a - 1, b, k + 1) + vertical(a, b, k + 1));
    const float rt = b + 1 < nb ? val(a, b + 1, k) : uc;
    const float lf = b > 0 ? val(a, b - 1, k) : uc;
    const float uy = (rt - lf) * h;
    const float vc = 0.25f * (transverse(a - 1, b, k) + transverse(a, b, k) + transverse(a - 1, b + 1, k) +
                              transverse(a, b + 1, k));
    out[f] = uc - dt * (uc * ux + wc * uz + vc * uy);
  }
}

// Vertical faces of advect_mobile3 (delta3d_mobile.rs:685-731); the bottom
// and top faces and closed faces keep their value.
__global__ void __launch_bounds__(256) k_advect_w(Fine G, const float* u, const float* v, const float* w, float dt,
                                                  float h, float* out) {
#pragma clang fp contract(off)
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const size_t nf = (size_t)nx * ny * (nz + 1);
  for (size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(f % nx);
    const size_t t = f / nx;
    const int j = (int)(t % ny), k = (int)(t / ny);
    const float wc = w[f];
    if (k == 0 || k == nz || G.ow[f] == 0) {
      out[f] = wc;
      continue;
    }
    const float rt = i + 1 < nx ? w[st_fw(G, i + 1, j, k)] : wc;
    const float lf = i > 0 ? w[st_fw(G, i - 1, j, k)] : wc;
    const float wx = (rt - lf) * h;
    const float wz = (w[st_fw(G, i, j, k + 1)] - w[st_fw(G, i, j, k - 1)]) * h;
    const float uc = 0.25f * (u[st_fu(G, i, j, k - 1)] + u[st_fu(G, i + 1, j, k - 1)] + u[st_fu(G, i, j, k)] +
                              u[st_fu(G, i + 1, j, k)]);
    const float bk = j + 1 < ny ? w[st_fw(G, i, j + 1, k)] : wc;
    const float fr = j > 0 ? w[st_fw(G, i, j - 1, k)] : wc;
    const float wy = (bk - fr) * h;
    const float vc = 0.25f * (v[st_fv(G, i, j, k - 1)] + v[st_fv(G, i, j + 1, k - 1)] + v[st_fv(G, i, j, k)] +
                              v[st_fv(G, i, j + 1, k)]);
    out[f] = wc - dt * (uc * wx + wc * wz + vc * wy);
  }
}

// extrapolate_mobile3 (delta3d_mobile.rs:345-394): one thread per (column,
// axis), x fastest across threads so the faces of one plane are read
// together; each marches up its column carrying the last solved velocity,
// EX levels at a time: the loads of a group of levels are all issued before
// the carried value walks through them.
constexpr int EX = 8;
__global__ void __launch_bounds__(256) k_extrapolate(Fine G, float* u, float* v, float* w) {
#pragma clang fp contract(off)
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const size_t cols = (size_t)nx * ny;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < 3 * cols; t += (size_t)gridDim.x * blockDim.x) {
    const int axis = (int)(t / cols);
    const size_t cc = t - axis * cols;
    const int i = (int)(cc % nx), j = (int)(cc / nx);
    if ((axis == 0 && i == 0) || (axis == 1 && j == 0)) continue;
    float* field = axis == 0 ? u : axis == 1 ? v : w;
    const uint8_t* open = axis == 0 ? G.ou : axis == 1 ? G.ov : G.ow;
    // the two cells a horizontal face joins (this column and the one at -x or -y);
    // a w face's cell below is in this column
    const int x = axis == 0 ? i - 1 : i, y = axis == 1 ? j - 1 : j;
    const float h_here = G.eta[(size_t)j * nx + i], h_prev = G.eta[(size_t)y * nx + x];
    bool have = false;
    float last = 0.f;
    const int k0 = axis == 2 ? 1 : 0, k1 = axis == 2 ? nz + 1 : nz;
    for (int kb = k0; kb < k1; kb += EX) {
      uint8_t op[EX], sa[EX], sb[EX];
      float val[EX];
#pragma unroll
      for (int e = 0; e < EX; ++e) {
        const int k = kb + e < k1 ? kb + e : k1 - 1;
        size_t f;
        if (axis == 0) {
          f = st_fu(G, i, j, k);
          sa[e] = G.solid[st_c(G, x, y, k)];
          sb[e] = G.solid[st_c(G, i, j, k)];
        } else if (axis == 1) {
          f = st_fv(G, i, j, k);
          sa[e] = G.solid[st_c(G, x, y, k)];
          sb[e] = G.solid[st_c(G, i, j, k)];
        } else {
          f = st_fw(G, i, j, k);
          sa[e] = G.solid[st_c(G, i, j, k - 1)];
          sb[e] = 0;
        }
        op[e] = open[f];
        val[e] = field[f];
      }
#pragma unroll
      for (int e = 0; e < EX; ++e) {
        const int k = kb + e;
        if (k >= k1) break;
        if (op[e] == 0) continue;
        bool solved;
        if (axis < 2) {
          // wet3 of both cells (delta3d_mobile.rs:97-103): not solid, centre below the surface
          const float zc = ((float)k + 0.5f) * G.dx;
          solved = !sa[e] && !sb[e] && (zc < h_prev || zc < h_here);
        } else {
          const float zc = ((float)(k - 1) + 0.5f) * G.dx;
          solved = !sa[e] && zc < h_here;
        }
        size_t f = axis == 0 ? st_fu(G, i, j, k) : axis == 1 ? st_fv(G, i, j, k) : st_fw(G, i, j, k);
        if (solved) {
          last = val[e];
          have = true;
        } else if (have) {
          field[f] = last;
        }
      }
    }
  }
}

// The wetted discharges of transport_mobile3 (delta3d_mobile.rs:741-770),
// every entry of flux_x ((nx+1)*ny) then flux_y (nx*(ny+1)); the wall
// entries are 0. Reads eta^n only.
__global__ void __launch_bounds__(256) k_flux(Fine G, const float* u, const float* v, float* flux_x, float* flux_y) {
#pragma clang fp contract(off)
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const float dx = G.dx;
  const size_t nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < nfx + nfy; t += (size_t)gridDim.x * blockDim.x) {
    const int axis = t < nfx ? 0 : 1;
    int i, j;
    if (axis == 0) {
      i = (int)(t % (nx + 1));
      j = (int)(t / (nx + 1));
    } else {
      const size_t q = t - nfx;
      i = (int)(q % nx);
      j = (int)(q / nx);
    }
    float* out = axis == 0 ? &flux_x[t] : &flux_y[t - nfx];
    if ((axis == 0 && (i == 0 || i == nx)) || (axis == 1 && (j == 0 || j == ny))) {
      *out = 0.f;
      continue;
    }
    const int x = axis == 0 ? i - 1 : i, y = axis == 0 ? j : j - 1;
    const float surface = 0.5f * (G.eta[(size_t)y * nx + x] + G.eta[(size_t)j * nx + i]);
    float q = 0.f;
    for (int k = 0; k < nz; ++k) {
      float wet = (surface - (float)k * dx) / dx;
      wet = wet < 0.f ? 0.f : (wet > 1.f ? 1.f : wet);  // Rust f32::clamp(0, 1): NaN passes through
      if (wet == 0.f) break;
      float vel, a;
      if (axis == 0) {
        const size_t f = st_fu(G, i, j, k);
        vel = u[f];
        a = G.ou[f] ? 1.f : 0.f;
      } else {
        const size_t f = st_fv(G, i, j, k);
        vel = v[f];
        a = G.ov[f] ? 1.f : 0.f;
      }
      q += a * vel * dx * wet;
    }
    *out = q;
  }
}

// The height update of transport_mobile3 (delta3d_mobile.rs:771-781), the
// compensated roundoff carried in eta_rt.
__global__ void __launch_bounds__(256) k_heights(Fine G, const float* flux_x, const float* flux_y, float transport,
                                                 float* eta, float* eta_rt) {
#pragma clang fp contract(off)
  const int nx = G.nx, ny = G.ny;
  const size_t cols = (size_t)nx * ny;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nx), j = (int)(c / nx);
    const float x = flux_x[(size_t)j * (nx + 1) + i + 1] - flux_x[(size_t)j * (nx + 1) + i];
    const float y = flux_y[(size_t)(j + 1) * nx + i] - flux_y[(size_t)j * nx + i];
    const float increment = -transport * (x + y) - eta_rt[c];
    const float e = eta[c];
    const float height = e + increment;
    eta_rt[c] = (height - e) - increment;
    eta[c] = height;
  }
}

// The step's scalars. One workgroup, columns in a fixed order, so every
// value is deterministic.
struct StepScalars {
  double volume;    // perturbation_volume (delta3d_balance.rs:105-113)
  double boundary;  // boundary_flux_mobile3 before the dt*dx factor (delta3d_mobile.rs:784-800)
  int in_bounds;    // mobile_in_bounds (delta3d_mobile.rs:801-813)
  int finite;       // 1 until a non-finite value is seen (k_finite)
};

constexpr int STEP_THREADS = 1024;

// volume and in_bounds into *s; with flux arrays, also the wall sum. floor:
// per column, the top of its highest solid cell (0 without one), as the
// reference's cut geometry raises it (delta3d_cut.rs:429-447).
__global__ void __launch_bounds__(STEP_THREADS) k_step_scalars(Fine G, const float* floor, const float* flux_x,
                                                               const float* flux_y, StepScalars* s) {
#pragma clang fp contract(off)
  __shared__ double red[STEP_THREADS];
  __shared__ int ok_all;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const size_t cols = (size_t)nx * ny;
  const float dx = G.dx, top = (float)(nz - 1) * dx;
  if (threadIdx.x == 0) ok_all = 1;
  __syncthreads();
  double vol = 0.0;
  int ok = 1;
  auto column = [&](float e, float rt, float fl) {
    vol += ((double)e - (double)rt) - (double)G.rest;
    if (!(e >= fl + 2.f * dx && e <= top)) ok = 0;
  };
  // eight columns' loads in flight per round, summed in the thread's order
  constexpr int U = 8;
  size_t c = threadIdx.x;
  for (; c + (size_t)(U - 1) * STEP_THREADS < cols; c += (size_t)U * STEP_THREADS) {
    float e[U], rt[U], fl[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      e[u] = G.eta[c + (size_t)u * STEP_THREADS];
      rt[u] = G.eta_rt[c + (size_t)u * STEP_THREADS];
      fl[u] = floor[c + (size_t)u * STEP_THREADS];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) column(e[u], rt[u], fl[u]);
  }
  for (; c < cols; c += STEP_THREADS) column(G.eta[c], G.eta_rt[c], floor[c]);
  if (!ok) ok_all = 0;  // every writer stores the same value
  red[threadIdx.x] = vol;
  __syncthreads();
  for (int w = STEP_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double volume = red[0] * ((double)dx * (double)dx);
  __syncthreads();
  double bsum = 0.0;
  if (flux_x) {
    for (int j = threadIdx.x; j < ny; j += STEP_THREADS)
      bsum += (double)flux_x[(size_t)j * (nx + 1)] - (double)flux_x[(size_t)j * (nx + 1) + nx];
    for (int i = threadIdx.x; i < nx; i += STEP_THREADS)
      bsum += (double)flux_y[i] - (double)flux_y[(size_t)ny * nx + i];
  }
  red[threadIdx.x] = bsum;
  __syncthreads();
  for (int w = STEP_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    s->volume = volume;
    s->boundary = red[0];
    s->in_bounds = ok_all;
  }
}

// Clears s->finite to 0 when a value of a is not finite (every writer
// stores the same value).
// The fields a step writes, checked for non-finite values in one launch: the
// grid's row y takes field y.
constexpr int FINITE_FIELDS = 11;
struct FiniteSet {
  const float* a[FINITE_FIELDS];
  size_t n[FINITE_FIELDS];
};
__global__ void __launch_bounds__(256) k_finite(FiniteSet F, StepScalars* s) {
  const float* a = F.a[blockIdx.y];
  const size_t n = F.n[blockIdx.y];
  bool bad = false;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    if (!isfinite(a[q])) bad = true;
  if (bad) s->finite = 0;
}

__global__ void k_step_reset(StepScalars* s) {
  s->volume = 0.0;
  s->boundary = 0.0;
  s->in_bounds = 1;
  s->finite = 1;
}

}  // namespace fsp
