// The graded column (delta3d_graded.rs): below the cubic top layers the
// pressure lives at node layers only, the same for every column, linear
// between two nodes. P prolongs node values to the fine layers, P^T is its
// exact transpose (same coefficients, the reference's summation order), and
// the projection solves P^T A P x = P^T b by conjugate gradients with the
// condensed diagonal sum_k P_kj^2 A_kk as preconditioner. The reduced vectors
// are node-major: x[s * cols + column].
#pragma once
#include "kernels.h"

namespace fsp {

struct GradedCol {
  int cols, nn, nz, top;  // columns, nodes, fine layers, the top node's layer
  const int* node;        // nn: the nodes' fine layers, bottom up
  const int* seg;         // nz: the segment s of layer k (node[s] <= k < node[s + 1]), k < top
  const float* wa;        // nz: 1 - t of layer k in its segment
  const float* wb;        // nz: t = (k - a) / (b - a)
};

// fine = P x (delta3d_graded.rs:109-126), one thread per fine cell.
__global__ void __launch_bounds__(BLOCK) k_gr_expand(GradedCol D, const float* x, float* fine) {
#pragma clang fp contract(off)
  const size_t cols = (size_t)D.cols, n = cols * (size_t)D.nz;
  FSP_FOR(c, n) {
    const int k = (int)(c / cols);
    const size_t col = c - (size_t)k * cols;
    if (k == D.top) {
      fine[c] = x[(size_t)(D.nn - 1) * cols + col];
    } else {
      const size_t s = (size_t)D.seg[k];
      fine[c] = D.wa[k] * x[s * cols + col] + D.wb[k] * x[(s + 1) * cols + col];
    }
  }
}

// out = P^T fine (delta3d_graded.rs:128-145), one thread per (node, column),
// the reference's order: the t-terms of the segment below (k ascending),
// then the (1 - t)-terms of the segment above, then the top layer for the
// top node. CONDENSED: sum of (w w) A_kk with A_kk = 1 / prec_k on a wet row
// and 0 on a dry one, `fine` holding prec (the condensed diagonal,
// delta3d_graded.rs:357-377).
template <bool CONDENSED>
__global__ void __launch_bounds__(BLOCK) k_gr_restrict(GradedCol D, const float* fine, float* out) {
#pragma clang fp contract(off)
  const size_t cols = (size_t)D.cols, n = cols * (size_t)D.nn;
  auto term = [&](int k, size_t col) -> float {
    const float f = fine[(size_t)k * cols + col];
    return CONDENSED ? (f > 0.f ? 1.0f / f : 0.f) : f;
  };
  FSP_FOR(q, n) {
    const int s = (int)(q / cols);
    const size_t col = q - (size_t)s * cols;
    float acc = 0.f;
    if (s > 0)
      for (int k = D.node[s - 1]; k < D.node[s]; ++k) {
        const float w = CONDENSED ? D.wb[k] * D.wb[k] : D.wb[k];
        acc += w * term(k, col);
      }
    if (s + 1 < D.nn)
      for (int k = D.node[s]; k < D.node[s + 1]; ++k) {
        const float w = CONDENSED ? D.wa[k] * D.wa[k] : D.wa[k];
        acc += w * term(k, col);
      }
    if (s == D.nn - 1) acc += term(D.top, col);
    out[q] = acc;
  }
}

// x = the fine pressure at the nodes where the condensed diagonal is
// positive, else 0 (the warm start by injection, delta3d_graded.rs:378-383).
__global__ void __launch_bounds__(BLOCK) k_gr_inject(GradedCol D, const float* p, const float* m, float* x) {
  const size_t cols = (size_t)D.cols, n = cols * (size_t)D.nn;
  FSP_FOR(q, n) {
    const int s = (int)(q / cols);
    const size_t col = q - (size_t)s * cols;
    x[q] = m[q] > 0.f ? p[(size_t)D.node[s] * cols + col] : 0.f;
  }
}

// Per-block f64 partials of <a, b> and <c, e> (c null: the second is 0),
// into slot 0 of `partials`.
__global__ void __launch_bounds__(BLOCK) k_gr_dot(size_t n, const float* a, const float* b, const float* c,
                                                  const float* e, double* partials) {
  double s = 0.0, t = 0.0;
  FSP_FOR(q, n) {
    s += (double)a[q] * (double)b[q];
    if (c) t += (double)c[q] * (double)e[q];
  }
  block_partial<false>(s, t, 0, partials);
}

// Folds `blocks` partial pairs of slot 0 in a fixed order into out[0..1]
// (sums) or their maxima (MAX). One workgroup.
template <bool MAX>
__global__ void __launch_bounds__(BLOCK) k_gr_fold(int blocks, const double* partials, double* out) {
  double a = 0.0, b = 0.0;
  for (int q = (int)threadIdx.x; q < blocks; q += BLOCK) {
    a = MAX ? fmax(a, partials[2 * q]) : a + partials[2 * q];
    b = MAX ? fmax(b, partials[2 * q + 1]) : b + partials[2 * q + 1];
  }
  __shared__ double sa[BLOCK / 64], sb[BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  a = MAX ? wave_max(a) : wave_sum(a);
  b = MAX ? wave_max(b) : wave_sum(b);
  if (lane == 0) { sa[wid] = a; sb[wid] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = sa[0], tb = sb[0];
    for (int w = 1; w < BLOCK / 64; ++w) {
      ta = MAX ? fmax(ta, sa[w]) : ta + sa[w];
      tb = MAX ? fmax(tb, sb[w]) : tb + sb[w];
    }
    out[0] = ta;
    out[1] = tb;
  }
}

// r = b - q; z = r / m where m > 0, else 0; d = z (the true residual and
// the preconditioned direction that start a pass, delta3d_graded.rs:388-399).
__global__ void __launch_bounds__(BLOCK) k_gr_restart(size_t n, const float* b, const float* q, const float* m,
                                                      float* r, float* z, float* d) {
#pragma clang fp contract(off)
  FSP_FOR(c, n) {
    const float rc = b[c] - q[c];
    r[c] = rc;
    if (z) {
      const float zc = m[c] > 0.f ? rc / m[c] : 0.f;
      z[c] = zc;
      d[c] = zc;
    }
  }
}

// x += alpha d, r -= alpha q, z = r / m (delta3d_graded.rs:409-413).
__global__ void __launch_bounds__(BLOCK) k_gr_update(size_t n, float alpha, const float* d, const float* q,
                                                     const float* m, float* x, float* r, float* z) {
#pragma clang fp contract(off)
  FSP_FOR(c, n) {
    x[c] += alpha * d[c];
    const float rc = r[c] - alpha * q[c];
    r[c] = rc;
    z[c] = m[c] > 0.f ? rc / m[c] : 0.f;
  }
}

// d = z + beta d (delta3d_graded.rs:416-418).
__global__ void __launch_bounds__(BLOCK) k_gr_direction(size_t n, float beta, const float* z, float* d) {
#pragma clang fp contract(off)
  FSP_FOR(c, n) d[c] = z[c] + beta * d[c];
}

// The restricted divergence metric (delta3d_graded.rs:451-486): per block
// the maxima of |P^T div u| / weight over wet node cells (all) and over the
// plain ones (graded nodes and the lowest cubic node j0; above it, rows
// without a ghost face), from the row codes.
__global__ void __launch_bounds__(BLOCK) k_gr_metric(GradedCol D, const float* rdiv, const float* weight,
                                                     const uint8_t* mask, int j0, double* partials) {
#pragma clang fp contract(off)
  const size_t cols = (size_t)D.cols, n = cols * (size_t)D.nn;
  double all = 0.0, plain = 0.0;
  FSP_FOR(q, n) {
    const int s = (int)(q / cols);
    const size_t col = q - (size_t)s * cols;
    const uint8_t m = mask[(size_t)D.node[s] * cols + col];
    if (!(m & M_WET)) continue;
    const float d = fabsf(rdiv[q]) / weight[s];
    all = fmax(all, (double)d);
    if (s <= j0 || !(m & M_GHOST)) plain = fmax(plain, (double)d);
  }
  block_partial<true>(all, plain, 0, partials);
}

// Per-block max |x| over three arrays (the faces' largest speed).
__global__ void __launch_bounds__(BLOCK) k_gr_absmax(const float* u, size_t nu, const float* v, size_t nv,
                                                     const float* w, size_t nw, double* partials) {
  double mx = 0.0;
  FSP_FOR(q, nu + nv + nw) {
    const float x = q < nu ? u[q] : q < nu + nv ? v[q - nu] : w[q - nu - nv];
    mx = fmax(mx, (double)fabsf(x));
  }
  block_partial<true>(mx, 0.0, 0, partials);
}

// p64 = p (the fine pressure for the correction kernels).
__global__ void __launch_bounds__(BLOCK) k_gr_widen(size_t n, const float* p, double* p64) {
  FSP_FOR(c, n) p64[c] = (double)p[c];
}

}  // namespace fsp
