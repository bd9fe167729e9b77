// CPU ORACLE — test infrastructure only.
//
// A from-scratch C++ restatement of the reference's δ-3D pressure projection
// (`Volume3::project_mobile3` with its multigrid preconditioner). It exists to
// check the HIP library: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or as the timed CPU
// baseline. The product path never calls it.
//
// Template parameter T:
//   float  — "f32-faithful": the reference's loop order, summation order and
//            chunk-64 ordered f32 reductions; build with -ffp-contract=off.
//   double — the FP64 oracle used to state the error of mixed-precision runs.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference checkout, code/water-core/src/).
//
// Pinning (what checks each part, see tests/test_oracle_*.py):
//   - project/v_cycle/operator: the reference's published iteration counts of
//     its S385 bench (MULTIGRILLE-3D-S385.md:57-58: 9/10/11 with multigrid,
//     102/191/365 with Jacobi, nx = 32/64/128) reproduced exactly;
//   - operator/assembly/correction/divergence: dense brute-force solves on
//     small random geometries, DCT closed forms on box domains, symmetry,
//     positivity and prec == 1/diag bitwise (tests_delta3d.rs:352, :507);
//   - hierarchy: level_count3 known answers (tests_delta3d.rs:1814-1817);
//   - fingerprint: FNV-1a 64 published offset/prime known answers;
//   - V-cycle: symmetry and positivity (tests_delta3d.rs:1840-1872).
// Extensions that are not in the reference (RBGS smoother, level cap, coarse
// sweep count, pure-Neumann gauge) are marked EXTENSION; with their default
// settings the oracle is the reference algorithm.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

enum Status : int32_t { OK = 0, ERR_DOMAIN = 1, ERR_NOT_FINITE = 2, ERR_SHAPE = 3, ERR_CONVERGENCE = 4 };

// delta_projection.rs:164-188
struct Report {
  uint32_t refinements;
  uint32_t iterations;
  int32_t degraded;
  int32_t floor;
  double residual;
  double divergence;
  double backward_error;
  double divergence_plain;
};

struct Options {
  int32_t multigrid;      // 0: Jacobi-PCG (delta3d_mobile.rs:435-461), 1: V-cycle (delta3d_multigrid.rs:393-465)
  int32_t smoother;       // 0: damped Jacobi ω=6/7 (delta3d_multigrid.rs:28); 1: EXTENSION symmetric red-black Gauss-Seidel
  int32_t max_levels;     // EXTENSION: cap on coarse levels; <0 = the reference rule (delta3d_multigrid.rs:73-87)
  int32_t pre_sweeps;     // delta3d_multigrid.rs:32 (2)
  int32_t post_sweeps;    // delta3d_multigrid.rs:33 (2)
  int32_t coarse_sweeps;  // delta3d_multigrid.rs:34 (8)
  int32_t gauge;          // EXTENSION: 1 = remove the mean of b and of p on a pure-Neumann problem
  int32_t coarse_op;      // 0: re-discretised coarse levels (delta3d_multigrid.rs:141-255);
                          // 1: EXTENSION Galerkin P^T A P / 16 with piecewise-constant P
  int32_t storage;        // EXTENSION: V-cycle vectors rounded to 0 f32 (the reference), 1 IEEE binary16,
                          // 2 bfloat16 at every store; arithmetic stays in T
  int32_t flexible;       // EXTENSION: 1 = flexible CG, beta = <z', r' - r> / <z, r> (Polak-Ribiere)
  int32_t level_pre;      // EXTENSION: sweeps before restriction on the coarse levels (<0: pre_sweeps)
  int32_t level_post;     // EXTENSION: sweeps after prolongation on the coarse levels (<0: post_sweeps)
  double rel_tol;         // EXTENSION: relative residual target; 1e-6 is the reference's (delta3d_mobile.rs:507)
};

inline Options default_options() { return Options{1, 0, -1, 2, 2, 8, 1, 0, 0, 0, -1, -1, 1e-6}; }

// Round-to-nearest-even of an f32 value to bfloat16 (8 significant bits,
// f32 exponent range); the result is returned as the f32 it represents.
inline float round_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return x;  // inf / nan unchanged
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float y;
  std::memcpy(&y, &u, 4);
  return y;
}
// Round-to-nearest-even of an f32 value to IEEE binary16 (11 significant
// bits, normal range [2^-14, 65504], subnormal step 2^-24, overflow to inf).
inline float round_f16(float x) {
  const float ax = std::fabs(x);
  if (!(ax < std::numeric_limits<float>::infinity())) return x;
  if (ax < 6.103515625e-05f) {  // 2^-14: fixed step 2^-24
    const float y = std::nearbyint(ax * 16777216.0f) * 5.9604644775390625e-08f;
    return std::copysign(y, x);
  }
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0xFFFu + ((u >> 13) & 1u);
  u &= 0xFFFFE000u;
  float y;
  std::memcpy(&y, &u, 4);
  if (std::fabs(y) > 65504.0f) return std::copysign(std::numeric_limits<float>::infinity(), x);
  return y;
}
template <class T> inline T round_storage(int storage, T x) {
  if (storage == 1) return (T)round_f16((float)x);
  if (storage == 2) return (T)round_bf16((float)x);
  return x;
}

// Geometry: the subset of Volume3's state that the projection reads.
struct Geometry {
  int32_t nx, ny, nz;
  float dx;
  float rho, g_eff, rest;
  int32_t homogeneous_ghost;   // delta3d_mobile.rs:130, :140
  int32_t closed_lid;          // EXTENSION of the cut apertures: aperture 0 on every top face
  const float* eta;            // nx*ny column heights
  const float* eta_roundoff;   // nx*ny or null (zeros)
  const uint8_t* solid;        // cells, nonzero = solid (cut frac == 0), or null
  const float* open_u;         // apertures, or null (derived)
  const float* open_v;
  const float* open_w;
};

// delta3d.rs:69-72 and delta_projection.rs:200-203: gamma_{m+4} = (m+4)u/(1-(m+4)u)
// with u the unit roundoff of T (2^-24 in f32 as in the reference; 2^-53 in
// the f64 oracle, whose floor is that of f64).
template <class T> inline T unit_roundoff() { return std::numeric_limits<T>::epsilon() / T(2); }
template <class T> inline T roundoff_3d() { const T u = unit_roundoff<T>(); return T(10) * u / (T(1) - T(10) * u); }
template <class T> inline T roundoff_2d() { const T u = unit_roundoff<T>(); return T(8) * u / (T(1) - T(8) * u); }

// delta3d.rs:1360-1369 — FNV-1a over the little-endian bytes of each f32.
inline uint64_t fingerprint_f32(const float* v, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    uint32_t bits;
    std::memcpy(&bits, &v[i], 4);
    for (int b = 0; b < 4; ++b) {
      h ^= (uint64_t)((bits >> (8 * b)) & 0xffu);
      h *= 0x100000001b3ull;
    }
  }
  return h;
}

template <class T> inline uint64_t fingerprint(const std::vector<T>& v) {
  // The f64 oracle fingerprints the f32 image of p (the cycle test is about
  // a state that returns bit for bit; the f64 state rarely cycles).
  std::vector<float> f(v.size());
  for (size_t i = 0; i < v.size(); ++i) f[i] = (float)v[i];
  return fingerprint_f32(f.data(), f.size());
}

// delta3d_multigrid.rs:73-87
inline bool coarsens3(size_t nx, size_t ny, size_t nz) {
  return nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0 && nx >= 4 && ny >= 4 && nz >= 4;
}
inline size_t level_count3(size_t nx, size_t ny, size_t nz) {
  size_t n = 0;
  while (coarsens3(nx, ny, nz)) { nx /= 2; ny /= 2; nz /= 2; ++n; }
  return n;
}
// delta3d_multigrid.rs:91-105 (no overflow concerns at oracle sizes)
inline size_t hierarchy_floats3(size_t nx, size_t ny, size_t nz) {
  size_t total = 2 * nx * ny * nz;
  while (coarsens3(nx, ny, nz)) {
    size_t cx = nx / 2, cy = ny / 2, cz = nz / 2;
    size_t cells = cx * cy * cz;
    size_t faces = (cx + 1) * cy * cz + cx * (cy + 1) * cz + cx * cy * (cz + 1);
    total += cells * 5 + faces;
    nx = cx; ny = cy; nz = cz;
  }
  return total;
}

// delta3d_multigrid.rs:37-52
template <class T> struct Level3 {
  size_t nx, ny, nz;
  T inv;
  std::vector<T> open_u, open_v, open_w, kind, diag, x, r, t;
  std::vector<T> dd;   // Galerkin levels: Dirichlet part of the diagonal (units of inv)
  bool galerkin = false;
  int storage = 0;     // Options::storage of the vectors x and r
  size_t cells() const { return nx * ny * nz; }
  size_t c(size_t i, size_t j, size_t k) const { return (k * ny + j) * nx + i; }
};

// delta3d_multigrid.rs:209-211
static const int CHILDREN[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
static const int QUARTER[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};

struct RowEntry {
  int64_t n;  // neighbour cell, -1 = none
  float a_f;
  double a_d, it_d, val_d;
};

template <class T> class Volume {
 public:
  // ---------------------------------------------------------------- state
  size_t nx, ny, nz;
  T dx, rho, g, rest;
  bool homogeneous;
  bool cut;  // cut semantics: aperture-weighted divergence (delta3d.rs:806-855)
  std::vector<T> eta, eta_roundoff;
  std::vector<uint8_t> solid;
  std::vector<T> open_u, open_v, open_w;
  std::vector<T> us, vs, ws, u, v, w;
  std::vector<T> p, rhs, res, dir, tmp, prec;
  // coupled surface (delta3d.rs:109-114, set by prepare_background3,
  // delta3d_coupling.rs:97-189): heights of the total surface and the
  // background's ghost terms
  bool coupled = false;
  std::vector<T> surface_total, ghost_bg_up, ghost_bg_x, ghost_bg_y;
  // multigrid (delta3d_multigrid.rs:65-69)
  bool mg_on = false;
  std::vector<Level3<T>> levels;
  std::vector<T> mz, mt;
  Options opt;
  bool pure_neumann = false;

  size_t cells() const { return nx * ny * nz; }
  size_t c(size_t i, size_t j, size_t k) const { return (k * ny + j) * nx + i; }            // delta3d.rs:188
  size_t fu(size_t i, size_t j, size_t k) const { return (k * ny + j) * (nx + 1) + i; }     // delta3d.rs:192
  size_t fv(size_t i, size_t j, size_t k) const { return (k * (ny + 1) + j) * nx + i; }     // delta3d.rs:197
  size_t fw(size_t i, size_t j, size_t k) const { return (k * ny + j) * nx + i; }           // delta3d.rs:201
  size_t col(size_t i, size_t j) const { return j * nx + i; }                               // delta3d.rs:205
  size_t nu() const { return (nx + 1) * ny * nz; }
  size_t nv() const { return nx * (ny + 1) * nz; }
  size_t nw() const { return nx * ny * (nz + 1); }

  Status configure(const Geometry& g_, const Options& o) {
    if (g_.nx <= 0 || g_.ny <= 0 || g_.nz <= 0 || !(g_.dx > 0.f) || !std::isfinite(g_.dx)) return ERR_DOMAIN;
    if (!std::isfinite(g_.rho) || !(g_.rho > 0.f) || !std::isfinite(g_.g_eff)) return ERR_NOT_FINITE;
    nx = g_.nx; ny = g_.ny; nz = g_.nz;
    dx = (T)g_.dx; rho = (T)g_.rho; g = (T)g_.g_eff; rest = (T)g_.rest;
    homogeneous = g_.homogeneous_ghost != 0;
    opt = o;
    size_t cols = nx * ny;
    eta.assign(cols, T(0)); eta_roundoff.assign(cols, T(0));
    for (size_t q = 0; q < cols; ++q) {
      if (!std::isfinite(g_.eta[q])) return ERR_NOT_FINITE;
      eta[q] = (T)g_.eta[q];
      if (g_.eta_roundoff) eta_roundoff[q] = (T)g_.eta_roundoff[q];
    }
    solid.assign(cells(), 0);
    if (g_.solid) for (size_t q = 0; q < cells(); ++q) solid[q] = g_.solid[q] ? 1 : 0;
    cut = g_.solid || g_.open_u || g_.open_v || g_.open_w || g_.closed_lid;
    open_u.assign(nu(), T(1)); open_v.assign(nv(), T(1)); open_w.assign(nw(), T(1));
    if (cut) derive_apertures(g_);
    us.assign(nu(), T(0)); vs.assign(nv(), T(0)); ws.assign(nw(), T(0));
    u = us; v = vs; w = ws;
    p.assign(cells(), T(0)); rhs = p; res = p; dir = p; tmp = p; prec = p;
    mg_on = o.multigrid != 0;
    if (mg_on) build_hierarchy();
    return OK;
  }

  // Apertures when none are given: a face touching a solid cell or the box
  // walls is closed; the lid (top faces) is open unless closed_lid. With the
  // reference's cut geometry (delta3d_cut.rs) a face covered by solid has
  // aperture 0 in the same way.
  void derive_apertures(const Geometry& g_) {
    auto S = [&](size_t i, size_t j, size_t k) { return solid[c(i, j, k)] != 0; };
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i <= nx; ++i) {
          float a = (i == 0 || i == nx || S(i - 1, j, k) || S(i, j, k)) ? 0.f : 1.f;
          open_u[fu(i, j, k)] = g_.open_u ? (T)g_.open_u[fu(i, j, k)] : (T)a;
        }
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j <= ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          float a = (j == 0 || j == ny || S(i, j - 1, k) || S(i, j, k)) ? 0.f : 1.f;
          open_v[fv(i, j, k)] = g_.open_v ? (T)g_.open_v[fv(i, j, k)] : (T)a;
        }
    for (size_t k = 0; k <= nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          float a;
          if (k == 0) a = 0.f;
          else if (k == nz) a = (S(i, j, nz - 1) || g_.closed_lid) ? 0.f : 1.f;
          else a = (S(i, j, k - 1) || S(i, j, k)) ? 0.f : 1.f;
          T val = g_.open_w ? (T)g_.open_w[fw(i, j, k)] : (T)a;
          if (k == nz && g_.closed_lid) val = T(0);
          open_w[fw(i, j, k)] = val;
        }
  }

  // ---------------------------------------------------------------- classification
  T height3(size_t i, size_t j) const {                                                 // delta3d_mobile.rs:92-95
    return coupled ? surface_total[col(i, j)] : eta[col(i, j)];
  }
  bool solid3(size_t cc) const { return solid[cc] != 0; }                               // delta3d_mobile.rs:122-124
  bool wet3(size_t i, size_t j, size_t k) const {                                       // delta3d_mobile.rs:97-103
    if (solid3(c(i, j, k))) return false;
    return ((T)k + T(0.5)) * dx < height3(i, j);
  }
  T open3(int axis, size_t f) const {                                                   // delta3d_mobile.rs:107-119
    if (!cut) return T(1);
    return axis == 0 ? open_u[f] : axis == 1 ? open_v[f] : open_w[f];
  }
  static T theta_min() { return (T)1e-3; }                                               // delta_mobile.rs:14
  static T tmax(T a, T b) { return (std::isnan(a) || a < b) ? b : a; }                  // Rust f32::max

  // delta3d_mobile.rs:126-134
  void ghost_up3(size_t i, size_t j, size_t k, T& it, T& value) const {
    size_t cc = col(i, j);
    T theta = tmax((height3(i, j) - ((T)k + T(0.5)) * dx) / dx, theta_min());
    it = T(1) / theta;
    if (homogeneous) { value = T(0); return; }
    value = rho * g * ((eta[cc] - rest) - eta_roundoff[cc]);
    if (coupled) value += ghost_bg_up[cc];
  }
  // delta3d_mobile.rs:136-147
  void ghost_side3(size_t i, size_t j, size_t k, size_t x, size_t y, T& it, T& value) const {
    T zc = ((T)k + T(0.5)) * dx;
    T h = height3(i, j);
    T theta = tmax((h - zc) / (h - height3(x, y)), theta_min());
    it = T(1) / theta;
    if (homogeneous) { value = T(0); return; }
    value = rho * g * (zc - rest);
    if (coupled) value += i != x ? ghost_bg_x[fu(std::max(i, x), j, k)] : ghost_bg_y[fv(i, std::max(j, y), k)];
  }

  struct Row { int64_t n[6]; T a[6], it[6], val[6]; };

  // delta3d_mobile.rs:152-188 — order x−, x+, y−, y+, z−, z+.
  Row mobile_row(size_t i, size_t j, size_t k) const {
    Row row;
    for (int f = 0; f < 6; ++f) { row.n[f] = -1; row.a[f] = T(0); row.it[f] = T(0); row.val[f] = T(0); }
    struct Nb { bool has; size_t x, y; T a; };
    Nb nb[4] = {
        {i > 0, i > 0 ? i - 1 : 0, j, open3(0, fu(i, j, k))},
        {i + 1 < nx, i + 1, j, open3(0, fu(i + 1, j, k))},
        {j > 0, i, j > 0 ? j - 1 : 0, open3(1, fv(i, j, k))},
        {j + 1 < ny, i, j + 1, open3(1, fv(i, j + 1, k))},
    };
    for (int f = 0; f < 4; ++f) {
      if (!nb[f].has) continue;
      size_t x = nb[f].x, y = nb[f].y;
      T a = nb[f].a;
      if (a == T(0) || solid3(c(x, y, k))) continue;
      if (wet3(x, y, k)) {
        row.n[f] = (int64_t)c(x, y, k); row.a[f] = a;
      } else {
        T it, b;
        ghost_side3(i, j, k, x, y, it, b);
        row.a[f] = a; row.it[f] = it; row.val[f] = b;
      }
    }
    T down = open3(2, fw(i, j, k));
    if (k > 0 && down > T(0) && !solid3(c(i, j, k - 1))) { row.n[4] = (int64_t)c(i, j, k - 1); row.a[4] = down; }
    T up = open3(2, fw(i, j, k + 1));
    if (up > T(0)) {
      if (k + 1 < nz && wet3(i, j, k + 1)) {
        row.n[5] = (int64_t)c(i, j, k + 1); row.a[5] = up;
      } else {
        T it, b;
        ghost_up3(i, j, k, it, b);
        row.a[5] = up; row.it[5] = it; row.val[5] = b;
      }
    }
    return row;
  }

  T inv() const { return T(1) / (dx * dx); }

  // delta3d_mobile.rs:190-213
  void apply_mobile3(const std::vector<T>& x, std::vector<T>& out) const {
    const T iv = inv();
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          size_t cc = c(i, j, k);
          if (!wet3(i, j, k)) { out[cc] = T(0); continue; }
          Row row = mobile_row(i, j, k);
          T acc = T(0);
          for (int f = 0; f < 6; ++f) {
            if (row.n[f] >= 0) acc += row.a[f] * (x[cc] - x[row.n[f]]);
            else if (row.a[f] > T(0)) acc += row.a[f] * x[cc] * row.it[f];
          }
          out[cc] = acc * iv;
        }
  }

  // delta3d_mobile.rs:215-241 (rhs holds the divergence on entry)
  void rhs_mobile3(T scale) {
    const T iv = inv();
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          size_t cc = c(i, j, k);
          if (!wet3(i, j, k)) { rhs[cc] = T(0); prec[cc] = T(0); continue; }
          T b = scale * rhs[cc], diag = T(0);
          Row row = mobile_row(i, j, k);
          for (int f = 0; f < 6; ++f) {
            if (row.n[f] >= 0) diag += row.a[f];
            else if (row.a[f] > T(0)) {
              diag += row.a[f] * row.it[f];
              b += row.a[f] * row.val[f] * row.it[f] * iv;
            }
          }
          rhs[cc] = b;
          prec[cc] = diag > T(0) ? T(1) / (diag * iv) : T(0);
        }
  }

  // delta3d.rs:747-763 (plain) and :806-823 (cut, fixed solids)
  void divergence(const std::vector<T>& uu, const std::vector<T>& vv, const std::vector<T>& ww, std::vector<T>& out) const {
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          size_t cc = c(i, j, k);
          if (cut) {
            if (solid3(cc)) { out[cc] = T(0); continue; }
            T fl = open_u[fu(i, j, k)] * uu[fu(i, j, k)], fr = open_u[fu(i + 1, j, k)] * uu[fu(i + 1, j, k)];
            T ff = open_v[fv(i, j, k)] * vv[fv(i, j, k)], fk = open_v[fv(i, j + 1, k)] * vv[fv(i, j + 1, k)];
            T fb = open_w[fw(i, j, k)] * ww[fw(i, j, k)], ft = open_w[fw(i, j, k + 1)] * ww[fw(i, j, k + 1)];
            out[cc] = ((fr - fl + ft - fb) + (fk - ff)) / dx;
          } else {
            T fl = uu[fu(i, j, k)], fr = uu[fu(i + 1, j, k)];
            T ff = vv[fv(i, j, k)], fk = vv[fv(i, j + 1, k)];
            T fb = ww[fw(i, j, k)], ft = ww[fw(i, j, k + 1)];
            out[cc] = ((fr - fl + ft - fb) + (fk - ff)) / dx;
          }
        }
  }

  // delta3d.rs:860-868 with SequentialJobs (host_impl.rs:76-93): f32 sums of
  // 64-element chunks, merged in index order in f32.
  T dot(const std::vector<T>& a, const std::vector<T>& b) const {
    T accum = T(0);
    size_t n = a.size();
    for (size_t s = 0; s < n; s += 64) {
      size_t e = s + 64 < n ? s + 64 : n;
      T chunk = T(0);
      for (size_t q = s; q < e; ++q) chunk += a[q] * b[q];
      accum = accum + chunk;
    }
    return accum;
  }
  // delta3d.rs:871-876
  Status norm2(const std::vector<T>& a, T& out) const {
    T n = dot(a, a);
    if (!std::isfinite(n)) return ERR_NOT_FINITE;
    if (n == T(0)) for (T x : a) if (x != T(0)) return ERR_NOT_FINITE;
    out = n;
    return OK;
  }
  // delta3d_mobile.rs:447-461
  Status dot_prec3(T& out) const {
    T accum = T(0);
    size_t n = p.size();
    for (size_t s = 0; s < n; s += 64) {
      size_t e = s + 64 < n ? s + 64 : n;
      T chunk = T(0);
      for (size_t q = s; q < e; ++q) chunk += res[q] * prec[q] * res[q];
      accum = accum + chunk;
    }
    if (!std::isfinite(accum)) return ERR_NOT_FINITE;
    out = accum;
    return OK;
  }
  // delta3d_mobile.rs:435-445
  Status prime_mobile3(T beta, T& rz) {
    for (size_t q = 0; q < dir.size(); ++q) {
      T z = prec[q] * res[q];
      dir[q] = beta == T(0) ? z : z + beta * dir[q];
    }
    return dot_prec3(rz);
  }

  // delta3d_mobile.rs:243-271
  T backward_error_mobile3() const {
    const T iv = inv();
    T worst = T(0);
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          if (!wet3(i, j, k)) continue;
          size_t cc = c(i, j, k);
          T pc = std::fabs(p[cc]);
          T acc = T(0);
          Row row = mobile_row(i, j, k);
          for (int f = 0; f < 6; ++f) {
            if (row.n[f] >= 0) acc += row.a[f] * (pc + std::fabs(p[row.n[f]]));
            else if (row.a[f] > T(0)) acc += row.a[f] * pc * row.it[f];
          }
          T scale = std::fabs(rhs[cc]) + acc * iv;
          if (scale > T(0)) worst = tmax(worst, std::fabs(res[cc]) / scale);
        }
    return worst;
  }
  // delta3d_mobile.rs:273-280
  T mobile_roundoff() const { return ny == 1 ? roundoff_2d<T>() : roundoff_3d<T>(); }

  // delta3d_mobile.rs:284-343
  void correct_mobile3(T k1) {
    u = us; v = vs; w = ws;
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i)
          for (int axis = 0; axis < 2; ++axis) {
            if ((axis == 0 && i == 0) || (axis == 1 && j == 0)) continue;
            size_t x = axis == 0 ? i - 1 : i, y = axis == 0 ? j : j - 1;
            size_t l = c(x, y, k), r = c(i, j, k);
            size_t face = axis == 0 ? fu(i, j, k) : fv(i, j, k);
            if (open3(axis, face) == T(0) || solid3(l) || solid3(r)) continue;
            bool wl = wet3(x, y, k), wr = wet3(i, j, k);
            T change;
            if (wl && wr) change = k1 * (p[r] - p[l]) / dx;
            else if (wl && !wr) { T a, b; ghost_side3(x, y, k, i, j, a, b); change = k1 * (b - p[l]) * a / dx; }
            else if (!wl && wr) { T a, b; ghost_side3(i, j, k, x, y, a, b); change = k1 * (p[r] - b) * a / dx; }
            else continue;
            if (axis == 0) u[face] -= change; else v[face] -= change;
          }
    for (size_t k = 1; k <= nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          size_t f = fw(i, j, k);
          if (open3(2, f) == T(0) || !wet3(i, j, k - 1)) continue;
          size_t below = c(i, j, k - 1);
          if (k < nz && wet3(i, j, k)) {
            w[f] -= k1 * (p[c(i, j, k)] - p[below]) / dx;
          } else {
            T a, b;
            ghost_up3(i, j, k - 1, a, b);
            w[f] -= k1 * (b - p[below]) * a / dx;
          }
        }
  }

  // delta3d_mobile.rs:398-433 (PROJECTION_VELOCITY_FLOOR = 1e-4, delta3d.rs:41)
  void divergence_mobile3(double& all_out, double& plain_out, T* all_raw = nullptr, T* plain_raw = nullptr) {
    divergence(u, v, w, tmp);
    T all = T(0), plain = T(0);
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          if (!wet3(i, j, k)) continue;
          T d = std::fabs(tmp[c(i, j, k)]);
          all = tmax(all, d);
          Row row = mobile_row(i, j, k);
          bool ghost = false;
          for (int f = 0; f < 6; ++f) if (row.n[f] < 0 && row.a[f] > T(0)) ghost = true;
          if (!ghost) plain = tmax(plain, d);
        }
    T mx = T(0);
    for (T x : u) mx = tmax(mx, std::fabs(x));
    for (T x : v) mx = tmax(mx, std::fabs(x));
    for (T x : w) mx = tmax(mx, std::fabs(x));
    T echelle = tmax(mx, (T)1e-4);
    T ratio = mx > T(0) ? dx / echelle : T(0);
    all_out = (double)(all * ratio);
    plain_out = (double)(plain * ratio);
    if (all_raw) *all_raw = all;
    if (plain_raw) *plain_raw = plain;
  }

  // ---------------------------------------------------------------- multigrid
  // delta3d_multigrid.rs:109-136 (EXTENSION: max_levels cap)
  void build_hierarchy() {
    levels.clear();
    size_t x = nx, y = ny, z = nz;
    T h = dx;
    while (coarsens3(x, y, z) && (opt.max_levels < 0 || (int)levels.size() < opt.max_levels)) {
      size_t cx = x / 2, cy = y / 2, cz = z / 2;
      h *= T(2);
      Level3<T> L;
      L.nx = cx; L.ny = cy; L.nz = cz;
      L.inv = T(1) / (h * h);
      L.open_u.assign((cx + 1) * cy * cz, T(0));
      L.open_v.assign(cx * (cy + 1) * cz, T(0));
      L.open_w.assign(cx * cy * (cz + 1), T(0));
      size_t n = cx * cy * cz;
      L.kind.assign(n, T(0)); L.diag = L.kind; L.x = L.kind; L.r = L.kind; L.t = L.kind; L.dd = L.kind;
      L.galerkin = opt.coarse_op == 1;
      L.storage = opt.storage;
      levels.push_back(std::move(L));
      x = cx; y = cy; z = cz;
    }
    mz.assign(cells(), T(0));
    mt.assign(cells(), T(0));
  }

  // delta3d_multigrid.rs:216-255
  static void row_sums(const Level3<T>& L, const std::vector<T>* pp, std::vector<T>& out) {
    const size_t nx_ = L.nx, ny_ = L.ny, nz_ = L.nz;
    for (size_t k = 0; k < nz_; ++k)
      for (size_t j = 0; j < ny_; ++j)
        for (size_t i = 0; i < nx_; ++i) {
          size_t cc = L.c(i, j, k);
          if (L.kind[cc] <= T(0)) { out[cc] = T(0); continue; }
          T pc = pp ? (*pp)[cc] : T(0);
          T acc = T(0);
          const bool gal = L.galerkin;
          auto face = [&](T a, int64_t n, bool top) {
            if (a == T(0)) return;
            if (n >= 0 && L.kind[n] > T(0)) acc += pp ? a * (pc - (*pp)[n]) : a;
            else if (gal) {}
            else if (n >= 0 && L.kind[n] < T(0)) acc += T(2) * a * (pp ? pc : T(1));
            else if (n >= 0) {}
            else if (top) acc += T(2) * a * (pp ? pc : T(1));
          };
          face(L.open_u[(k * ny_ + j) * (nx_ + 1) + i], i > 0 ? (int64_t)cc - 1 : -1, false);
          face(L.open_u[(k * ny_ + j) * (nx_ + 1) + i + 1], i + 1 < nx_ ? (int64_t)cc + 1 : -1, false);
          face(L.open_v[(k * (ny_ + 1) + j) * nx_ + i], j > 0 ? (int64_t)(cc - nx_) : -1, false);
          face(L.open_v[(k * (ny_ + 1) + j + 1) * nx_ + i], j + 1 < ny_ ? (int64_t)(cc + nx_) : -1, false);
          face(L.open_w[(k * ny_ + j) * nx_ + i], k > 0 ? (int64_t)(cc - nx_ * ny_) : -1, false);
          face(L.open_w[((k + 1) * ny_ + j) * nx_ + i], k + 1 < nz_ ? (int64_t)(cc + nx_ * ny_) : -1, true);
          if (gal) acc += pp ? L.dd[cc] * pc : L.dd[cc];
          out[cc] = acc * L.inv;
        }
  }

  // delta3d_multigrid.rs:141-206
  template <class OpenF, class KindF>
  static void coarsen3(size_t fnx, size_t fny, OpenF open, KindF kind, Level3<T>& L) {
    const size_t cx = L.nx, cy = L.ny, cz = L.nz;
    auto fc = [&](size_t i, size_t j, size_t k) { return (k * fny + j) * fnx + i; };
    auto fu_ = [&](size_t i, size_t j, size_t k) { return (k * fny + j) * (fnx + 1) + i; };
    auto fv_ = [&](size_t i, size_t j, size_t k) { return (k * (fny + 1) + j) * fnx + i; };
    auto fw_ = [&](size_t i, size_t j, size_t k) { return (k * fny + j) * fnx + i; };
    for (size_t k = 0; k < cz; ++k)
      for (size_t j = 0; j < cy; ++j)
        for (size_t i = 0; i < cx; ++i) {
          bool active = false, air = false;
          for (auto& ch : CHILDREN) {
            T vk = kind(fc(2 * i + ch[0], 2 * j + ch[1], 2 * k + ch[2]));
            if (vk > T(0)) active = true; else if (vk < T(0)) air = true;
          }
          L.kind[L.c(i, j, k)] = active ? T(1) : air ? T(-1) : T(0);
        }
    for (size_t k = 0; k < cz; ++k)
      for (size_t j = 0; j < cy; ++j)
        for (size_t i = 0; i <= cx; ++i) {
          T s = T(0);
          for (auto& q : QUARTER) s += open(0, fu_(2 * i, 2 * j + q[0], 2 * k + q[1]));
          L.open_u[(k * cy + j) * (cx + 1) + i] = T(0.25) * s;
        }
    for (size_t k = 0; k < cz; ++k)
      for (size_t j = 0; j <= cy; ++j)
        for (size_t i = 0; i < cx; ++i) {
          T s = T(0);
          for (auto& q : QUARTER) s += open(1, fv_(2 * i + q[0], 2 * j, 2 * k + q[1]));
          L.open_v[(k * (cy + 1) + j) * cx + i] = T(0.25) * s;
        }
    for (size_t k = 0; k <= cz; ++k)
      for (size_t j = 0; j < cy; ++j)
        for (size_t i = 0; i < cx; ++i) {
          T s = T(0);
          for (auto& q : QUARTER) s += open(2, fw_(2 * i + q[0], 2 * j + q[1], 2 * k));
          L.open_w[(k * cy + j) * cx + i] = T(0.25) * s;
        }
    row_sums(L, nullptr, L.diag);
  }

  // delta3d_multigrid.rs:357-367
  T fine_kind3(size_t cc) const {
    size_t i = cc % nx, j = (cc / nx) % ny, k = cc / (nx * ny);
    if (wet3(i, j, k)) return T(1);
    if (solid3(cc)) return T(0);
    return T(-1);
  }

  // delta3d_multigrid.rs:370-387
  // EXTENSION (coarse_op = 1): Galerkin coarse levels L_{l+1} = P^T L_l P / 16
  // with P the piecewise-constant prolongation onto wet cells. A coarse face
  // carries a quarter of the couplings of the four fine faces it covers
  // (only faces between two active cells couple); the Dirichlet (ghost) part
  // of the diagonal of the eight children is summed and divided by four. In
  // the interior this is the reference's re-discretised operator exactly.
  T ghost_diag(size_t i, size_t j, size_t k) const {
    if (!wet3(i, j, k)) return T(0);
    Row row = mobile_row(i, j, k);
    T g = T(0);
    for (int f = 0; f < 6; ++f)
      if (row.n[f] < 0 && row.a[f] > T(0)) g += row.a[f] * row.it[f];
    return g;
  }
  T fine_coupling(int axis, size_t i, size_t j, size_t k) const {
    // aperture of the face below (i, j, k) along `axis` if it couples two wet cells
    if (axis == 0 && (i == 0 || i >= nx)) return T(0);
    if (axis == 1 && (j == 0 || j >= ny)) return T(0);
    if (axis == 2 && (k == 0 || k >= nz)) return T(0);
    if (!wet3(i, j, k)) return T(0);
    Row row = mobile_row(i, j, k);
    const int f = axis == 0 ? 0 : axis == 1 ? 2 : 4;
    return row.n[f] >= 0 ? row.a[f] : T(0);
  }
  void prepare_galerkin() {
    Level3<T>& L0 = levels[0];
    coarsen3(nx, ny, [&](int, size_t) { return T(1); }, [&](size_t cc) { return fine_kind3(cc); }, L0);
    const size_t cx = L0.nx, cy = L0.ny, cz = L0.nz;
    for (size_t k = 0; k < cz; ++k)
      for (size_t j = 0; j < cy; ++j)
        for (size_t i = 0; i <= cx; ++i) {
          T s = T(0);
          for (auto& q : QUARTER) s += fine_coupling(0, 2 * i, 2 * j + q[0], 2 * k + q[1]);
          L0.open_u[(k * cy + j) * (cx + 1) + i] = T(0.25) * s;
        }
    for (size_t k = 0; k < cz; ++k)
      for (size_t j = 0; j <= cy; ++j)
        for (size_t i = 0; i < cx; ++i) {
          T s = T(0);
          for (auto& q : QUARTER) s += fine_coupling(1, 2 * i + q[0], 2 * j, 2 * k + q[1]);
          L0.open_v[(k * (cy + 1) + j) * cx + i] = T(0.25) * s;
        }
    for (size_t k = 0; k <= cz; ++k)
      for (size_t j = 0; j < cy; ++j)
        for (size_t i = 0; i < cx; ++i) {
          T s = T(0);
          for (auto& q : QUARTER) s += fine_coupling(2, 2 * i + q[0], 2 * j + q[1], 2 * k);
          L0.open_w[(k * cy + j) * cx + i] = T(0.25) * s;
        }
    for (size_t k = 0; k < cz; ++k)
      for (size_t j = 0; j < cy; ++j)
        for (size_t i = 0; i < cx; ++i) {
          T s = T(0);
          for (auto& ch : CHILDREN) s += ghost_diag(2 * i + ch[0], 2 * j + ch[1], 2 * k + ch[2]);
          L0.dd[L0.c(i, j, k)] = T(0.25) * s;
        }
    row_sums(L0, nullptr, L0.diag);
    for (size_t l = 1; l < levels.size(); ++l) {
      const Level3<T>& F = levels[l - 1];
      Level3<T>& L = levels[l];
      coarsen3(F.nx, F.ny,
               [&](int axis, size_t f) { return axis == 0 ? F.open_u[f] : axis == 1 ? F.open_v[f] : F.open_w[f]; },
               [&](size_t cc) { return F.kind[cc]; }, L);
      for (size_t k = 0; k < L.nz; ++k)
        for (size_t j = 0; j < L.ny; ++j)
          for (size_t i = 0; i < L.nx; ++i) {
            T s = T(0);
            for (auto& ch : CHILDREN) s += F.dd[F.c(2 * i + ch[0], 2 * j + ch[1], 2 * k + ch[2])];
            L.dd[L.c(i, j, k)] = T(0.25) * s;
          }
      row_sums(L, nullptr, L.diag);
    }
  }

  void prepare_multigrid3() {
    if (levels.empty()) return;
    if (opt.coarse_op == 1) { prepare_galerkin(); return; }
    coarsen3(nx, ny, [&](int axis, size_t f) { return open3(axis, f); }, [&](size_t cc) { return fine_kind3(cc); }, levels[0]);
    for (size_t l = 1; l < levels.size(); ++l) {
      const Level3<T>& F = levels[l - 1];
      coarsen3(F.nx, F.ny,
               [&](int axis, size_t f) { return axis == 0 ? F.open_u[f] : axis == 1 ? F.open_v[f] : F.open_w[f]; },
               [&](size_t cc) { return F.kind[cc]; }, levels[l]);
    }
  }

  // delta3d_multigrid.rs:258-272
  static void restrict3(size_t fnx, size_t fny, const std::vector<T>& fine, Level3<T>& L) {
    for (size_t k = 0; k < L.nz; ++k)
      for (size_t j = 0; j < L.ny; ++j)
        for (size_t i = 0; i < L.nx; ++i) {
          T s = T(0);
          for (auto& ch : CHILDREN) s += fine[((2 * k + ch[2]) * fny + 2 * j + ch[1]) * fnx + 2 * i + ch[0]];
          L.r[L.c(i, j, k)] = round_storage(L.storage, T(0.125) * s);
        }
  }
  // delta3d_multigrid.rs:275-287
  static void prolong_add3(size_t fnx, size_t fny, std::vector<T>& fine, const Level3<T>& L) {
    for (size_t k = 0; k < L.nz; ++k)
      for (size_t j = 0; j < L.ny; ++j)
        for (size_t i = 0; i < L.nx; ++i) {
          T val = L.x[L.c(i, j, k)];
          for (auto& ch : CHILDREN) {
            T& f = fine[((2 * k + ch[2]) * fny + 2 * j + ch[1]) * fnx + 2 * i + ch[0]];
            f = round_storage(L.storage, f + val);
          }
        }
  }

  static T damping() { return T(6) / T(7); }  // delta3d_multigrid.rs:28

  // delta3d_multigrid.rs:290-299
  static void smooth3(Level3<T>& L) {
    row_sums(L, &L.x, L.t);
    for (size_t cc = 0; cc < L.cells(); ++cc)
      if (L.diag[cc] > T(0)) L.x[cc] = round_storage(L.storage, L.x[cc] + damping() * (L.r[cc] - L.t[cc]) / L.diag[cc]);
  }
  // EXTENSION: one colour of a Gauss-Seidel sweep on a coarse level. Cells of
  // one colour do not couple, so the update uses the other colour's values.
  static void gs_colour(Level3<T>& L, int colour) {
    row_sums(L, &L.x, L.t);
    for (size_t k = 0; k < L.nz; ++k)
      for (size_t j = 0; j < L.ny; ++j)
        for (size_t i = 0; i < L.nx; ++i) {
          if ((int)((i + j + k) & 1) != colour) continue;
          size_t cc = L.c(i, j, k);
          if (L.diag[cc] > T(0)) L.x[cc] = round_storage(L.storage, L.x[cc] + (L.r[cc] - L.t[cc]) / L.diag[cc]);
        }
  }
  // A sweep: Jacobi (reference) or one red-black pass; `forward` = red then
  // black (pre-smoothing), else black then red (post-smoothing, the adjoint).
  void level_sweep(Level3<T>& L, bool forward) const {
    if (opt.smoother == 0) { smooth3(L); return; }
    gs_colour(L, forward ? 0 : 1);
    gs_colour(L, forward ? 1 : 0);
  }

  // delta3d_multigrid.rs:303-332
  void coarse_cycle3() {
    size_t last = levels.size() - 1;
    for (size_t l = 0; l <= last; ++l) {
      Level3<T>& L = levels[l];
      std::fill(L.x.begin(), L.x.end(), T(0));
      if (l == last) {
        if (opt.smoother == 0) {
          for (int s = 0; s < opt.coarse_sweeps; ++s) smooth3(L);
        } else {
          // palindromic red-black sequence so the coarsest correction is symmetric
          int half = opt.coarse_sweeps / 2;
          for (int s = 0; s < half; ++s) level_sweep(L, true);
          if (opt.coarse_sweeps & 1) gs_colour(L, 0);
          for (int s = 0; s < half; ++s) level_sweep(L, false);
        }
      } else {
        const int pre = opt.level_pre >= 0 ? opt.level_pre : opt.pre_sweeps;
        for (int s = 0; s < pre; ++s) level_sweep(L, true);
      }
      if (l < last) {
        row_sums(L, &L.x, L.t);
        for (size_t cc = 0; cc < L.cells(); ++cc) L.t[cc] = L.r[cc] - L.t[cc];
        restrict3(L.nx, L.ny, L.t, levels[l + 1]);
      }
    }
    for (size_t l = last; l-- > 0;) {
      Level3<T>& L = levels[l];
      prolong_add3(L.nx, L.ny, L.x, levels[l + 1]);
      const int post = opt.level_post >= 0 ? opt.level_post : opt.post_sweeps;
      for (int s = 0; s < post; ++s) level_sweep(L, false);
    }
  }

  // Fine-level sweep on mz with the exact operator and prec
  // (delta3d_multigrid.rs:398-405, :420-427; RBGS is the EXTENSION).
  void fine_sweep(bool forward) {
    if (opt.smoother == 0) {
      apply_mobile3(mz, mt);
      for (size_t cc = 0; cc < cells(); ++cc)
        if (prec[cc] > T(0)) mz[cc] = round_storage(opt.storage, mz[cc] + damping() * prec[cc] * (res[cc] - mt[cc]));
      return;
    }
    for (int pass = 0; pass < 2; ++pass) {
      int colour = forward ? pass : 1 - pass;
      apply_mobile3(mz, mt);
      for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
          for (size_t i = 0; i < nx; ++i) {
            if ((int)((i + j + k) & 1) != colour) continue;
            size_t cc = c(i, j, k);
            if (prec[cc] > T(0)) mz[cc] = round_storage(opt.storage, mz[cc] + prec[cc] * (res[cc] - mt[cc]));
          }
    }
  }

  // delta3d_multigrid.rs:393-429
  void v_cycle3() {
    std::fill(mz.begin(), mz.end(), T(0));
    for (int s = 0; s < opt.pre_sweeps; ++s) fine_sweep(true);
    if (!levels.empty()) {
      apply_mobile3(mz, mt);
      for (size_t cc = 0; cc < cells(); ++cc) mt[cc] = res[cc] - mt[cc];
      restrict3(nx, ny, mt, levels[0]);
      coarse_cycle3();
      prolong_add3(nx, ny, mz, levels[0]);
      for (size_t cc = 0; cc < cells(); ++cc) if (prec[cc] == T(0)) mz[cc] = T(0);
    }
    for (int s = 0; s < opt.post_sweeps; ++s) fine_sweep(false);
  }

  // delta3d_multigrid.rs:433-445
  Status prime_multigrid3(T beta, T& rz) {
    v_cycle3();
    rz = dot(res, mz);
    for (size_t q = 0; q < dir.size(); ++q) dir[q] = beta == T(0) ? mz[q] : mz[q] + beta * dir[q];
    if (!std::isfinite(rz)) return ERR_NOT_FINITE;
    return OK;
  }
  // delta3d_multigrid.rs:448-457
  Status precondition_multigrid3(T& rz) {
    v_cycle3();
    rz = dot(res, mz);
    if (!std::isfinite(rz)) return ERR_NOT_FINITE;
    return OK;
  }
  // delta3d_multigrid.rs:460-465
  void direction_multigrid3(T beta) {
    for (size_t q = 0; q < dir.size(); ++q) dir[q] = mz[q] + beta * dir[q];
  }

  Status prime(T& rz) {
    if (mg_on) return prime_multigrid3(T(0), rz);
    return prime_mobile3(T(0), rz);
  }

  // EXTENSION: a problem with no Dirichlet face anywhere (no ghost face on
  // the fine level) is singular with the constants as null space.
  bool detect_pure_neumann() const {
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          if (!wet3(i, j, k)) continue;
          Row row = mobile_row(i, j, k);
          for (int f = 0; f < 6; ++f) if (row.n[f] < 0 && row.a[f] > T(0)) return false;
        }
    return true;
  }
  void remove_wet_mean(std::vector<T>& x) const {
    double s = 0; size_t n = 0;
    for (size_t cc = 0; cc < cells(); ++cc) if (prec[cc] > T(0)) { s += (double)x[cc]; ++n; }
    if (!n) return;
    T m = (T)(s / (double)n);
    for (size_t cc = 0; cc < cells(); ++cc) if (prec[cc] > T(0)) x[cc] -= m;
  }

  // delta3d_mobile.rs:465-609. On entry us/vs/ws hold the predicted faces and
  // p the warm start.
  Status project_mobile3(T scale, T k1, uint32_t max_iters, Report& rep) {
    if (!gnodes.empty()) return project_mobile3_graded(scale, k1, max_iters, rep);  // delta3d_mobile.rs:473-475
    divergence(us, vs, ws, rhs);
    rhs_mobile3(scale);
    if (mg_on) prepare_multigrid3();
    pure_neumann = opt.gauge && detect_pure_neumann();
    if (pure_neumann) remove_wet_mean(rhs);
    T b2;
    Status st = norm2(rhs, b2);
    if (st) return st;
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i)
          if (!wet3(i, j, k)) p[c(i, j, k)] = T(0);
    apply_mobile3(p, tmp);
    for (size_t q = 0; q < res.size(); ++q) res[q] = rhs[q] - tmp[q];
    T rr;
    if ((st = norm2(res, rr))) return st;
    T rz = T(0);
    bool primed = false;
    uint32_t it = 0;
    // rr > 1e-12 b2 (:507, :515); another target only through the rel_tol extension
    const T tol = opt.rel_tol == 1e-6 ? (T)1e-12 : (T)(opt.rel_tol * opt.rel_tol);
    bool floor_stop = false;
    T physical_target = std::numeric_limits<T>::infinity();
    bool have_checkpoint = false;
    uint64_t checkpoint = 0;
    uint32_t power = 1, since = 0;
    T actual_rr;
    bool settled = false;
    double set_all = 0, set_plain = 0;
    for (;;) {
      uint32_t before = it;
      while (b2 > T(0) && (rr > tol * b2 || rr > physical_target) && it < max_iters) {
        if (!primed) {
          if ((st = prime(rz))) return st;
          primed = true;
        }
        apply_mobile3(dir, tmp);
        T dq = dot(dir, tmp);
        if (!(dq > T(0))) break;
        T alpha = rz / dq;
        for (size_t q = 0; q < p.size(); ++q) {
          p[q] += alpha * dir[q];
          res[q] -= alpha * tmp[q];
        }
        T rn;
        if ((st = norm2(res, rn))) return st;
        if (mg_on) {
          T zn;
          if ((st = precondition_multigrid3(zn))) return st;
          if (opt.flexible) {
            // <z', r' - r> = -alpha <z', q> since r' = r - alpha q
            const T qz = dot(tmp, mz);
            direction_multigrid3(-(alpha * qz) / rz);
          } else {
            direction_multigrid3(zn / rz);
          }
          rz = zn;
        } else {
          T zn;
          if ((st = dot_prec3(zn))) return st;
          T beta = zn / rz;
          T unused;
          if ((st = prime_mobile3(beta, unused))) return st;
          rz = zn;
        }
        rr = rn;
        ++it;
      }
      apply_mobile3(p, tmp);
      for (size_t q = 0; q < res.size(); ++q) res[q] = rhs[q] - tmp[q];
      T actual;
      if ((st = norm2(res, actual))) return st;
      bool exhausted = it >= max_iters || it == before;
      bool have_measured = false;
      double m_all = 0, m_plain = 0;
      if (actual <= tol * b2) {
        if (exhausted) { actual_rr = actual; break; }
        if (pure_neumann) remove_wet_mean(p);
        correct_mobile3(k1);
        double r_all, r_plain;
        divergence_mobile3(r_all, r_plain);
        if (r_plain <= 1e-5) { actual_rr = actual; settled = true; set_all = r_all; set_plain = r_plain; break; }
        have_measured = true; m_all = r_all; m_plain = r_plain;
        T ratio = (T)(1e-5 / r_plain);
        physical_target = actual * ratio * ratio;
      } else if (exhausted) {
        actual_rr = actual;
        break;
      }
      if (backward_error_mobile3() <= mobile_roundoff()) {
        floor_stop = true;
        actual_rr = actual;
        if (have_measured) { settled = true; set_all = m_all; set_plain = m_plain; }
        break;
      }
      uint64_t state = fingerprint(p);
      if (have_checkpoint && checkpoint == state) {
        floor_stop = true;
        actual_rr = actual;
        if (have_measured) { settled = true; set_all = m_all; set_plain = m_plain; }
        break;
      }
      ++since;
      if (since == power) {
        have_checkpoint = true;
        checkpoint = state;
        power = power > 0x7fffffffu ? 0xffffffffu : power * 2;
        since = 0;
      }
      rr = actual;
      if ((st = prime(rz))) return st;
    }
    T residual = b2 > T(0) ? std::sqrt(actual_rr / b2) : T(0);
    double d_all, d_plain;
    if (settled) { d_all = set_all; d_plain = set_plain; }
    else {
      if (pure_neumann) remove_wet_mean(p);
      correct_mobile3(k1);
      divergence_mobile3(d_all, d_plain);
    }
    bool accepted = (actual_rr <= tol * b2 || floor_stop) && d_plain <= 1e-5;
    rep.refinements = 0;
    rep.iterations = it;
    rep.degraded = (b2 > T(0) && !accepted) ? 1 : 0;
    rep.residual = (double)residual;
    rep.divergence = d_all;
    rep.floor = floor_stop ? 1 : 0;
    rep.backward_error = (double)backward_error_mobile3();
    rep.divergence_plain = d_plain;
    return OK;
  }

  // delta3d_coupling.rs:226-249: the corrected faces as the predicted ones,
  // a solve from p = 0 with homogeneous ghosts, added to the base pressure.
  Status refine_mobile3(T scale, T k1, uint32_t max_iters, Report& rep) {
    const std::vector<T> base = p;
    us = u; vs = v; ws = w;
    std::fill(p.begin(), p.end(), T(0));
    const bool saved = homogeneous;
    homogeneous = true;
    const Status st = project_mobile3(scale, k1, max_iters, rep);
    homogeneous = saved;
    if (st) return st;
    for (size_t q = 0; q < p.size(); ++q) p[q] += base[q];
    rep.refinements = 1;
    return OK;
  }

  // ---------------------------------------------------------------- the graded column
  // delta3d_graded.rs: below the cubic top layers the pressure lives only at
  // node layers (the same for every column) and is linear between two nodes;
  // the projection solves P^T A P x = P^T b (SPD), preconditioned by the
  // condensed diagonal sum_k P_kj^2 A_kk, and the velocities stay fine.
  std::vector<size_t> gnodes;  // empty: the fine scheme
  std::vector<T> gweight;      // sum_k P_kj per node

  // delta3d_graded.rs:48-53: (1 - t, t), t = (k - a) / (b - a)
  static void gweights(size_t k, size_t a, size_t b, T& wa, T& wb) {
    const T t = (T)(k - a) / (T)(b - a);
    wa = T(1) - t;
    wb = t;
  }
  // enable_graded (delta3d_graded.rs:64-100)
  Status enable_graded(const size_t* nodes, size_t count) {
    if (count == 0 || nodes[0] != 0 || nodes[count - 1] != nz - 1) return ERR_SHAPE;
    for (size_t q = 1; q < count; ++q)
      if (nodes[q] <= nodes[q - 1]) return ERR_SHAPE;
    if (cut) return ERR_DOMAIN;
    gnodes.assign(nodes, nodes + count);
    gweight.assign(count, T(0));
    for (size_t s = 0; s + 1 < count; ++s)
      for (size_t k = gnodes[s]; k < gnodes[s + 1]; ++k) {
        T wa, wb;
        gweights(k, gnodes[s], gnodes[s + 1], wa, wb);
        gweight[s] += wa;
        gweight[s + 1] += wb;
      }
    gweight[count - 1] += T(1);
    return OK;
  }
  size_t graded_len() const { return nx * ny * gnodes.size(); }
  // fine = P x (delta3d_graded.rs:109-126)
  void graded_expand(const std::vector<T>& x, std::vector<T>& fine) const {
    const size_t cols = nx * ny, last = gnodes.size() - 1;
    for (size_t s = 0; s < last; ++s)
      for (size_t k = gnodes[s]; k < gnodes[s + 1]; ++k) {
        T wa, wb;
        gweights(k, gnodes[s], gnodes[s + 1], wa, wb);
        for (size_t cc = 0; cc < cols; ++cc) fine[k * cols + cc] = wa * x[s * cols + cc] + wb * x[(s + 1) * cols + cc];
      }
    for (size_t cc = 0; cc < cols; ++cc) fine[gnodes[last] * cols + cc] = x[last * cols + cc];
  }
  // out = P^T fine, the reference's summation order (delta3d_graded.rs:128-145);
  // square: out = sum_k (w w) fine_k, the condensed diagonal's weights (:363-377)
  void graded_restrict(const std::vector<T>& fine, std::vector<T>& out, bool square = false) const {
    const size_t cols = nx * ny, last = gnodes.size() - 1;
    std::fill(out.begin(), out.end(), T(0));
    for (size_t s = 0; s < last; ++s)
      for (size_t k = gnodes[s]; k < gnodes[s + 1]; ++k) {
        T wa, wb;
        gweights(k, gnodes[s], gnodes[s + 1], wa, wb);
        if (square) { wa = wa * wa; wb = wb * wb; }
        for (size_t cc = 0; cc < cols; ++cc) {
          out[s * cols + cc] += wa * fine[k * cols + cc];
          out[(s + 1) * cols + cc] += wb * fine[k * cols + cc];
        }
      }
    for (size_t cc = 0; cc < cols; ++cc) out[last * cols + cc] += fine[gnodes[last] * cols + cc];
  }
  // out = P^T A P x through the fine grid (:147-162), the mobile operator
  void graded_apply(const std::vector<T>& x, std::vector<T>& out) {
    graded_expand(x, dir);
    apply_mobile3(dir, tmp);
    graded_restrict(tmp, out);
  }
  static double dot_f64(const std::vector<T>& a, const std::vector<T>& b) {
    double s = 0.0;
    for (size_t q = 0; q < a.size(); ++q) s += (double)a[q] * (double)b[q];
    return s;
  }
  // the lowest cubic node: from it the nodes are consecutive to the top (:300-307)
  size_t cubic_node() const {
    size_t j = gnodes.size() - 1;
    while (j > 0 && gnodes[j - 1] + 1 == gnodes[j]) --j;
    return j;
  }
  // graded_divergence_mobile (:451-486): P^T div u over each node's weight,
  // on wet node cells; plain rows: graded nodes and the lowest cubic node, and
  // above it the rows without a ghost face
  void graded_divergence_mobile(size_t j0, std::vector<T>& restricted, double& all_out, double& plain_out) {
    const size_t cols = nx * ny;
    divergence(u, v, w, tmp);
    graded_restrict(tmp, restricted);
    T all = T(0), plain = T(0);
    for (size_t s = 0; s < gnodes.size(); ++s) {
      const size_t k = gnodes[s];
      for (size_t cc = 0; cc < cols; ++cc) {
        const size_t i = cc % nx, j = cc / nx;
        if (!wet3(i, j, k)) continue;
        const T d = std::fabs(restricted[s * cols + cc]) / gweight[s];
        all = tmax(all, d);
        bool ghost = false;
        if (s > j0) {
          Row row = mobile_row(i, j, k);
          for (int f = 0; f < 6; ++f) if (row.n[f] < 0 && row.a[f] > T(0)) ghost = true;
        }
        if (!ghost) plain = tmax(plain, d);
      }
    }
    T mx = T(0);
    for (T x : u) mx = tmax(mx, std::fabs(x));
    for (T x : v) mx = tmax(mx, std::fabs(x));
    for (T x : w) mx = tmax(mx, std::fabs(x));
    const T ratio = mx > T(0) ? dx / tmax(mx, (T)1e-4) : T(0);
    all_out = (double)(all * ratio);
    plain_out = (double)(plain * ratio);
  }
  // project_mobile3_graded (delta3d_graded.rs:310-447); p holds the warm start
  Status project_mobile3_graded(T scale, T k1, uint32_t max_iters, Report& rep) {
    const size_t cols = nx * ny, n = graded_len(), j0 = cubic_node(), k0 = gnodes[j0];
    bool inside = k0 + 1 < nz;
    for (size_t j = 0; j < ny && inside; ++j)
      for (size_t i = 0; i < nx && inside; ++i)
        if (!wet3(i, j, k0) || !wet3(i, j, k0 + 1)) inside = false;
    if (!inside) return ERR_DOMAIN;
    divergence(us, vs, ws, rhs);
    rhs_mobile3(scale);
    std::vector<T> gx(n), gr(n), gd(n), gq(n), gb(n), gm(n), gz(n);
    graded_restrict(rhs, gb);
    // condensed diagonal: A_kk = 1 / prec_k on a wet cell, 0 on a dry one
    for (size_t cc = 0; cc < cells(); ++cc) tmp[cc] = prec[cc] > T(0) ? T(1) / prec[cc] : T(0);
    graded_restrict(tmp, gm, true);
    // warm start by injection; a dry node starts from 0
    for (size_t s = 0; s < gnodes.size(); ++s)
      for (size_t cc = 0; cc < cols; ++cc)
        gx[s * cols + cc] = gm[s * cols + cc] > T(0) ? p[gnodes[s] * cols + cc] : T(0);
    const double b2 = dot_f64(gb, gb);
    const double tol = opt.rel_tol == 1e-6 ? 1e-12 : opt.rel_tol * opt.rel_tol;
    double target = tol * b2, rr = 0.0;
    uint32_t it = 0;
    bool floor_stop = false, have_ck = false;
    uint64_t ck = 0;
    uint32_t power = 1, since = 0;
    double d_all = 0.0, d_plain = 0.0;
    for (;;) {
      // the true residual, then PCG to the target
      graded_apply(gx, gq);
      for (size_t q = 0; q < n; ++q) gr[q] = gb[q] - gq[q];
      rr = dot_f64(gr, gr);
      for (size_t q = 0; q < n; ++q) {
        gz[q] = gm[q] > T(0) ? gr[q] / gm[q] : T(0);
        gd[q] = gz[q];
      }
      double rz = dot_f64(gr, gz);
      while (b2 > 0.0 && rr > target && it < max_iters) {
        graded_apply(gd, gq);
        const double dq = dot_f64(gd, gq);
        if (!(dq > 0.0)) break;
        const T alpha = (T)(rz / dq);
        for (size_t q = 0; q < n; ++q) {
          gx[q] += alpha * gd[q];
          gr[q] -= alpha * gq[q];
          gz[q] = gm[q] > T(0) ? gr[q] / gm[q] : T(0);
        }
        const double zn = dot_f64(gr, gz);
        const T beta = (T)(zn / rz);
        for (size_t q = 0; q < n; ++q) gd[q] = gz[q] + beta * gd[q];
        rz = zn;
        rr = dot_f64(gr, gr);
        ++it;
      }
      graded_apply(gx, gq);
      for (size_t q = 0; q < n; ++q) gr[q] = gb[q] - gq[q];
      graded_expand(gx, p);
      rr = dot_f64(gr, gr);
      correct_mobile3(k1);
      graded_divergence_mobile(j0, gz, d_all, d_plain);
      const bool converged = rr <= tol * b2;
      if (b2 == 0.0 || (converged && d_plain <= 1e-5) || it >= max_iters) break;
      const uint64_t state = fingerprint(gx);
      if (have_ck && ck == state) { floor_stop = true; break; }
      ++since;
      if (since == power) {
        have_ck = true;
        ck = state;
        power = power > 0x7fffffffu ? 0xffffffffu : power * 2;
        since = 0;
      }
      if (converged) {
        const double ratio = 1e-5 / std::max(d_plain, std::numeric_limits<double>::min());
        target = rr * ratio * ratio;
      } else {
        target = tol * b2;
      }
    }
    const bool accepted = b2 == 0.0 || ((rr <= tol * b2 || floor_stop) && d_plain <= 1e-5);
    rep.refinements = 0;
    rep.iterations = it;
    rep.degraded = accepted ? 0 : 1;
    rep.residual = b2 > 0.0 ? std::sqrt(rr / b2) : 0.0;
    rep.divergence = d_all;
    rep.floor = floor_stop ? 1 : 0;
    rep.backward_error = 0.0;
    rep.divergence_plain = d_plain;
    return OK;
  }

  // ---------------------------------------------------------------- the surface step
  // The steps either side of the projection in step_surface_mobile
  // (delta3d_mobile.rs:819-927). Scope: no sparse column set
  // (sparse_outside3 is false, delta3d_sparse.rs:306-309) and the second-order
  // advection term off (its default, delta3d_advection.rs:31-36).
  std::vector<T> flux_x, flux_y;  // (nx+1)*ny and nx*(ny+1) wetted-column discharges

  static T clamp01(T x) { return x < T(0) ? T(0) : (x > T(1) ? T(1) : x); }  // Rust f32::clamp (NaN passes)

  // delta3d_mobile.rs:615-732: centred MAC advection of u, v, w into us, vs,
  // ws (Euler step). Closed faces and the faces on the box walls keep their
  // value; the top w face is not advected.
  void advect_mobile3(T dt) {
    const T h = T(0.5) / dx;
    us = u; vs = v; ws = w;
    for (int axis = 0; axis < 2; ++axis) {
      const size_t na = axis == 0 ? nx : ny, nb = axis == 0 ? ny : nx;
      auto val = [&](size_t a, size_t b, size_t k) { return axis == 0 ? u[fu(a, b, k)] : v[fv(b, a, k)]; };
      auto vertical = [&](size_t a, size_t b, size_t k) { return axis == 0 ? w[fw(a, b, k)] : w[fw(b, a, k)]; };
      auto transverse = [&](size_t a, size_t b, size_t k) { return axis == 0 ? v[fv(a, b, k)] : u[fu(b, a, k)]; };
      for (size_t b = 0; b < nb; ++b)
        for (size_t a = 1; a < na; ++a)
          for (size_t k = 0; k < nz; ++k) {
            const size_t i = axis == 0 ? a : b, j = axis == 0 ? b : a;
            const size_t face = axis == 0 ? fu(i, j, k) : fv(i, j, k);
            if (open3(axis, face) == T(0)) continue;
            const T uc = val(a, b, k);
            const T ux = (val(a + 1, b, k) - val(a - 1, b, k)) * h;
            const T up = k + 1 < nz ? val(a, b, k + 1) : uc;
            const T dn = k > 0 ? val(a, b, k - 1) : uc;
            const T uz = (up - dn) * h;
            const T wc = T(0.25) * (vertical(a - 1, b, k) + vertical(a, b, k) + vertical(a - 1, b, k + 1) +
                                    vertical(a, b, k + 1));
            const T rt = b + 1 < nb ? val(a, b + 1, k) : uc;
            const T lf = b > 0 ? val(a, b - 1, k) : uc;
            const T uy = (rt - lf) * h;
            const T vc = T(0.25) * (transverse(a - 1, b, k) + transverse(a, b, k) + transverse(a - 1, b + 1, k) +
                                    transverse(a, b + 1, k));
            const T next = uc - dt * (uc * ux + wc * uz + vc * uy);
            if (axis == 0) us[face] = next;
            else vs[face] = next;
          }
    }
    for (size_t j = 0; j < ny; ++j)
      for (size_t i = 0; i < nx; ++i)
        for (size_t k = 1; k < nz; ++k) {
          const size_t f = fw(i, j, k);
          if (open3(2, f) == T(0)) continue;
          const T wc = w[f];
          const T rt = i + 1 < nx ? w[fw(i + 1, j, k)] : wc;
          const T lf = i > 0 ? w[fw(i - 1, j, k)] : wc;
          const T wx = (rt - lf) * h;
          const T wz = (w[fw(i, j, k + 1)] - w[fw(i, j, k - 1)]) * h;
          const T uc = T(0.25) * (u[fu(i, j, k - 1)] + u[fu(i + 1, j, k - 1)] + u[fu(i, j, k)] + u[fu(i + 1, j, k)]);
          const T bk = j + 1 < ny ? w[fw(i, j + 1, k)] : wc;
          const T fr = j > 0 ? w[fw(i, j - 1, k)] : wc;
          const T wy = (bk - fr) * h;
          const T vc = T(0.25) * (v[fv(i, j, k - 1)] + v[fv(i, j + 1, k - 1)] + v[fv(i, j, k)] + v[fv(i, j + 1, k)]);
          ws[f] = wc - dt * (uc * wx + wc * wz + vc * wy);
        }
  }

  // delta3d_mobile.rs:345-394: up every column, each open face that is not
  // solved (open, between two non-solid cells, at least one of them wet; a w
  // face: open above a wet cell) takes the velocity of the last solved face
  // below it. Closed faces are skipped and do not interrupt the run.
  void extrapolate_mobile3() {
    for (size_t j = 0; j < ny; ++j)
      for (size_t i = 0; i < nx; ++i)
        for (int axis = 0; axis < 3; ++axis) {
          if ((axis == 0 && i == 0) || (axis == 1 && j == 0)) continue;
          bool have = false;
          T last = T(0);
          const size_t k0 = axis == 2 ? 1 : 0, k1 = axis == 2 ? nz + 1 : nz;
          for (size_t k = k0; k < k1; ++k) {
            size_t f;
            bool solved;
            if (axis == 0) {
              f = fu(i, j, k);
              solved = open3(0, f) > T(0) && !solid3(c(i - 1, j, k)) && !solid3(c(i, j, k)) &&
                       (wet3(i - 1, j, k) || wet3(i, j, k));
            } else if (axis == 1) {
              f = fv(i, j, k);
              solved = open3(1, f) > T(0) && !solid3(c(i, j - 1, k)) && !solid3(c(i, j, k)) &&
                       (wet3(i, j - 1, k) || wet3(i, j, k));
            } else {
              f = fw(i, j, k);
              solved = open3(2, f) > T(0) && wet3(i, j, k - 1);
            }
            if (open3(axis, f) == T(0)) continue;
            std::vector<T>& field = axis == 0 ? u : axis == 1 ? v : w;
            if (solved) { last = field[f]; have = true; }
            else if (have) field[f] = last;
          }
        }
  }

  // delta3d_mobile.rs:734-782: wetted discharges through the inner vertical
  // faces from eta^n (every discharge before any height changes), then each
  // column's height moves by -transport * (flux divergence) with the
  // compensated (Kahan-style) roundoff carried in eta_roundoff.
  void transport_mobile3(T transport) {
    flux_x.assign((nx + 1) * ny, T(0));
    flux_y.assign(nx * (ny + 1), T(0));
    for (size_t j = 0; j < ny; ++j)
      for (size_t i = 0; i < nx; ++i)
        for (int axis = 0; axis < 2; ++axis) {
          if ((axis == 0 && i == 0) || (axis == 1 && j == 0)) continue;
          const size_t x = axis == 0 ? i - 1 : i, y = axis == 0 ? j : j - 1;
          const T surface = T(0.5) * (eta[col(x, y)] + eta[col(i, j)]);
          T q = T(0);
          for (size_t k = 0; k < nz; ++k) {
            const T wet = clamp01((surface - (T)k * dx) / dx);
            if (wet == T(0)) break;
            T vel, a;
            if (axis == 0) { const size_t f = fu(i, j, k); vel = u[f]; a = open3(0, f); }
            else { const size_t f = fv(i, j, k); vel = v[f]; a = open3(1, f); }
            q += a * vel * dx * wet;
          }
          if (axis == 0) flux_x[j * (nx + 1) + i] = q;
          else flux_y[j * nx + i] = q;
        }
    for (size_t j = 0; j < ny; ++j)
      for (size_t i = 0; i < nx; ++i) {
        const size_t cc = col(i, j);
        const T x = flux_x[j * (nx + 1) + i + 1] - flux_x[j * (nx + 1) + i];
        const T y = flux_y[(j + 1) * nx + i] - flux_y[j * nx + i];
        const T increment = -transport * (x + y) - eta_roundoff[cc];
        const T height = eta[cc] + increment;
        eta_roundoff[cc] = (height - eta[cc]) - increment;
        eta[cc] = height;
      }
  }

  // delta3d_balance.rs:105-113
  double perturbation_volume() const {
    const double area = (double)dx * (double)dx, r = (double)rest;
    double total = 0.0;
    for (size_t q = 0; q < eta.size(); ++q) total += ((double)eta[q] - (double)eta_roundoff[q]) - r;
    return total * area;
  }

  // delta3d_mobile.rs:784-800: discharge through the four outer walls,
  // positive inwards, in m^3 over dt (zero by construction here).
  double boundary_flux_mobile3(double dt) const {
    double total = 0.0;
    for (size_t j = 0; j < ny; ++j) total += (double)flux_x[j * (nx + 1)] - (double)flux_x[j * (nx + 1) + nx];
    for (size_t i = 0; i < nx; ++i) total += (double)flux_y[i] - (double)flux_y[ny * nx + i];
    return total * dt * (double)dx;
  }

  // delta3d_mobile.rs:801-813 with the floor of delta3d_cut.rs:429-447: the
  // surface stays at least two cells above the top of the highest solid cell
  // of its column (the flat bottom at 0 otherwise) and at most at (nz-1) dx.
  bool mobile_in_bounds() const {
    const T top = (T)(nz - 1) * dx;
    for (size_t j = 0; j < ny; ++j)
      for (size_t i = 0; i < nx; ++i) {
        T floor = T(0);
        for (size_t k = 0; k < nz; ++k)
          if (solid3(c(i, j, k))) { const T t = (T)(k + 1) * dx; if (t > floor) floor = t; }
        const T e = height3(i, j);
        if (!(e >= floor + T(2) * dx && e <= top)) return false;
      }
    return true;
  }

  // delta3d_balance.rs Balance3 of the uncoupled step (delta3d_mobile.rs:866-877)
  struct Balance { double volume, delta, band_in, perturbation_in, sponge_out, residual, outgoing; };

  static bool all_finite(const std::vector<T>& f) {
    for (const T& x : f) if (!std::isfinite(x)) return false;
    return true;
  }

  // delta3d_mobile.rs:819-927: advect, project (warm start from p), extrapolate,
  // transport, balance; a refusal restores u, v, w, p, eta and eta_roundoff.
  Status step_surface_mobile(uint64_t duration_us, uint32_t max_iters, Report& rep, Balance& bal) {
    if (duration_us == 0 || duration_us > (1ull << 53)) return ERR_NOT_FINITE;
    const double dt = (double)duration_us * 1e-6;
    const float dxf = (float)dx, rhof = (float)rho, gf = (float)g;
    if (gf <= 0.f || dt * dt * (double)gf / (double)dxf > 1.0 || !mobile_in_bounds()) return ERR_DOMAIN;
    const float scale = (float)(-(double)rhof / dt);
    const float correction = (float)(dt / (double)rhof);
    const float transport = (float)(dt / (double)dxf);
    const float advection = (float)dt;
    if (!std::isfinite(scale) || !std::isfinite(correction) || correction == 0.f || !std::isfinite(transport) ||
        transport == 0.f || !(advection > 0.f))
      return ERR_NOT_FINITE;
    const std::vector<T> su = u, sv = v, sw = w, sp = p, se = eta, sr = eta_roundoff;
    const double volume_before = perturbation_volume();
    auto body = [&]() -> Status {
      advect_mobile3((T)advection);
      Status st = project_mobile3((T)scale, (T)correction, max_iters, rep);
      if (st) return st;
      if (rep.degraded) return ERR_CONVERGENCE;
      extrapolate_mobile3();
      transport_mobile3((T)transport);
      const double volume = perturbation_volume();
      const double delta = volume - volume_before;
      const double pin = boundary_flux_mobile3(dt);
      for (const std::vector<T>* f : {&u, &v, &w, &p, &eta, &eta_roundoff, &us, &vs, &ws, &rhs, &res, &dir, &tmp, &prec})
        if (!all_finite(*f)) return ERR_NOT_FINITE;
      if (!std::isfinite(rep.residual) || !std::isfinite(rep.divergence)) return ERR_NOT_FINITE;
      if (!mobile_in_bounds()) return ERR_DOMAIN;
      bal = Balance{volume, delta, 0.0, pin, 0.0, delta - pin, 0.0};
      return OK;
    };
    const Status st = body();
    if (st) { u = su; v = sv; w = sw; p = sp; eta = se; eta_roundoff = sr; }
    return st;
  }
};

}  // namespace orc
