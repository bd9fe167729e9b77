// C ABI and host driver of the MI355X pressure projection.
//
// The driver restates the control flow of the reference's
// Volume3::project_mobile3 (code/water-core/src/delta3d_mobile.rs:465-609):
// assembly, warm start, conjugate-gradient passes, true residual recomputed
// after each pass, acceptance on the plain-row divergence with target
// tightening, floor stops on the backward error or a repeated state of p.
// The iterations themselves run on the device: every scalar of the CG
// (alpha, beta, the loop test) is computed by the last block of the kernel
// that produced its reduction, so the host only polls a small control block
// between batches of iterations.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/fsp.h"
#include "../../include/fsp_learned.h"
#include "kernels.h"
#include "kernels_fast.h"
#include "kernels_coarse.h"
#include "kernels_step.h"
#include "kernels_graded.h"
#include "learned_internal.h"

using namespace fsp;

struct fsp_solver {
  fsp_domain dom{};
  fsp_options opt{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  size_t n = 0, nu = 0, nv = 0, nw = 0, ncols = 0;
  bool geom = false;
  int cut = 0;
  fsp_surface surf{};
  uint8_t *solid = nullptr, *ou = nullptr, *ov = nullptr, *ow = nullptr, *mask = nullptr;
  // fractional apertures (code 2 in ou/ov/ow), reserved by the first
  // fsp_set_geometry that has one
  float *au = nullptr, *av = nullptr, *aw = nullptr;
  int frac = 0;
  // level-set geometry (fsp_set_level_set): water labels and phi, reserved
  // by its first call
  uint8_t* water = nullptr;
  float* phi = nullptr;
  float theta_min = 0.f;
  int levelset = 0;
  // coupled surface (fsp_set_coupling): the total surface per column and the
  // background's ghost terms, reserved by its first call
  float *total = nullptr, *bg_up = nullptr, *bg_x = nullptr, *bg_y = nullptr;
  int coupled = 0;
  // graded column (fsp_enable_graded): node layers, per-layer segment and
  // weights and per-node weights on the device, the reduced vectors (columns
  // x nodes each), partials and folded scalars; reserved by its call
  std::vector<int> gnodes;  // empty: the fine scheme
  int *g_node = nullptr, *g_seg = nullptr;
  float *g_wa = nullptr, *g_wb = nullptr, *g_weight = nullptr;
  float *gx = nullptr, *gr = nullptr, *gd = nullptr, *gq = nullptr, *gb = nullptr, *gm = nullptr, *gz = nullptr;
  size_t gcap = 0;
  double *gpart = nullptr, *gout = nullptr, *hgout = nullptr;
  float *eta = nullptr, *eta_rt = nullptr;
  float* col_floor = nullptr;  // per column: top of its highest solid cell (0 without one)
  uint16_t* gk0 = nullptr;  // per column: lowest ghost row
  float* gtab = nullptr;    // per column: 1/theta of G_SLOTS ghost rows
  float* precv = nullptr;   // per cell: row_prec (quad V-cycle)
  float *b = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *zt = nullptr, *d0 = nullptr,
        *d1 = nullptr;
  float *us = nullptr, *vs = nullptr, *ws = nullptr, *u = nullptr, *v = nullptr, *w = nullptr;
  double* p64 = nullptr;  // the pressure, accumulated in f64
  std::vector<Level> levels;
  // fast paths: quad kernels on the fine level (nx % 4 == 0), one-workgroup
  // tail for the coarse levels from tail_first on
  bool quad = false;
  QuadGrid qfine{}, qrestr{};
  BrickGrid brick{};  // fused red-black sweeps on bricks (bx == 0: off)
  std::vector<QuadGrid> cq, cqr;  // compact levels: quads of the level, and of its restriction
  int tail_first = 0;
  size_t tail_lds = 0;
  float* tail_pack = nullptr;  // Galerkin tail: couplings of the tail levels in LDS order
  float* zres = nullptr;  // buffer holding M^-1 r after precondition()
  Ctl* ctl = nullptr;
  double* partials = nullptr;
  unsigned long long *hashes = nullptr, *hash_out = nullptr;
  Ctl* hctl = nullptr;
  unsigned long long* hhash = nullptr;
  std::vector<void*> allocs;
  // live timing of one kernel (fsp_profile_begin / fsp_profile_end)
  // one CG iteration captured as a graph, replayed per iteration
  // one CG iteration per (phase, kernel form): [2 * phase + table_only],
  // phase 0 as configured, 1 the learned phase
  hipGraph_t iter_graph[4] = {};
  hipGraphExec_t iter_exec[4] = {};
  int table_only = 0;                 // 1: the fine kernels' table-only forms (set per projection)
  fsp_net* net = nullptr;             // learned preconditioner (not owned)
  uint64_t net_gen = 0;               // the network's buffer generation iter_graph[1] captured
  int net_phase = 0;                  // 1 while the current pass uses the network
  uint32_t net_until = 0;             // iteration at which the learned phase ends
  // surface step (allocated on its first call): saved state, discharges, scalars
  float *sv_u = nullptr, *sv_v = nullptr, *sv_w = nullptr, *sv_p = nullptr, *sv_eta = nullptr, *sv_rt = nullptr;
  float *flux_x = nullptr, *flux_y = nullptr;
  StepScalars* step_sc = nullptr;
  StepScalars* hstep = nullptr;
  bool prof_on = false;
  std::string prof_name;
  std::vector<hipEvent_t> prof_events;
  size_t prof_used = 0;
  int32_t* prof_active = nullptr;  // per timed launch: the loop-active flag it started with

  GradedCol graded_col() const {
    GradedCol D;
    D.cols = (int)ncols;
    D.nn = (int)gnodes.size();
    D.nz = dom.nz;
    D.top = gnodes.empty() ? 0 : gnodes.back();
    D.node = g_node;
    D.seg = g_seg;
    D.wa = g_wa;
    D.wb = g_wb;
    return D;
  }

  Fine fine() const {
    Fine G;
    G.nx = dom.nx; G.ny = dom.ny; G.nz = dom.nz;
    G.cut = cut;
    G.n = n;
    G.dx = dom.dx;
    G.inv = 1.0f / (dom.dx * dom.dx);
    G.rho = surf.rho; G.g = surf.g_eff; G.rest = surf.rest;
    G.homogeneous = surf.homogeneous_ghost;
    G.solid = solid; G.ou = ou; G.ov = ov; G.ow = ow;
    G.frac = frac;
    G.au = frac ? au : nullptr; G.av = frac ? av : nullptr; G.aw = frac ? aw : nullptr;
    G.water = levelset ? water : nullptr;
    G.phi = levelset ? phi : nullptr;
    G.theta_min = theta_min;
    G.eta = coupled ? total : eta; G.eta_rt = eta_rt;
    G.eta_val = eta;
    G.bg_up = coupled ? bg_up : nullptr;
    G.bg_x = coupled ? bg_x : nullptr;
    G.bg_y = coupled ? bg_y : nullptr;
    G.mask = mask;
    G.gk0 = gk0;
    G.gtab = gtab;
    G.precv = precv;
    // prec of a plain row with n faces, 1/(n * inv), as rhs_mobile3 computes it
    G.prec_tab[0] = 0.f;
    for (int q = 1; q < 7; ++q) G.prec_tab[q] = 1.0f / ((float)q * G.inv);
    return G;
  }
};

namespace {

#define FSP_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) {                            \
      std::fprintf(stderr, "fsp: %s failed: %s\n", #call, hipGetErrorString(e_)); \
      return FSP_DEVICE;                               \
    }                                                  \
  } while (0)

// One thread per item up to the partial-sum capacity (for kernels whose
// per-item work is a chain of dependent loads).
unsigned grid_wide(size_t n) {
  size_t b = (n + BLOCK - 1) / BLOCK;
  if (b > (size_t)MAX_PARTIALS) b = MAX_PARTIALS;
  if (b == 0) b = 1;
  return (unsigned)b;
}

unsigned grid_for(size_t n) {
  size_t b = (n + BLOCK - 1) / BLOCK;
  if (b > (size_t)MAX_BLOCKS) b = MAX_BLOCKS;
  if (b == 0) b = 1;
  return (unsigned)b;
}

// One item per thread for a grid-stride kernel that reduces nothing (no
// per-block partial): every item's loads are in flight at once instead of
// a thread walking several items one after the other (the advection's
// gathers: 55-62 -> 47-52 us; plain streams ran slower so, 17.5 -> 20 us).
unsigned grid_items(size_t n) {
  size_t b = (n + BLOCK - 1) / BLOCK;
  if (b > (size_t)(1u << 24)) b = 1u << 24;
  if (b == 0) b = 1;
  return (unsigned)b;
}

// The loop-active flag a timed launch starts with (stream order: the value
// the kernel itself reads), recorded before the launch.
__global__ void k_prof_flag(const Ctl* ctl, int32_t* out) {
  if (threadIdx.x == 0) *out = ctl->active;
}

// When `name` is the kernel being timed (fsp_profile_begin), records its
// loop-active flag and hands out the event pair of its launch: the launch
// then goes through hipExtLaunchKernelGGL, whose events take the dispatch's
// own start and end timestamps (no marker packets around it).
inline bool prof_begin(fsp_solver* S, const char* name, hipStream_t stream, hipEvent_t* e0, hipEvent_t* e1) {
  const bool on = S->prof_on && S->prof_name == name && S->prof_used + 2 <= S->prof_events.size();
  if (on) {
    k_prof_flag<<<1, 64, 0, stream>>>(S->ctl, S->prof_active + S->prof_used / 2);
    *e0 = S->prof_events[S->prof_used];
    *e1 = S->prof_events[S->prof_used + 1];
    S->prof_used += 2;
  }
  return on;
}

// Launch with BLOCK threads (FSP_LAUNCH_NB: with `block`, FSP_LAUNCH_NS: with
// `lds` bytes of dynamic LDS); `name` is what fsp_profile_begin matches (the
// kernel without its storage argument).
#define FSP_LAUNCH_NBS(S, name, kern, grid, block, lds, stream, ...)                            \
  do {                                                                                         \
    hipEvent_t e0_ = nullptr, e1_ = nullptr;                                                   \
    if (prof_begin(S, name, stream, &e0_, &e1_))                                               \
      hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), (uint32_t)(lds), stream, e0_, e1_, 0, \
                            __VA_ARGS__);                                                      \
    else                                                                                       \
      kern<<<grid, dim3(block), lds, stream>>>(__VA_ARGS__);                                   \
  } while (0)
#define FSP_LAUNCH_N(S, name, kern, grid, stream, ...) \
  FSP_LAUNCH_NBS(S, name, kern, grid, BLOCK, 0, stream, __VA_ARGS__)
#define FSP_LAUNCH_NB(S, name, kern, grid, block, stream, ...) \
  FSP_LAUNCH_NBS(S, name, kern, grid, block, 0, stream, __VA_ARGS__)
#define FSP_LAUNCH_NS(S, name, kern, grid, lds, stream, ...) \
  FSP_LAUNCH_NBS(S, name, kern, grid, BLOCK, lds, stream, __VA_ARGS__)
#define FSP_LAUNCH(S, kern, grid, stream, ...) FSP_LAUNCH_N(S, #kern, kern, grid, stream, __VA_ARGS__)

// A fine-level kernel in its table-only form (template argument GEN false:
// every ghost row of the geometry lies in its column's table, known after
// the projection's first sync) or its general form.
#define FSP_LAUNCH_GEN(S, name, KT, KG, grid, block, stream, ...)                             \
  do {                                                                                     \
    if ((S)->table_only) FSP_LAUNCH_NBS(S, name, KT, grid, block, 0, stream, __VA_ARGS__); \
    else FSP_LAUNCH_NBS(S, name, KG, grid, block, 0, stream, __VA_ARGS__);                 \
  } while (0)

// Kernels templated on the V-cycle storage ST of the enclosing function.
#define K_CGS (k_cgs<ST>)
#define K_CRESTRICT (k_crestrict<ST>)
#define K_CPROLONG (k_cprolong<ST>)
#define K_GCGS(M) (k_gc_gs<M, ST>)
template <int ST> using TS = typename Store<ST>::T;

// Calls f with the storage type of `storage` (FSP_STORE_*) as a constant.
template <class F>
auto with_storage(int storage, F&& f) {
  switch (storage) {
    case FSP_STORE_F16: return f(std::integral_constant<int, ST_F16>{});
    case FSP_STORE_BF16: return f(std::integral_constant<int, ST_BF16>{});
    default: return f(std::integral_constant<int, ST_F32>{});
  }
}

// Fold the per-block partials of a reducing kernel (k_finish) right after it.
void launch_finish(hipStream_t st, int op, unsigned nblocks, bool is_max, int gated, Ctl* ctl, double* partials) {
  if (is_max) hipLaunchKernelGGL(k_finish<true>, dim3(1), dim3(FIN_THREADS), 0, st, op, nblocks, gated, ctl, partials);
  else hipLaunchKernelGGL(k_finish<false>, dim3(1), dim3(FIN_THREADS), 0, st, op, nblocks, gated, ctl, partials);
}
#define FSP_FINISH(S, stream, op, grid, is_max, gated) \
  launch_finish(stream, op, (grid).x, is_max, gated, (S)->ctl, (S)->partials)

template <class T>
fsp_status dalloc(fsp_solver* S, T** out, size_t count) {
  void* ptr = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 256;
  FSP_HIP(hipMalloc(&ptr, bytes));
  S->allocs.push_back(ptr);
  *out = (T*)ptr;
  return FSP_OK;
}

bool coarsens3(size_t nx, size_t ny, size_t nz) {
  return nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0 && nx >= 4 && ny >= 4 && nz >= 4;
}

__global__ void k_reset(Ctl* ctl, int gauge) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned char* bytes = (unsigned char*)ctl;
  for (size_t q = 0; q < sizeof(Ctl); ++q) bytes[q] = 0;
  ctl->first_dir = 1;
  ctl->physical_target = __longlong_as_double(0x7ff0000000000000ll);
  ctl->gauge = gauge;
}

__global__ void __launch_bounds__(BLOCK) k_fill_gated(float* x, size_t n, float value, Ctl* ctl) {
  if (!ctl->active) return;
  FSP_FOR(c, n) x[c] = value;
}
__global__ void __launch_bounds__(BLOCK) k_copy_gated(const float* a, float* b, size_t n, Ctl* ctl) {
  if (!ctl->active) return;
  FSP_FOR(c, n) b[c] = a[c];
}

// The tail of Galerkin levels with red-black smoothing runs in the compact
// form (couplings and dd only).
// Caller buffers read or written with 16-byte vector accesses must be 16-byte aligned.
inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

bool gc_tail_ok(const fsp_options& o) {
  return o.coarse_operator == FSP_COARSE_GALERKIN && o.smoother == FSP_SMOOTH_RBGS;
}

// The fine V-cycle runs on quads (red-black smoothing, at least one pre- and
// one post-sweep).
bool quad_vcycle_ok(bool quad, const fsp_options& o) {
  return quad && o.smoother == FSP_SMOOTH_RBGS && o.pre_sweeps >= 1 && o.post_sweeps >= 1;
}

struct Launch {
  fsp_solver* S;
  hipStream_t st;
  Fine G;
  explicit Launch(fsp_solver* s) : S(s), st(s->stream), G(s->fine()) {}

  // prec: also the per-row factors k_q_gs_first reads (k_prec), unless the
  // caller's assembly writes them.
  void mask(bool prec = true) {
    const dim3 g(grid_for(G.n));
    if (S->levelset) {
      // labels and level set: no column structure, no ghost-row table
      FSP_LAUNCH(S, k_mask_ls, g, st, G, S->mask, S->ctl, S->partials);
      FSP_FINISH(S, st, OP_COUNT, g, false, 0);
      if (S->precv) FSP_LAUNCH(S, k_prec, g, st, G, S->precv);
      return;
    }
    if (S->quad) {
      const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
      FSP_LAUNCH(S, k_mask_q, gq, st, G, S->qfine, S->mask, S->ctl, S->partials);
      FSP_FINISH(S, st, OP_COUNT, gq, false, 0);
    } else {
      FSP_LAUNCH(S, k_mask, g, st, G, S->mask, S->ctl, S->partials);
      FSP_FINISH(S, st, OP_COUNT, g, false, 0);
    }
    // ghost-row table from the fresh row codes (k_mask's finish does not touch it)
    Fine Gm = G;
    FSP_LAUNCH(S, k_ghost_table, dim3(grid_for(S->ncols)), st, Gm, S->gk0, S->gtab, S->ctl);
    if (S->precv && prec) FSP_LAUNCH(S, k_prec, g, st, Gm, S->precv);
  }

  // A level whose V-cycle kernels read only couplings and Dirichlet parts:
  // a compact level, or a level of the Galerkin tail.
  bool kindless(int l) const {
    return S->levels[l].compact || (gc_tail_ok(S->opt) && l >= S->tail_first);
  }

  // Coarse geometry of every level (prepare_multigrid3, delta3d_multigrid.rs:370-387).
  // Cell kinds and diagonals are built only up to the last level whose
  // kernels read them (kinds of a level derive from those of the level above).
  void hierarchy() {
    int last_generic = -1;
    for (int l = 0; l < (int)S->levels.size(); ++l)
      if (!kindless(l)) last_generic = l;
    for (size_t l = 0; l < S->levels.size(); ++l) {
      Level& L = S->levels[l];
      const bool kinds = (int)l <= last_generic, diag = !kindless((int)l);
      const size_t nf = (size_t)(L.nx + 1) * L.ny * L.nz + (size_t)L.nx * (L.ny + 1) * L.nz +
                        (size_t)L.nx * L.ny * (L.nz + 1);
      if (l == 0) {
        if (kinds) FSP_LAUNCH(S, k_ckind_fine, dim3(grid_for(L.n)), st, G, L);
        if (L.galerkin) {
          FSP_LAUNCH(S, k_cfaces_fine_gal, dim3(grid_for(nf)), st, G, L);
          FSP_LAUNCH(S, k_cdd_fine, dim3(grid_for(L.n)), st, G, L);
        } else {
          FSP_LAUNCH(S, k_cfaces_fine, dim3(grid_for(nf)), st, G, L);
        }
      } else {
        if (kinds) FSP_LAUNCH(S, k_ckind, dim3(grid_for(L.n)), st, S->levels[l - 1], L);
        if (L.galerkin) FSP_LAUNCH(S, k_cgal, dim3(grid_for(std::max(nf, L.n))), st, S->levels[l - 1], L);
        else FSP_LAUNCH(S, k_cfaces, dim3(grid_for(nf)), st, S->levels[l - 1], L);
      }
      if (diag) FSP_LAUNCH(S, k_cdiag, dim3(grid_for(L.n)), st, L);
      if (L.compact) FSP_LAUNCH(S, k_gc_build, dim3(grid_for(L.n)), st, L);
    }
    if (S->tail_pack) {
      const TailSet T = tail_set();
      FSP_LAUNCH(S, k_gct_pack, dim3(gct_pack_blocks(T)), st, T, S->tail_pack);
    }
  }

  TailSet tail_set() const {
    TailSet T{};
    T.count = (int)S->levels.size() - S->tail_first;
    for (int l = 0; l < T.count; ++l) T.lv[l] = S->levels[S->tail_first + l];
    T.pack = S->tail_pack;
    T.smoother = S->opt.smoother;
    T.pre = S->opt.pre_sweeps;
    T.post = S->opt.post_sweeps;
    T.coarse = S->opt.coarse_sweeps;
    return T;
  }

  // r = b - A p64 and |r|^2 into `op`, on quads where the grid has them.
  // replacement: inside the CG loop, only in the iterations that replace the
  // recurrence residual by the true one (replace_every).
  void residual(const double* p64, int op, bool replacement = false) {
    const int gated = replacement ? 1 : 0, fin_gate = replacement ? 3 : 0;
    if (S->quad) {
      const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
      FSP_LAUNCH_GEN(S, "k_residual_q", k_residual_q<false>, k_residual_q<true>, gq, QBLOCK, st, G, S->qfine, p64,
                     S->b, S->r, op, S->ctl, S->partials, gated);
      FSP_FINISH(S, st, op, gq, false, fin_gate);
    } else {
      const dim3 g(grid_for(G.n));
      FSP_LAUNCH(S, k_residual, g, st, G, p64, S->b, S->r, op, S->ctl, S->partials, gated);
      FSP_FINISH(S, st, op, g, false, fin_gate);
    }
  }

  // The true residual into OP_RR with the backward error of the same r into
  // OP_BACKWARD, in one pass over p64 (quads only).
  void residual_backward(const double* p64) {
    const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
    FSP_LAUNCH_GEN(S, "k_residual_q", (k_residual_q<false, true>), (k_residual_q<true, true>), gq, QBLOCK, st, G,
                   S->qfine, p64, S->b, S->r, OP_RR, S->ctl, S->partials, 0);
    FSP_FINISH(S, st, OP_RR, gq, false, 0);
    FSP_FINISH(S, st, OP_BACKWARD, gq, true, 0);
  }

  void fill_gated(float* x, size_t n, float v) {
    FSP_LAUNCH(S, k_fill_gated, dim3(grid_for(n)), st, x, n, v, S->ctl);
  }

  // Sweeps on coarse level l; `x` is the current iterate (nullptr = zero),
  // returns the buffer holding the result. Jacobi sweeps exist in f32 only.
  template <int ST>
  TS<ST>* coarse_sweeps(int l, TS<ST>* x, int count, bool forward, bool coarsest) {
    Level& L = S->levels[l];
    TS<ST>* lx = reinterpret_cast<TS<ST>*>(L.x);
    const int smoother = S->opt.smoother;
    if (count == 0) {
      if (!x) { fill_gated(L.x, L.n, 0.f); return lx; }
      return x;
    }
    if constexpr (ST == ST_F32) {
      if (smoother == FSP_SMOOTH_JACOBI) {
        float* cur = x;
        for (int s = 0; s < count; ++s) {
          float* out = (cur == L.x) ? L.t : L.x;
          FSP_LAUNCH(S, k_cjacobi, dim3(grid_for(L.n)), st, L, cur, out, S->ctl);
          cur = out;
        }
        return cur;
      }
      if (x && x != L.x) FSP_LAUNCH(S, k_copy_gated, dim3(grid_for(L.n)), st, x, L.x, L.n, S->ctl);
    }
    // red-black Gauss-Seidel, in place on L.x
    int first = x ? 0 : 1;
    auto colour = [&](int c) {
      FSP_LAUNCH_N(S, "k_cgs", K_CGS, dim3(grid_for(L.n)), st, L, lx, c, first, S->ctl);
      first = 0;
    };
    if (coarsest) {
      const int half = count / 2;
      for (int s = 0; s < half; ++s) { colour(0); colour(1); }
      if (count & 1) colour(0);
      for (int s = 0; s < half; ++s) { colour(1); colour(0); }
    } else {
      for (int s = 0; s < count; ++s) {
        if (forward) { colour(0); colour(1); } else { colour(1); colour(0); }
      }
    }
    return lx;
  }

  // The tail levels in one workgroup: level tail_first's r is restricted;
  // leaves its x in levels[tail_first].x.
  template <int ST>
  void tail() {
    const TailSet T = tail_set();
    if (gc_tail_ok(S->opt)) k_gc_tail<ST><<<dim3(1), dim3(TAIL_THREADS), S->tail_lds, st>>>(T, S->ctl);
    else k_tail<ST><<<dim3(1), dim3(TAIL_THREADS), S->tail_lds, st>>>(T, S->ctl);
  }

  static dim3 qgrid(const QuadGrid& Q) { return dim3((unsigned)(Q.nbx * Q.nby * Q.planes)); }

  // Pre-smoothing of a compact level from zero, in place on L.x: forward
  // red-black sweeps, or the palindromic sequence on the coarsest level.
  template <int ST>
  TS<ST>* gc_sweeps(int l, int count, bool coarsest) {
    Level& L = S->levels[l];
    TS<ST>* x = reinterpret_cast<TS<ST>*>(L.x);
    if (count == 0) { fill_gated(L.x, L.n, 0.f); return x; }
    const dim3 g = qgrid(S->cq[l]);
    const TS<ST>* none = nullptr;
    bool first = true;
    auto colour = [&](int c) {
      if (first) FSP_LAUNCH_N(S, "k_gc_gs<1>", K_GCGS(1), g, st, L, S->cq[l], (const TS<ST>*)x, x, c, L, none, S->ctl);
      else FSP_LAUNCH_N(S, "k_gc_gs<0>", K_GCGS(0), g, st, L, S->cq[l], (const TS<ST>*)x, x, c, L, none, S->ctl);
      first = false;
    };
    if (coarsest) {
      const int half = count / 2;
      for (int s = 0; s < half; ++s) { colour(0); colour(1); }
      if (count & 1) colour(0);
      for (int s = 0; s < half; ++s) { colour(1); colour(0); }
    } else {
      // the first sweep's two passes in one launch
      FSP_LAUNCH_N(S, "k_gc_gs_first", k_gc_gs_first<ST>, g, st, L, S->cq[l], x, S->ctl);
      first = false;
      for (int s = 1; s < count; ++s) { colour(0); colour(1); }
    }
    return x;
  }

  // Prolongation from the coarser level's iterate `xp` fused with the first
  // post-smoothing pass (L.x -> L.t), then the remaining passes on L.t.
  template <int ST>
  TS<ST>* gc_ascend(int l, const TS<ST>* xp) {
    Level& L = S->levels[l];
    Level& C = S->levels[l + 1];
    TS<ST>* x = reinterpret_cast<TS<ST>*>(L.x);
    TS<ST>* t = reinterpret_cast<TS<ST>*>(L.t);
    const dim3 g = qgrid(S->cq[l]);
    const TS<ST>* none = nullptr;
    const int passes = 2 * S->opt.post_sweeps;
    FSP_LAUNCH_N(S, "k_gc_gs<2>", K_GCGS(2), g, st, L, S->cq[l], (const TS<ST>*)x, t, 1, C, xp, S->ctl);
    for (int pz = 1; pz < passes; ++pz) {
      const int colour = (pz % 2 == 0) ? 1 : 0;
      FSP_LAUNCH_N(S, "k_gc_gs<0>", K_GCGS(0), g, st, L, S->cq[l], (const TS<ST>*)t, t, colour, L, none, S->ctl);
    }
    return t;
  }

  // A compact level with a coarser level below it, small enough to be bound by
  // launch latency, is smoothed in tiles (k_gb_pre / k_gb_post) when its sweep
  // counts fit the tile's halo.
  bool tiled(int l) const {
    const Level& L = S->levels[l];
    return L.compact && l + 1 < (int)S->levels.size() && L.n <= (size_t)FSP_GB_MAX_CELLS &&
           S->opt.pre_sweeps >= 1 && S->opt.pre_sweeps <= GB_MAX_SWEEPS && S->opt.post_sweeps >= 1 &&
           S->opt.post_sweeps <= GB_MAX_SWEEPS;
  }

  // coarse_cycle3 (delta3d_multigrid.rs:303-332); returns the level-0 iterate.
  template <int ST>
  TS<ST>* coarse_cycle() {
    const int last = (int)S->levels.size() - 1;
    const int tf = S->tail_first;
    std::vector<TS<ST>*> xcur(S->levels.size(), nullptr);
    for (int l = 0; l < tf; ++l) {
      Level& L = S->levels[l];
      if (tiled(l)) {
        if (S->opt.pre_sweeps == 2)
          FSP_LAUNCH_NB(S, "k_gb_pre", (k_gb_pre<ST, 2>), dim3(gb_blocks(L)), GB_THREADS, st, L, S->levels[l + 1],
                        S->ctl);
        else
          FSP_LAUNCH_NB(S, "k_gb_pre", (k_gb_pre<ST, 1>), dim3(gb_blocks(L)), GB_THREADS, st, L, S->levels[l + 1],
                        S->ctl);
        xcur[l] = reinterpret_cast<TS<ST>*>(L.x);
        continue;
      }
      const int sweeps = (l == last) ? S->opt.coarse_sweeps : S->opt.pre_sweeps;
      if (L.compact) xcur[l] = gc_sweeps<ST>(l, sweeps, l == last);
      else xcur[l] = coarse_sweeps<ST>(l, nullptr, sweeps, true, l == last);
      if (l < last) {
        Level& C = S->levels[l + 1];
        if (L.compact) {
          // the restriction's grid for 128-thread halves (k_gc_restrict2)
          QuadGrid H = S->cqr[l];
          H.RY = (QBLOCK / 2) / H.QX;
          H.nby = (H.rows + H.RY - 1) / H.RY;
          FSP_LAUNCH_N(S, "k_gc_restrict", (k_gc_restrict2<ST>), qgrid(H), st, L, H, (const TS<ST>*)xcur[l], C,
                       S->ctl);
        } else {
          FSP_LAUNCH_N(S, "k_crestrict", K_CRESTRICT, dim3(grid_for(C.n)), st, L, C, (const TS<ST>*)xcur[l], S->ctl);
        }
      }
    }
    if (tf <= last) {
      tail<ST>();
      xcur[tf] = reinterpret_cast<TS<ST>*>(S->levels[tf].x);
    }
    for (int l = std::min(tf, last) - 1; l >= 0; --l) {
      Level& L = S->levels[l];
      Level& C = S->levels[l + 1];
      if (tiled(l)) {
        TS<ST>* t = reinterpret_cast<TS<ST>*>(L.t);
        if (S->opt.post_sweeps == 2)
          FSP_LAUNCH_NB(S, "k_gb_post", (k_gb_post<ST, 2>), dim3(gb_blocks(L)), GB_THREADS, st, L, C,
                        (const TS<ST>*)xcur[l], (const TS<ST>*)xcur[l + 1], t, S->ctl);
        else
          FSP_LAUNCH_NB(S, "k_gb_post", (k_gb_post<ST, 1>), dim3(gb_blocks(L)), GB_THREADS, st, L, C,
                        (const TS<ST>*)xcur[l], (const TS<ST>*)xcur[l + 1], t, S->ctl);
        xcur[l] = t;
        continue;
      }
      if (L.compact) {
        xcur[l] = gc_ascend<ST>(l, xcur[l + 1]);
        continue;
      }
      FSP_LAUNCH_N(S, "k_cprolong", K_CPROLONG, dim3(grid_for(L.n)), st, L, C, (const TS<ST>*)xcur[l + 1], xcur[l],
                   S->ctl);
      xcur[l] = coarse_sweeps<ST>(l, xcur[l], S->opt.post_sweeps, false, false);
    }
    return xcur[0];
  }

  // Fine sweeps; `cur` is the current iterate (nullptr = zero); returns the result buffer.
  float* fine_sweeps(float* cur, int count, bool forward) {
    if (count == 0) {
      if (!cur) { fill_gated(S->z, G.n, 0.f); return S->z; }
      return cur;
    }
    if (S->opt.smoother == FSP_SMOOTH_JACOBI) {
      for (int s = 0; s < count; ++s) {
        float* out = (cur == S->z) ? S->zt : S->z;
        FSP_LAUNCH(S, k_fjacobi, dim3(grid_for(G.n)), st, G, cur, out, S->r, S->ctl);
        cur = out;
      }
      return cur;
    }
    int first = cur ? 0 : 1;
    if (cur && cur != S->z) {
      FSP_LAUNCH(S, k_copy_gated, dim3(grid_for(G.n)), st, cur, S->z, G.n, S->ctl);
    }
    for (int s = 0; s < count; ++s) {
      const int c0 = forward ? 0 : 1;
      FSP_LAUNCH(S, k_fgs, dim3(grid_for(G.n)), st, G, S->z, S->r, c0, first, S->ctl);
      first = 0;
      FSP_LAUNCH(S, k_fgs, dim3(grid_for(G.n)), st, G, S->z, S->r, 1 - c0, 0, S->ctl);
    }
    return S->z;
  }

  // z = M^-1 r (v_cycle3, delta3d_multigrid.rs:393-429) into S->z; f32 storage.
  void vcycle() {
    // With Jacobi sweeps the iterate alternates between z and zt; start so
    // that an even total ends in z.
    float* cur = nullptr;
    cur = fine_sweeps(nullptr, S->opt.pre_sweeps, true);
    if (!S->levels.empty()) {
      Level& L0 = S->levels[0];
      FSP_LAUNCH(S, k_restrict_fine, dim3(grid_for(L0.n)), st, G, L0, cur, S->r, S->ctl);
      float* x0 = coarse_cycle<ST_F32>();
      FSP_LAUNCH(S, k_prolong_fine, dim3(grid_for(G.n)), st, G, L0, x0, cur, S->ctl);
    }
    cur = fine_sweeps(cur, S->opt.post_sweeps, false);
    if (cur != S->z) {
      FSP_LAUNCH(S, k_copy_gated, dim3(grid_for(G.n)), st, cur, S->z, G.n, S->ctl);
    }
  }

  bool quad_vcycle() const { return quad_vcycle_ok(S->quad, S->opt); }

  // The V-cycle on quads with red-black smoothing: first pass from zero (or,
  // with from_guess, from the values z already holds), residual fused with
  // the restriction, prolongation fused with the first post-smoothing pass,
  // <r, z> fused with the last one. Returns the buffer holding z (storage ST).
  template <int ST>
  TS<ST>* vcycle_quad(int dot_op, bool from_guess = false) {
    const QuadGrid& Q = S->qfine;
    const dim3 g((unsigned)(Q.nbx * Q.nby * Q.planes));
    Level L0 = S->levels.empty() ? Level{} : S->levels[0];
    TS<ST>* z = reinterpret_cast<TS<ST>*>(S->z);
    TS<ST>* zt = reinterpret_cast<TS<ST>*>(S->zt);
    const TS<ST>* none = nullptr;
    if (S->brick.bx > 0) return vcycle_fused<ST>(dot_op, from_guess);
    auto pass = [&](TS<ST>* x, int colour, int op) {
      FSP_LAUNCH_GEN(S, "k_q_gs<0>", (k_q_gs<0, ST, false>), (k_q_gs<0, ST, true>), g, BLOCK, st, G, Q,
                     (const TS<ST>*)x, x, S->r, colour, L0, none, op, S->ctl, S->partials);
      if (op >= 0) FSP_FINISH(S, st, op, g, false, 1);
    };
    if (from_guess) {
      pass(z, 0, -1);
      pass(z, 1, -1);
    } else {
      FSP_LAUNCH_GEN(S, "k_q_gs_first", (k_q_gs_first<ST, false>), (k_q_gs_first<ST, true>), g, BLOCK, st, G, Q, z,
                     S->r, S->ctl);
    }
    for (int s = 1; s < S->opt.pre_sweeps; ++s) {
      pass(z, 0, -1);
      pass(z, 1, -1);
    }
    TS<ST>* out = z;
    int passes = 2 * S->opt.post_sweeps;
    int done = 0;
    if (!S->levels.empty()) {
      // the restriction's grid for 128-thread halves (k_q_restrict2)
      QuadGrid H = S->qrestr;
      H.RY = (QBLOCK / 2) / H.QX;
      H.nby = (H.rows + H.RY - 1) / H.RY;
      const dim3 gr((unsigned)(H.nbx * H.nby * H.planes));
      FSP_LAUNCH_GEN(S, "k_q_restrict", (k_q_restrict2<ST, false>), (k_q_restrict2<ST, true>), gr, BLOCK, st, G, H,
                     (const TS<ST>*)z, (const float*)S->r, L0, S->ctl);
      const TS<ST>* x0 = coarse_cycle<ST>();
      const int op = passes == 1 ? dot_op : -1;
      FSP_LAUNCH_GEN(S, "k_q_gs<2>", (k_q_gs<2, ST, false>), (k_q_gs<2, ST, true>), g, BLOCK, st, G, Q,
                     (const TS<ST>*)z, zt, S->r, 1, L0, x0, op, S->ctl, S->partials);
      if (op >= 0) FSP_FINISH(S, st, op, g, false, 1);
      out = zt;
      done = 1;
    }
    for (int pz = done; pz < passes; ++pz) pass(out, (pz % 2 == 0) ? 1 : 0, pz == passes - 1 ? dot_op : -1);
    return out;
  }

  // The same V-cycle with each red-black sweep (both colours) in one
  // launch (k_q_brick): pre-smoothing red-black from the first pair of
  // passes on, post-smoothing black-red with the prolongation fused into
  // its first sweep. A sweep reads one buffer and writes the other, so the
  // iterate alternates between z and zt. Bit-identical to vcycle_quad's
  // single passes.
  template <int ST>
  TS<ST>* vcycle_fused(int dot_op, bool from_guess) {
    const QuadGrid& Q = S->qfine;
    const dim3 g((unsigned)(Q.nbx * Q.nby * Q.planes));
    const dim3 gb((unsigned)(S->brick.bx * S->brick.by * S->brick.bz));
    Level L0 = S->levels.empty() ? Level{} : S->levels[0];
    TS<ST>* a = reinterpret_cast<TS<ST>*>(S->z);
    TS<ST>* b = reinterpret_cast<TS<ST>*>(S->zt);
    const TS<ST>* none = nullptr;
    auto sweep = [&](int colour, const TS<ST>* xc, int op) {
      if (xc)
        FSP_LAUNCH_GEN(S, "k_q_brick<2>", (k_q_brick<2, ST, false>), (k_q_brick<2, ST, true>), gb, BR_THREADS, st, G,
                       S->brick, (const TS<ST>*)a, b, S->r, colour, L0, xc, op, S->ctl, S->partials);
      else
        FSP_LAUNCH_GEN(S, "k_q_brick<0>", (k_q_brick<0, ST, false>), (k_q_brick<0, ST, true>), gb, BR_THREADS, st, G,
                       S->brick, (const TS<ST>*)a, b, S->r, colour, L0, none, op, S->ctl, S->partials);
      if (op >= 0) FSP_FINISH(S, st, op, gb, false, 1);
      std::swap(a, b);
    };
    if (from_guess) sweep(0, none, -1);
    else
      FSP_LAUNCH_GEN(S, "k_q_gs_first", (k_q_gs_first<ST, false>), (k_q_gs_first<ST, true>), g, BLOCK, st, G, Q, a,
                     S->r, S->ctl);
    for (int s = 1; s < S->opt.pre_sweeps; ++s) sweep(0, none, -1);
    const TS<ST>* x0 = none;
    if (!S->levels.empty()) {
      QuadGrid H = S->qrestr;
      H.RY = (QBLOCK / 2) / H.QX;
      H.nby = (H.rows + H.RY - 1) / H.RY;
      const dim3 gr((unsigned)(H.nbx * H.nby * H.planes));
      FSP_LAUNCH_GEN(S, "k_q_restrict", (k_q_restrict2<ST, false>), (k_q_restrict2<ST, true>), gr, BLOCK, st, G, H,
                     (const TS<ST>*)a, (const float*)S->r, L0, S->ctl);
      x0 = coarse_cycle<ST>();
    }
    for (int s = 0; s < S->opt.post_sweeps; ++s)
      sweep(1, s == 0 ? x0 : none, s == S->opt.post_sweeps - 1 ? dot_op : -1);
    return a;
  }

  void precondition(int op) {
    if (S->net_phase) {
      // z = network(r) inside flexible CG; sigma from |r|^2 on the device
      const double* rr = op == OP_RZ ? &S->ctl->rn : &S->ctl->rr;
      (void)net_apply(S->net, st, S->r, S->mask, S->solid, rr, &S->ctl->wet_cells, &S->ctl->active, S->z);
      float* zout = S->z;
      if (S->opt.learned_vcycle) {
        // one V-cycle from the network's output: the network carries the
        // error the multigrid hierarchy leaves, the V-cycle the rest
        zout = reinterpret_cast<float*>(vcycle_quad<ST_F32>(-1, true));
      } else if (S->quad) {
        // red-black sweeps of A z = r from the network's output: the network
        // carries the smooth part of the error, the sweeps the rough part
        const QuadGrid& Q = S->qfine;
        const dim3 g((unsigned)(Q.nbx * Q.nby * Q.planes));
        const Level none_l{};
        const float* none = nullptr;
        for (int s = 0; s < S->opt.learned_smooth; ++s)
          for (int c = 0; c < 2; ++c)
            FSP_LAUNCH_GEN(S, "k_q_gs<0>", (k_q_gs<0, ST_F32, false>), (k_q_gs<0, ST_F32, true>), g, BLOCK, st,
                           G, Q, (const float*)S->z, S->z, S->r, c, none_l, none, -1, S->ctl, S->partials);
      }
      FSP_LAUNCH(S, k_dot_rz_qz, dim3(grid_for(G.n)), st, G.n, (const float*)S->r, (const float*)zout,
                 (const float*)S->q, op, S->ctl, S->partials);
      FSP_FINISH(S, st, op, dim3(grid_for(G.n)), false, 1);
      S->zres = zout;
      return;
    }
    if (S->opt.preconditioner == FSP_PRECOND_JACOBI) {
      FSP_LAUNCH(S, k_jacobi_prec, dim3(grid_for(G.n)), st, G, S->r, S->z, op, S->ctl,
                         S->partials);
      FSP_FINISH(S, st, op, dim3(grid_for(G.n)), false, 1);
      S->zres = S->z;
      return;
    }
    if (quad_vcycle()) {
      S->zres = with_storage(S->opt.vcycle_storage, [&](auto c) {
        return reinterpret_cast<float*>(vcycle_quad<decltype(c)::value>(op));
      });
      return;
    }
    vcycle();
    FSP_LAUNCH(S, k_dot, dim3(grid_for(G.n)), st, G.n, S->r, S->z, op, 1, S->ctl, S->partials);
    FSP_FINISH(S, st, op, dim3(grid_for(G.n)), false, 1);
    S->zres = S->z;
  }

  // The preconditioned vector in f32 (tests and the device entry point).
  void export_z(float* out) {
    with_storage(S->opt.vcycle_storage, [&](auto c) {
      constexpr int ST = decltype(c)::value;
      FSP_LAUNCH(S, k_unpack<ST>, dim3(grid_for(G.n)), st, G.n, (const TS<ST>*)S->zres, out);
      return 0;
    });
  }

  void iteration() {
    if (S->quad) {
      const QuadGrid& Q = S->qfine;
      const dim3 g((unsigned)(Q.nbx * Q.nby * Q.planes));
      const int storage = (S->opt.preconditioner != FSP_PRECOND_JACOBI && !S->net_phase) ? S->opt.vcycle_storage
                                                                                           : FSP_STORE_F32;
      with_storage(storage, [&](auto c) {
        constexpr int ST = decltype(c)::value;
        FSP_LAUNCH_GEN(S, "k_q_spmv_dir", (k_q_spmv_dir<ST, false>), (k_q_spmv_dir<ST, true>), g, BLOCK, st, G, Q,
                       (const TS<ST>*)S->zres, S->d0, S->d1, S->q, S->ctl, S->partials);
        return 0;
      });
      FSP_FINISH(S, st, OP_DQ, g, false, 1);
      const size_t nquads = G.n / 4;
      const dim3 gu((unsigned)((nquads + QBLOCK - 1) / QBLOCK));
      FSP_LAUNCH(S, k_q_update, gu, st, nquads, S->p64, S->r, (const float*)S->d0, (const float*)S->d1,
                 (const float*)S->q, S->ctl, S->partials);
      FSP_FINISH(S, st, OP_RN, gu, false, 1);
    } else {
      FSP_LAUNCH(S, k_spmv_dir, dim3(grid_for(G.n)), st, G, S->zres, S->d0, S->d1, S->q, S->ctl, S->partials);
      FSP_FINISH(S, st, OP_DQ, dim3(grid_for(G.n)), false, 1);
      FSP_LAUNCH(S, k_update, dim3(grid_for(G.n)), st, G.n, S->p64, S->r, S->d0, S->d1, S->q, S->ctl,
                 S->partials);
      FSP_FINISH(S, st, OP_RN, dim3(grid_for(G.n)), false, 1);
    }
    // true-residual replacement: every replace_every-th iteration r (and
    // |r|^2, which the loop test reads) becomes b - A p64 formed in f64; the
    // preconditioned residual, beta and the direction then follow from it
    if (S->opt.replace_every > 0) residual(S->p64, OP_RN_TRUE, true);
    precondition(OP_RZ);
  }

  void set_pass(int active, double rr, double tol_b2, double target, uint32_t it, uint32_t max_iters) {
    const uint32_t every = S->opt.replace_every > 0 ? (uint32_t)S->opt.replace_every : 0u;
    uint32_t limit = 0xffffffffu;
    if (S->net_phase) limit = std::min(limit, S->net_until);
    hipLaunchKernelGGL(k_set_pass, dim3(1), dim3(64), 0, st, S->ctl, active, rr, tol_b2, target, it, max_iters, limit,
                       S->net_phase, every);
  }
};

fsp_status sync_ctl(fsp_solver* S) {
  FSP_HIP(hipMemcpyAsync(S->hctl, S->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

void cleanup(fsp_solver* S) {
  for (int g = 0; g < 4; ++g) {
    if (S->iter_exec[g]) (void)hipGraphExecDestroy(S->iter_exec[g]);
    if (S->iter_graph[g]) (void)hipGraphDestroy(S->iter_graph[g]);
    S->iter_exec[g] = nullptr;
    S->iter_graph[g] = nullptr;
  }
  for (hipEvent_t e : S->prof_events) (void)hipEventDestroy(e);
  S->prof_events.clear();
  if (S->prof_active) (void)hipFree(S->prof_active);
  S->prof_active = nullptr;
  for (void* p : S->allocs) (void)hipFree(p);
  S->allocs.clear();
  if (S->hctl) (void)hipHostFree(S->hctl);
  if (S->hhash) (void)hipHostFree(S->hhash);
  if (S->hstep) (void)hipHostFree(S->hstep);
  if (S->hgout) (void)hipHostFree(S->hgout);
  if (S->ev0) (void)hipEventDestroy(S->ev0);
  if (S->ev1) (void)hipEventDestroy(S->ev1);
  if (S->stream) (void)hipStreamDestroy(S->stream);
}

// The correction and its metric, launched (read_measure reads the metric
// after the next sync of the control block). gauge: the mean-free pressure
// of a pure-Neumann problem first (the kernels return at once otherwise;
// a caller that has read the problem's flag passes it and launches none).
void launch_correct(fsp_solver* S, const float* us, const float* vs, const float* ws, double* p, float k1, float* u,
                    float* v, float* w, bool gauge) {
  Launch K(S);
  if (gauge) {
    FSP_LAUNCH(S, k_sum_active<double>, dim3(grid_for(K.G.n)), K.st, K.G, (const double*)p,
                       OP_MEAN_P, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_MEAN_P, dim3(grid_for(K.G.n)), false, 2);
    FSP_LAUNCH(S, k_sub_mean<double>, dim3(grid_for(K.G.n)), K.st, K.G, p, 1, -1, S->ctl,
                       S->partials);
  }
  if (S->quad && aligned16(vs) && aligned16(ws) && aligned16(v) && aligned16(w)) {
    const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
    FSP_LAUNCH(S, k_correct_q, gq, K.st, K.G, S->qfine, us, vs, ws, (const double*)p, k1, u, v, w, S->ctl,
               S->partials);
    FSP_FINISH(S, K.st, OP_UMAX, gq, true, 0);
  } else {
    FSP_LAUNCH(S, k_correct, dim3(grid_wide(K.G.n)), K.st, K.G, us, vs, ws,
               (const double*)p, k1, u, v, w, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_UMAX, dim3(grid_wide(K.G.n)), true, 0);
  }
  if (S->quad && aligned16(v) && aligned16(w)) {
    const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
    FSP_LAUNCH(S, k_divmetric_q, gq, K.st, K.G, S->qfine, u, v, w, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_DIVMAX, gq, true, 0);
  } else {
    FSP_LAUNCH(S, k_divmetric, dim3(grid_for(K.G.n)), K.st, K.G, u, v, w, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_DIVMAX, dim3(grid_for(K.G.n)), true, 0);
  }
}

void read_measure(fsp_solver* S, double& all, double& plain) {
  // delta3d_mobile.rs:422-432 (f32 arithmetic)
  const float dmax_all = (float)S->hctl->red[OP_DIVMAX][0];
  const float dmax_plain = (float)S->hctl->red[OP_DIVMAX][1];
  const float mx = (float)S->hctl->red[OP_UMAX][0];
  const float echelle = std::max(mx, 1e-4f);
  const float ratio = mx > 0.f ? S->dom.dx / echelle : 0.f;
  all = (double)(dmax_all * ratio);
  plain = (double)(dmax_plain * ratio);
}

// Correction and acceptance metric (delta3d_mobile.rs:560-561); returns
// (divergence over all wet rows, over plain rows), normalised as the reference.
fsp_status correct_and_measure(fsp_solver* S, const float* us, const float* vs, const float* ws, double* p, float k1,
                               float* u, float* v, float* w, double& all, double& plain, bool gauge) {
  launch_correct(S, us, vs, ws, p, k1, u, v, w, gauge);
  fsp_status st = sync_ctl(S);
  if (st) return st;
  read_measure(S, all, plain);
  return FSP_OK;
}

// The backward error of the true residual, launched (read from the control
// block after the next sync).
void launch_backward(fsp_solver* S, const double* p) {
  Launch K(S);
  if (S->quad) {
    const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
    FSP_LAUNCH_GEN(S, "k_backward_q", k_backward_q<false>, k_backward_q<true>, gq, QBLOCK, K.st, K.G, S->qfine,
                   (const double*)p, S->b, S->r, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_BACKWARD, gq, true, 0);
  } else {
    FSP_LAUNCH(S, k_backward, dim3(grid_for(K.G.n)), K.st, K.G, p, S->b, S->r, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_BACKWARD, dim3(grid_for(K.G.n)), true, 0);
  }
}

fsp_status backward_error(fsp_solver* S, const double* p, double& be) {
  launch_backward(S, p);
  fsp_status st = sync_ctl(S);
  if (st) return st;
  be = S->hctl->red[OP_BACKWARD][0];
  return FSP_OK;
}

fsp_status fingerprint(fsp_solver* S, const double* p, uint64_t& h) {
  // one partial per block; S->hashes holds (n + 1023) / 1024 of them
  const size_t chunks = std::min((size_t)grid_for(S->n), (S->n + 1023) / 1024);
  FSP_LAUNCH(S, k_hash_chunks, dim3((unsigned)chunks), S->stream, p, S->n, S->hashes);
  hipLaunchKernelGGL(k_hash_combine, dim3(1), dim3(HASH_THREADS), 0, S->stream, S->hashes, chunks, S->hash_out);
  FSP_HIP(hipMemcpyAsync(S->hhash, S->hash_out, sizeof(unsigned long long), hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  h = *S->hhash;
  return FSP_OK;
}

// Captures one CG iteration (every kernel reads its scalars from the device
// control block and every buffer is the solver's own, so the graph stays
// valid for the solver's lifetime) and instantiates it once.
fsp_status capture_iteration(fsp_solver* S) {
  const int slot = 2 * S->net_phase + S->table_only;
  if (S->iter_exec[slot]) return FSP_OK;
  const bool prof = S->prof_on;
  S->prof_on = false;
  FSP_HIP(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
  {
    Launch K(S);
    K.iteration();
  }
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(S->stream, &g);
  S->prof_on = prof;
  if (e != hipSuccess || !g) {
    std::fprintf(stderr, "fsp: graph capture failed: %s\n", hipGetErrorString(e));
    return FSP_DEVICE;
  }
  S->iter_graph[slot] = g;
  FSP_HIP(hipGraphInstantiate(&S->iter_exec[slot], g, nullptr, nullptr, 0));
  // device resources of the graph now, not at its first replay
  FSP_HIP(hipGraphUpload(S->iter_exec[slot], S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  if (S->net_phase == 1) S->net_gen = net_generation(S->net);
  return FSP_OK;
}

// Drops both kernel forms' graphs of phase `phase`.
void drop_graph(fsp_solver* S, int phase) {
  for (int t = 0; t < 2; ++t) {
    const int slot = 2 * phase + t;
    if (S->iter_exec[slot]) (void)hipGraphExecDestroy(S->iter_exec[slot]);
    if (S->iter_graph[slot]) (void)hipGraphDestroy(S->iter_graph[slot]);
    S->iter_exec[slot] = nullptr;
    S->iter_graph[slot] = nullptr;
  }
}

// Captures and instantiates the iteration graph of phase `phase` (0: as
// configured, 1: the learned phase) ahead of the projections, so that a
// projection allocates nothing (the reference reserves every buffer before
// seal(), delta3d.rs:225-241). A throwaway capture of the prime first sets
// which buffer holds M^-1 r, as the prime of a projection does before its
// first iteration. Both kernel forms (general and table-only) are captured;
// each projection replays the one its geometry allows.
fsp_status prepare_graph(fsp_solver* S, int phase) {
  if (!S->opt.graph || (S->iter_exec[2 * phase] && S->iter_exec[2 * phase + 1])) return FSP_OK;
  if (phase == 1 && !S->net) return FSP_OK;
  const int saved = S->net_phase, saved_table = S->table_only;
  const bool prof = S->prof_on;
  S->net_phase = phase;
  S->prof_on = false;
  hipError_t e = hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    {
      Launch K(S);
      K.precondition(OP_RZ_PRIME);
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(S->stream, &g);
    if (g) (void)hipGraphDestroy(g);
  }
  S->prof_on = prof;
  fsp_status st = e == hipSuccess ? FSP_OK : FSP_DEVICE;
  for (int t = 0; t < 2 && st == FSP_OK; ++t) {
    S->table_only = t;
    st = capture_iteration(S);
  }
  S->table_only = saved_table;
  S->net_phase = saved;
  if (e != hipSuccess) std::fprintf(stderr, "fsp: prime capture failed: %s\n", hipGetErrorString(e));
  return st;
}

// The iteration graphs hold the fine geometry's kernel arguments: after a
// change of geometry mode, capture them again.
fsp_status regraph(fsp_solver* S) {
  for (int gi = 0; gi < 2; ++gi) drop_graph(S, gi);
  fsp_status gst = prepare_graph(S, 0);
  if (gst == FSP_OK && S->net && S->opt.preconditioner == FSP_PRECOND_LEARNED) gst = prepare_graph(S, 1);
  return gst;
}

// Geometry rebuilt from the surface at the start of every projection
// (delta3d_mobile.rs:479-484).
void build_geometry(fsp_solver* S, bool hierarchy, bool prec = true) {
  Launch K(S);
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, K.st, S->ctl, S->opt.gauge);
  K.mask(prec);
  if (hierarchy && S->opt.preconditioner != FSP_PRECOND_JACOBI) K.hierarchy();
}

// project_mobile3_graded (delta3d_graded.rs:310-447) on device buffers:
// conjugate gradients on P^T A P x = P^T b preconditioned by the condensed
// diagonal, f64 dot products folded on the device and read by the host each
// iteration, the true residual and the restricted divergence metric deciding
// each pass, the fine pressure P x corrected into u, v, w.
fsp_status run_graded(fsp_solver* S, const float* us, const float* vs, const float* ws, float* p, int warm, float* u,
                      float* v, float* w, float scale, float k1, double rel_tol, uint32_t max_iters,
                      fsp_report* rep) {
  // a cut or level-set geometry is refused (delta3d_graded.rs:70-72)
  if (S->cut || S->levelset) return FSP_DOMAIN;
  const GradedCol D = S->graded_col();
  const size_t cols = S->ncols, n = cols * (size_t)D.nn, N = S->n;
  const int nz = S->dom.nz;
  int j0 = D.nn - 1;
  while (j0 > 0 && S->gnodes[j0 - 1] + 1 == S->gnodes[j0]) --j0;
  const int k0 = S->gnodes[j0];
  Launch K(S);
  const Fine& G = K.G;
  hipStream_t st = K.st;
  FSP_HIP(hipEventRecord(S->ev0, st));
  build_geometry(S, false);
  // the guard (delta3d_graded.rs:316-322): the lowest cubic node's layer and
  // the one above it wet in every column
  if (k0 + 1 >= nz) return FSP_DOMAIN;
  {
    std::vector<uint8_t> rows(2 * cols);
    FSP_HIP(hipMemcpyAsync(rows.data(), S->mask + (size_t)k0 * cols, 2 * cols, hipMemcpyDeviceToHost, st));
    FSP_HIP(hipStreamSynchronize(st));
    for (uint8_t m : rows)
      if (!(m & M_WET)) return FSP_DOMAIN;
  }
  // right-hand side and prec on the fine grid (rhs_mobile3), restricted; the
  // condensed diagonal from prec
  if (S->quad && aligned16(vs) && aligned16(ws)) {
    const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
    FSP_LAUNCH(S, k_assemble_q, gq, st, G, S->qfine, us, vs, ws, (const float*)nullptr, scale, S->b, S->d0, S->ctl,
               S->partials);
    FSP_FINISH(S, st, OP_B2, gq, false, 0);
  } else {
    FSP_LAUNCH(S, k_assemble, dim3(grid_for(N)), st, G, us, vs, ws, (const float*)nullptr, scale, S->b, S->d0,
               S->ctl, S->partials);
    FSP_FINISH(S, st, OP_B2, dim3(grid_for(N)), false, 0);
  }
  fsp_status sts = sync_ctl(S);
  if (sts) return sts;
  if (S->hctl->nonfinite) return FSP_NOT_FINITE;
  const dim3 gn(grid_for(n)), gN(grid_for(N));
  FSP_LAUNCH(S, k_gr_restrict<false>, gn, st, D, (const float*)S->b, S->gb);
  FSP_LAUNCH(S, k_gr_restrict<true>, gn, st, D, (const float*)S->d0, S->gm);
  // warm start by injection; a dry node starts from 0 (:378-383)
  if (warm) FSP_LAUNCH(S, k_gr_inject, gn, st, D, (const float*)p, (const float*)S->gm, S->gx);
  else FSP_LAUNCH(S, k_fill, gn, st, S->gx, n, 0.f);
  auto dots = [&](const float* a, const float* b, const float* c, const float* e, double& s1,
                  double& s2) -> fsp_status {
    hipLaunchKernelGGL(k_gr_dot, gn, dim3(BLOCK), 0, st, n, a, b, c, e, S->gpart);
    hipLaunchKernelGGL(k_gr_fold<false>, dim3(1), dim3(BLOCK), 0, st, (int)gn.x, (const double*)S->gpart, S->gout);
    FSP_HIP(hipMemcpyAsync(S->hgout, S->gout, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    FSP_HIP(hipStreamSynchronize(st));
    s1 = S->hgout[0];
    s2 = S->hgout[1];
    return FSP_OK;
  };
  // out = P^T A P x through the fine grid (:147-162)
  auto gapply = [&](const float* x, float* out) {
    FSP_LAUNCH(S, k_gr_expand, gN, st, D, x, S->d1);
    FSP_LAUNCH(S, k_apply, gN, st, G, (const float*)S->d1, S->q);
    FSP_LAUNCH(S, k_gr_restrict<false>, gn, st, D, (const float*)S->q, out);
  };
  double b2, unused;
  if ((sts = dots(S->gb, S->gb, nullptr, nullptr, b2, unused))) return sts;
  if (!std::isfinite(b2)) return FSP_NOT_FINITE;
  const double tol = rel_tol * rel_tol;
  double target = tol * b2, rr = 0.0, rz = 0.0;
  uint32_t it = 0;
  bool floor_stop = false, have_ck = false;
  uint64_t ck = 0;
  uint32_t power = 1, since = 0;
  double d_all = 0.0, d_plain = 0.0;
  std::vector<float> hx;
  for (;;) {
    // the true residual, then PCG to the target
    gapply(S->gx, S->gq);
    FSP_LAUNCH(S, k_gr_restart, gn, st, n, (const float*)S->gb, (const float*)S->gq, (const float*)S->gm, S->gr,
               S->gz, S->gd);
    if ((sts = dots(S->gr, S->gr, S->gr, S->gz, rr, rz))) return sts;
    while (b2 > 0.0 && rr > target && it < max_iters) {
      gapply(S->gd, S->gq);
      double dq;
      if ((sts = dots(S->gd, S->gq, nullptr, nullptr, dq, unused))) return sts;
      if (!(dq > 0.0)) break;
      const float alpha = (float)(rz / dq);
      FSP_LAUNCH(S, k_gr_update, gn, st, n, alpha, (const float*)S->gd, (const float*)S->gq, (const float*)S->gm,
                 S->gx, S->gr, S->gz);
      double zn, rn;
      if ((sts = dots(S->gr, S->gz, S->gr, S->gr, zn, rn))) return sts;
      const float beta = (float)(zn / rz);
      FSP_LAUNCH(S, k_gr_direction, gn, st, n, beta, (const float*)S->gz, S->gd);
      rz = zn;
      rr = rn;
      ++it;
    }
    gapply(S->gx, S->gq);
    FSP_LAUNCH(S, k_gr_restart, gn, st, n, (const float*)S->gb, (const float*)S->gq, (const float*)S->gm, S->gr,
               (float*)nullptr, (float*)nullptr);
    FSP_LAUNCH(S, k_gr_expand, gN, st, D, (const float*)S->gx, p);
    if ((sts = dots(S->gr, S->gr, nullptr, nullptr, rr, unused))) return sts;
    if (!std::isfinite(rr)) return FSP_NOT_FINITE;
    // correction from the fine pressure, then the restricted metric (:451-486)
    FSP_LAUNCH(S, k_gr_widen, gN, st, N, (const float*)p, S->p64);
    launch_correct(S, us, vs, ws, S->p64, k1, u, v, w, S->opt.gauge != 0);
    FSP_LAUNCH(S, k_divergence, gN, st, G, (const float*)u, (const float*)v, (const float*)w, S->d1);
    FSP_LAUNCH(S, k_gr_restrict<false>, gn, st, D, (const float*)S->d1, S->gz);
    hipLaunchKernelGGL(k_gr_metric, gn, dim3(BLOCK), 0, st, D, (const float*)S->gz, (const float*)S->g_weight,
                       (const uint8_t*)S->mask, j0, S->gpart);
    hipLaunchKernelGGL(k_gr_fold<true>, dim3(1), dim3(BLOCK), 0, st, (int)gn.x, (const double*)S->gpart, S->gout);
    const dim3 gf(grid_for(S->nu + S->nv + S->nw));
    hipLaunchKernelGGL(k_gr_absmax, gf, dim3(BLOCK), 0, st, (const float*)u, S->nu, (const float*)v, S->nv,
                       (const float*)w, S->nw, S->gpart);
    hipLaunchKernelGGL(k_gr_fold<true>, dim3(1), dim3(BLOCK), 0, st, (int)gf.x, (const double*)S->gpart,
                       S->gout + 2);
    FSP_HIP(hipMemcpyAsync(S->hgout, S->gout, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    FSP_HIP(hipStreamSynchronize(st));
    {
      const float mx = (float)S->hgout[2];
      const float ratio = mx > 0.f ? S->dom.dx / std::max(mx, 1e-4f) : 0.f;
      d_all = (double)((float)S->hgout[0] * ratio);
      d_plain = (double)((float)S->hgout[1] * ratio);
    }
    const bool converged = rr <= tol * b2;
    if (b2 == 0.0 || (converged && d_plain <= 1e-5) || it >= max_iters) break;
    // a state that comes back bit for bit is the floor (:430-438)
    hx.resize(n);
    FSP_HIP(hipMemcpy(hx.data(), S->gx, n * sizeof(float), hipMemcpyDeviceToHost));
    uint64_t state = 0xcbf29ce484222325ull;
    for (float xv : hx) {
      uint32_t bits;
      std::memcpy(&bits, &xv, 4);
      for (int b = 0; b < 4; ++b) {
        state ^= (bits >> (8 * b)) & 0xFFu;
        state *= 0x100000001b3ull;
      }
    }
    if (have_ck && ck == state) {
      floor_stop = true;
      break;
    }
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
  FSP_HIP(hipEventRecord(S->ev1, st));
  FSP_HIP(hipEventSynchronize(S->ev1));
  float ms = 0.f;
  FSP_HIP(hipEventElapsedTime(&ms, S->ev0, S->ev1));
  fsp_report R{};
  R.refinements = 0;
  R.iterations = it;
  R.degraded = accepted ? 0 : 1;
  R.residual = b2 > 0.0 ? std::sqrt(rr / b2) : 0.0;
  R.divergence = d_all;
  R.floor = floor_stop ? 1 : 0;
  R.backward_error = 0.0;
  R.divergence_plain = d_plain;
  R.ms_device = ms;
  if (rep) *rep = R;
  if (R.degraded && S->opt.refuse_degraded) return FSP_CONVERGENCE;
  return FSP_OK;
}

// The projection on device buffers (project_mobile3, delta3d_mobile.rs:465-609).
fsp_status run(fsp_solver* S, const float* us, const float* vs, const float* ws, float* p, int warm, float* u,
               float* v, float* w, float scale, float k1, double rel_tol, uint32_t max_iters, fsp_report* rep) {
  if (!S->geom) return FSP_DOMAIN;
  // the graded column takes the projection (delta3d_mobile.rs:473-475)
  if (!S->gnodes.empty()) return run_graded(S, us, vs, ws, p, warm, u, v, w, scale, k1, rel_tol, max_iters, rep);
  // the table-only kernel forms hold for this projection only
  struct TableGuard {
    fsp_solver* S;
    ~TableGuard() { S->table_only = 0; }
  } table_guard{S};
  S->table_only = 0;
  S->net_phase = 0;
  if (S->opt.preconditioner == FSP_PRECOND_LEARNED) {
    if (!S->net) return FSP_INVALID;
    const fsp_status ns = net_reserve(S->net, S->dom.nx, S->dom.ny, S->dom.nz);
    if (ns) return ns;
    // the network's buffers moved since the learned graph was captured (a
    // forward pass or another solver at another shape): capture it again
    if (S->iter_exec[2] && S->net_gen != net_generation(S->net)) {
      drop_graph(S, 1);
      if (const fsp_status gs = prepare_graph(S, 1)) return gs;
    }
    S->net_phase = 1;
    S->net_until = S->opt.learned_budget > 0 ? (uint32_t)S->opt.learned_budget : 50u;
  }
  Launch K(S);
  const Fine& G = K.G;
  FSP_HIP(hipEventRecord(S->ev0, S->stream));
  // On a column geometry with binary apertures the assembly writes the
  // per-row factors the first red-black pass reads (the diagonal it forms
  // anyway, the arithmetic of k_prec's row_prec), in place of k_prec's pass.
  float* const prec_out = (S->precv && !S->frac && !S->levelset) ? S->precv : nullptr;
  build_geometry(S, true, prec_out == nullptr);
  if (S->quad && aligned16(vs) && aligned16(ws)) {
    const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
    FSP_LAUNCH(S, k_assemble_q, gq, K.st, G, S->qfine, us, vs, ws, (const float*)nullptr, scale, S->b, prec_out,
               S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_B2, gq, false, 0);
  } else {
    FSP_LAUNCH(S, k_assemble, dim3(grid_for(G.n)), K.st, G, us, vs, ws, (const float*)nullptr,
               scale, S->b, prec_out, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_B2, dim3(grid_for(G.n)), false, 0);
  }
  if (S->opt.gauge) {
    FSP_LAUNCH(S, k_sum_active<float>, dim3(grid_for(G.n)), K.st, G, (const float*)S->b,
                       OP_MEAN_B, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_MEAN_B, dim3(grid_for(G.n)), false, 2);
    FSP_LAUNCH(S, k_sub_mean<float>, dim3(grid_for(G.n)), K.st, G, S->b, 0, OP_B2, S->ctl,
                       S->partials);
    FSP_FINISH(S, K.st, OP_B2, dim3(grid_for(G.n)), false, 2);
  }
  double* p64 = S->p64;
  if (!warm && S->quad) {
    const size_t nquads = G.n / 4;
    const dim3 gs((unsigned)((nquads + QBLOCK - 1) / QBLOCK));
    FSP_LAUNCH(S, k_start_cold, gs, K.st, nquads, (const float*)S->b, p64, S->r, OP_RR, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_RR, gs, false, 0);
  } else {
    FSP_LAUNCH(S, k_load_p, dim3(grid_for(G.n)), K.st, G, (const float*)p, p64, warm);
    K.residual(p64, OP_RR);
  }
  fsp_status st = sync_ctl(S);
  if (st) return st;
  if (S->hctl->nonfinite) return FSP_NOT_FINITE;
  // every ghost row in its column's table (binary apertures, column
  // geometry): the fine kernels' table-only forms
  S->table_only = (!S->frac && !S->levelset && !S->hctl->table_miss) ? 1 : 0;
  // the correction's gauge runs only for a pure-Neumann problem (known now)
  const bool gauge_now = S->opt.gauge && S->hctl->pure_neumann;

  const double b2 = S->hctl->b2;
  const double tol_b2 = rel_tol * rel_tol * b2;
  double rr = S->hctl->actual;
  double physical_target = std::numeric_limits<double>::infinity();
  uint32_t it = 0;
  bool primed = false, floor_stop = false;
  bool have_ck = false;
  uint64_t ck = 0;
  uint32_t power = 1, since = 0;
  double actual_rr = 0.0;
  bool settled = false;
  bool wrapped = false;  // the loop ended on a pass whose correction and backward error were launched with it
  double set_all = 0.0, set_plain = 0.0;
  // The floor certificate gamma_{m+4} (delta3d.rs:65-72, delta_projection.rs:194-203)
  // with the unit roundoff of the pressure's precision (f64 here).
  const double unit = 1.0 / 9007199254740992.0;
  const double gm = S->dom.ny == 1 ? 8.0 : 10.0;
  const double roundoff = gm * unit / (1.0 - gm * unit);
  auto loop_test = [&](double rrv, uint32_t itv) {
    return b2 > 0.0 && (rrv > tol_b2 || rrv > physical_target) && itv < max_iters;
  };
  for (;;) {
    const uint32_t before = it;
    bool net_failed = false;
    if (loop_test(rr, it)) {
      K.set_pass(1, rr, tol_b2, physical_target, it, max_iters);
      if (!primed) {
        K.precondition(OP_RZ_PRIME);
        primed = true;
      }
      // Batches of iterations between polls, sized from the observed rate.
      int batch = 1;
      double rr_prev = rr;
      uint32_t it_prev = it;
      for (;;) {
        if (S->opt.graph && !S->prof_on && (st = capture_iteration(S))) return st;
        for (int s = 0; s < batch; ++s) {
          if (hipGraphExec_t gx = S->iter_exec[2 * S->net_phase + S->table_only]; gx && !S->prof_on)
            FSP_HIP(hipGraphLaunch(gx, S->stream));
          else K.iteration();
        }
        if ((st = sync_ctl(S))) return st;
        if (S->hctl->nonfinite) {
          if (!S->net_phase) return FSP_NOT_FINITE;
          net_failed = true;  // the network's output: fall back to MGPCG from here
          break;
        }
        if (!S->hctl->active) break;
        const double rr_now = S->hctl->rr;
        const uint32_t it_now = S->hctl->it;
        if (it_now <= it_prev) {
          std::fprintf(stderr, "fsp: the iteration count did not advance (%u)\n", it_now);
          return FSP_DEVICE;
        }
        int next = 2;
        if (it_now > it_prev && rr_now > 0.0 && rr_prev > 0.0 && rr_now < rr_prev) {
          const double rate = std::pow(rr_now / rr_prev, 1.0 / (double)(it_now - it_prev));
          const double target = std::min(tol_b2, physical_target);
          if (rate < 1.0 && target > 0.0) {
            const double est = std::ceil(std::log(target / rr_now) / std::log(rate));
            next = (int)std::max(1.0, std::min(16.0, est));
          }
        }
        next = (int)std::min<uint64_t>((uint64_t)next, (uint64_t)(max_iters - it_now));
        batch = std::max(1, next);
        rr_prev = rr_now;
        it_prev = it_now;
      }
      it = S->hctl->it;
      rr = S->hctl->rr;
      if (net_failed) K.set_pass(0, rr, tol_b2, physical_target, it, max_iters);  // clears the flag
    }
    // True residual, recomputed: it decides, never the recurrence (:546-553).
    // The correction with its metric and the backward error ride on the same
    // sync as the true residual (read below if the pass ends in them; the
    // correction writes only u, v, w, which a later one overwrites). Not with
    // the gauge of a pure-Neumann problem, whose correction shifts p first.
    // On quads the backward error comes out of the residual's own pass.
    const bool spec = !(S->opt.gauge && S->hctl->pure_neumann);
    if (spec && S->quad) {
      K.residual_backward(p64);
      launch_correct(S, us, vs, ws, p64, k1, u, v, w, gauge_now);
    } else {
      K.residual(p64, OP_RR);
      if (spec) {
        launch_correct(S, us, vs, ws, p64, k1, u, v, w, gauge_now);
        launch_backward(S, p64);
      }
    }
    if ((st = sync_ctl(S))) return st;
    if (S->hctl->nonfinite) return FSP_NOT_FINITE;
    const double actual = S->hctl->actual;
    // The learned phase ends at its budget, on a breakdown (<d, A d> <= 0 or a
    // non-finite network output) or when a pass made no progress; MGPCG
    // then continues from the current pressure.
    bool switched = false;
    if (S->net_phase && (net_failed || S->hctl->stop_dq || it >= S->net_until || it == before)) {
      S->net_phase = 0;
      switched = true;
    }
    const bool exhausted = it >= max_iters || (it == before && !switched);
    bool have_measured = false;
    double m_all = 0.0, m_plain = 0.0;
    if (actual <= tol_b2) {
      if (exhausted) { actual_rr = actual; wrapped = spec; break; }
      double a, pl;
      if (spec) read_measure(S, a, pl);
      else if ((st = correct_and_measure(S, us, vs, ws, p64, k1, u, v, w, a, pl, gauge_now))) return st;
      if (pl <= 1e-5) { actual_rr = actual; settled = true; set_all = a; set_plain = pl; wrapped = spec; break; }
      have_measured = true; m_all = a; m_plain = pl;
      const float ratio = (float)(1e-5 / pl);
      physical_target = (double)((float)actual * ratio * ratio);
    } else if (exhausted) {
      actual_rr = actual;
      wrapped = spec;
      break;
    }
    double be;
    if (spec) be = S->hctl->red[OP_BACKWARD][0];
    else if ((st = backward_error(S, p64, be))) return st;
    if (be <= roundoff) {
      floor_stop = true;
      actual_rr = actual;
      if (have_measured) { settled = true; set_all = m_all; set_plain = m_plain; }
      wrapped = spec;
      break;
    }
    uint64_t state;
    if ((st = fingerprint(S, p64, state))) return st;
    if (have_ck && ck == state) {
      floor_stop = true;
      actual_rr = actual;
      if (have_measured) { settled = true; set_all = m_all; set_plain = m_plain; }
      break;
    }
    ++since;
    if (since == power) {
      have_ck = true;
      ck = state;
      power = power > 0x7fffffffu ? 0xffffffffu : power * 2;
      since = 0;
    }
    rr = actual;
    // Re-prime on the true residual (:587) when another pass will run.
    if (loop_test(rr, it)) {
      K.set_pass(1, rr, tol_b2, physical_target, it, max_iters);
      K.precondition(OP_RZ_PRIME);
    }
  }
  double d_all, d_plain;
  if (settled) { d_all = set_all; d_plain = set_plain; }
  else if (wrapped) read_measure(S, d_all, d_plain);
  else if ((st = correct_and_measure(S, us, vs, ws, p64, k1, u, v, w, d_all, d_plain, gauge_now))) return st;
  double be;
  if (wrapped) be = S->hctl->red[OP_BACKWARD][0];
  else if ((st = backward_error(S, p64, be))) return st;
  FSP_LAUNCH(S, k_store_p, dim3(grid_for(G.n)), K.st, G.n, (const double*)p64, p);
  const bool accepted = (actual_rr <= tol_b2 || floor_stop) && d_plain <= 1e-5;
  FSP_HIP(hipEventRecord(S->ev1, S->stream));
  FSP_HIP(hipEventSynchronize(S->ev1));
  float ms = 0.f;
  FSP_HIP(hipEventElapsedTime(&ms, S->ev0, S->ev1));
  fsp_report R{};
  R.refinements = 0;
  R.iterations = it;
  R.degraded = (b2 > 0.0 && !accepted) ? 1 : 0;
  R.residual = b2 > 0.0 ? (double)std::sqrt((float)(actual_rr / b2)) : 0.0;
  R.divergence = d_all;
  R.floor = floor_stop ? 1 : 0;
  R.backward_error = be;
  R.divergence_plain = d_plain;
  R.ms_device = ms;
  R.ms_transfer = 0.0;
  if (rep) *rep = R;
  if (R.degraded && S->opt.refuse_degraded) return FSP_CONVERGENCE;
  return FSP_OK;
}

fsp_status step_reserve(fsp_solver* S);

bool all_finite(const float* x, size_t n) {
  for (size_t q = 0; q < n; ++q) if (!std::isfinite(x[q])) return false;
  return true;
}

}  // namespace

extern "C" {

int32_t fsp_abi_version(void) { return FSP_ABI_VERSION; }

const char* fsp_status_string(fsp_status s) {
  switch (s) {
    case FSP_OK: return "ok";
    case FSP_DOMAIN: return "domain";
    case FSP_NOT_FINITE: return "not finite";
    case FSP_SHAPE: return "shape";
    case FSP_CONVERGENCE: return "convergence";
    case FSP_DEVICE: return "device";
    case FSP_INVALID: return "invalid";
  }
  return "unknown";
}

void fsp_default_options(fsp_options* o) {
  if (!o) return;
  o->preconditioner = FSP_PRECOND_MULTIGRID;
  o->smoother = FSP_SMOOTH_JACOBI;
  o->max_levels = -1;
  o->pre_sweeps = 2;
  o->post_sweeps = 2;
  o->coarse_sweeps = 8;
  o->gauge = 1;
  o->vcycle_storage = FSP_STORE_F32;
  o->replace_every = 0;
  o->refuse_degraded = 0;
  o->device = 0;
  o->coarse_operator = FSP_COARSE_REDISCRETISED;
  o->graph = 1;
  o->learned_budget = 0;
  o->learned_smooth = 1;
  o->learned_vcycle = 0;
  o->fine_sweeps = FSP_SWEEP_FUSED;
}

fsp_status fsp_create(const fsp_domain* d, const fsp_options* o, fsp_solver** out) {
  if (!d || !out) return FSP_INVALID;
  *out = nullptr;
  if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || !std::isfinite(d->dx) || !(d->dx > 0.f)) return FSP_DOMAIN;
  const uint64_t cells = (uint64_t)d->nx * (uint64_t)d->ny * (uint64_t)d->nz;
  if (cells >= (1ull << 31)) return FSP_DOMAIN;
  fsp_options opt;
  fsp_default_options(&opt);
  if (o) opt = *o;
  if (opt.preconditioner < 0 || opt.preconditioner > 2 || opt.smoother < 0 || opt.smoother > 1) return FSP_INVALID;
  if (opt.learned_budget < 0 || opt.learned_smooth < 0 || opt.learned_smooth > 64) return FSP_INVALID;
  if (opt.learned_vcycle < 0 || opt.learned_vcycle > 1) return FSP_INVALID;
  if (opt.fine_sweeps < 0 || opt.fine_sweeps > 1) return FSP_INVALID;
  if (opt.preconditioner == FSP_PRECOND_LEARNED && opt.learned_vcycle &&
      (opt.vcycle_storage != FSP_STORE_F32 || opt.smoother != FSP_SMOOTH_RBGS || opt.pre_sweeps < 1 ||
       opt.post_sweeps < 1))
    return FSP_INVALID;
  if (opt.pre_sweeps < 0 || opt.post_sweeps < 0 || opt.coarse_sweeps < 0 || opt.pre_sweeps > 64 ||
      opt.post_sweeps > 64 || opt.coarse_sweeps > 1024)
    return FSP_INVALID;
  if (opt.vcycle_storage < FSP_STORE_F32 || opt.vcycle_storage > FSP_STORE_BF16) return FSP_INVALID;
  if (opt.coarse_operator < 0 || opt.coarse_operator > 1) return FSP_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || opt.device < 0 || opt.device >= ndev) return FSP_DEVICE;
  fsp_solver* S = new (std::nothrow) fsp_solver();
  if (!S) return FSP_DEVICE;
  S->dom = *d;
  S->opt = opt;
  fsp_status st = FSP_OK;
  auto fail = [&](fsp_status s) { cleanup(S); delete S; return s; };
  if (hipSetDevice(opt.device) != hipSuccess) return fail(FSP_DEVICE);
  if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess) return fail(FSP_DEVICE);
  if (hipEventCreate(&S->ev0) != hipSuccess || hipEventCreate(&S->ev1) != hipSuccess) return fail(FSP_DEVICE);
  const size_t nx = d->nx, ny = d->ny, nz = d->nz;
  S->n = nx * ny * nz;
  S->nu = (nx + 1) * ny * nz;
  S->nv = nx * (ny + 1) * nz;
  S->nw = nx * ny * (nz + 1);
  S->ncols = nx * ny;
  if ((st = dalloc(S, &S->solid, S->n)) || (st = dalloc(S, &S->ou, S->nu)) || (st = dalloc(S, &S->ov, S->nv)) ||
      (st = dalloc(S, &S->ow, S->nw)) || (st = dalloc(S, &S->mask, S->n)) || (st = dalloc(S, &S->eta, S->ncols)) ||
      (st = dalloc(S, &S->eta_rt, S->ncols)) || (st = dalloc(S, &S->gk0, S->ncols)) ||
      (st = dalloc(S, &S->col_floor, S->ncols)) ||
      (st = dalloc(S, &S->gtab, S->ncols * (size_t)G_SLOTS * 8)))
    return fail(st);
  float** cellbufs[] = {&S->b, &S->r, &S->p, &S->q, &S->z, &S->zt, &S->d0, &S->d1};
  for (float** bp : cellbufs) if ((st = dalloc(S, bp, S->n))) return fail(st);
  if ((st = dalloc(S, &S->p64, S->n))) return fail(st);
  if ((st = dalloc(S, &S->us, S->nu)) || (st = dalloc(S, &S->vs, S->nv)) || (st = dalloc(S, &S->ws, S->nw)) ||
      (st = dalloc(S, &S->u, S->nu)) || (st = dalloc(S, &S->v, S->nv)) || (st = dalloc(S, &S->w, S->nw)))
    return fail(st);
  if ((st = dalloc(S, &S->ctl, 1)) || (st = dalloc(S, &S->partials, (size_t)NUM_SLOTS * MAX_PARTIALS * 2)) ||
      (st = dalloc(S, &S->hashes, (S->n + 1023) / 1024)) || (st = dalloc(S, &S->hash_out, 1)))
    return fail(st);
  if (hipMemset(S->ctl, 0, sizeof(Ctl)) != hipSuccess) return fail(FSP_DEVICE);
  if (hipHostMalloc((void**)&S->hctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess) return fail(FSP_DEVICE);
  if (hipHostMalloc((void**)&S->hhash, sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
    return fail(FSP_DEVICE);
  if (opt.preconditioner != FSP_PRECOND_JACOBI) {  // the learned preconditioner falls back to the V-cycle
    size_t x = nx, y = ny, z = nz;
    float h = d->dx;
    while (coarsens3(x, y, z) && (opt.max_levels < 0 || (int)S->levels.size() < opt.max_levels)) {
      x /= 2; y /= 2; z /= 2;
      h *= 2.f;
      Level L{};
      L.nx = (int)x; L.ny = (int)y; L.nz = (int)z;
      L.n = x * y * z;
      L.inv = 1.0f / (h * h);
      if ((st = dalloc(S, &L.kind, L.n)) || (st = dalloc(S, &L.ou, (x + 1) * y * z)) ||
          (st = dalloc(S, &L.ov, x * (y + 1) * z)) || (st = dalloc(S, &L.ow, x * y * (z + 1))) ||
          (st = dalloc(S, &L.diag, L.n)) || (st = dalloc(S, &L.x, L.n)) || (st = dalloc(S, &L.r, L.n)) ||
          (st = dalloc(S, &L.t, L.n)) || (st = dalloc(S, &L.dd, L.n)))
        return fail(st);
      L.galerkin = opt.coarse_operator == FSP_COARSE_GALERKIN ? 1 : 0;
      S->levels.push_back(L);
    }
  }
  // fast paths
  auto make_q = [](int nq, int rows, int planes) {
    QuadGrid Q{};
    int qx = 1;
    while (qx < nq && qx < FSP_QX_MAX) qx <<= 1;
    Q.nq = nq; Q.QX = qx; Q.RY = QBLOCK / qx;
    Q.nbx = (nq + qx - 1) / qx;
    Q.nby = (rows + Q.RY - 1) / Q.RY;
    Q.rows = rows; Q.planes = planes;
    return Q;
  };
  if (nx % 4 == 0) {
    S->qfine = make_q((int)nx / 4, (int)ny, (int)nz);
    S->qrestr = make_q((int)nx / 4, (int)ny / 2, (int)nz / 2);
    const size_t blocks = (size_t)S->qfine.nbx * S->qfine.nby * S->qfine.planes;
    S->quad = blocks <= (size_t)MAX_PARTIALS && (S->n / 4 + QBLOCK - 1) / QBLOCK <= (size_t)MAX_PARTIALS;
    if (S->quad && opt.fine_sweeps == FSP_SWEEP_FUSED) {
      S->brick.bx = ((int)nx / 4 + BR_QX - 1) / BR_QX;
      S->brick.by = ((int)ny + BR_Y - 1) / BR_Y;
      S->brick.bz = ((int)nz + BR_Z - 1) / BR_Z;
      if ((size_t)S->brick.bx * S->brick.by * S->brick.bz > (size_t)MAX_PARTIALS) S->brick = BrickGrid{};
    }
  }
  if (opt.preconditioner == FSP_PRECOND_LEARNED && opt.learned_vcycle && !quad_vcycle_ok(S->quad, opt))
    return fail(FSP_INVALID);
  if (opt.preconditioner != FSP_PRECOND_JACOBI && quad_vcycle_ok(S->quad, opt) && (st = dalloc(S, &S->precv, S->n)))
    return fail(st);
  S->tail_first = (int)S->levels.size();
  for (int t = (int)S->levels.size() - 1; t >= 0; --t) {
    if (S->levels[t].n > 4096 || (int)S->levels.size() - t > TAIL_MAX_LEVELS) break;
    const int count = (int)S->levels.size() - t;
    const size_t bytes = gc_tail_ok(opt) ? gc_tail_lds_bytes(&S->levels[t], count)
                                         : tail_lds_bytes(&S->levels[t], count, opt.smoother);
    if (bytes > TAIL_LDS_FLOATS * sizeof(float)) break;
    S->tail_first = t;
    S->tail_lds = bytes;
  }
  if (gc_tail_ok(opt) && S->tail_first < (int)S->levels.size()) {
    size_t floats = 0;
    for (int l = S->tail_first; l < (int)S->levels.size(); ++l) floats += 4 * S->levels[l].n;
    if ((st = dalloc(S, &S->tail_pack, floats))) return fail(st);
  }
  // Compact Galerkin levels: every level above the tail whose couplings are
  // exact in binary16 (level <= 5) and whose rows split into quads.
  S->cq.assign(S->levels.size(), QuadGrid{});
  S->cqr.assign(S->levels.size(), QuadGrid{});
  if (opt.coarse_operator == FSP_COARSE_GALERKIN && opt.preconditioner != FSP_PRECOND_JACOBI &&
      quad_vcycle_ok(S->quad, opt)) {
    for (int l = 0; l < S->tail_first && l < 5; ++l) {
      Level& L = S->levels[l];
      if (L.nx % 4 != 0) continue;
      QuadGrid q = make_q(L.nx / 4, L.ny, L.nz);
      const size_t blocks = (size_t)q.nbx * q.nby * q.planes;
      if (blocks > 0xFFFFFFFFu) continue;
      if ((st = dalloc(S, &L.cx, L.n)) || (st = dalloc(S, &L.cy, L.n)) || (st = dalloc(S, &L.cz, L.n)) ||
          (st = dalloc(S, &L.dg, L.n)))
        return fail(st);
      L.compact = 1;
      S->cq[l] = q;
      if (l + 1 < (int)S->levels.size()) S->cqr[l] = make_q(L.nx / 4, L.ny / 2, L.nz / 2);
    }
  }
  // 16-bit V-cycle storage exists on the red-black quad path only
  if (opt.vcycle_storage != FSP_STORE_F32 &&
      !(opt.preconditioner != FSP_PRECOND_JACOBI && quad_vcycle_ok(S->quad, opt)))
    return fail(FSP_INVALID);
  // the network's output is an f32 vector; it shares z with the V-cycle
  if (opt.preconditioner == FSP_PRECOND_LEARNED && opt.vcycle_storage != FSP_STORE_F32) return fail(FSP_INVALID);
  S->zres = S->z;
  // the surface step's buffers and the iteration graph: nothing is reserved
  // after this call (I-06)
  if ((st = step_reserve(S)) || (st = prepare_graph(S, 0))) return fail(st);
  *out = S;
  return FSP_OK;
}

void fsp_destroy(fsp_solver* s) {
  if (!s) return;
  (void)hipSetDevice(s->opt.device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  cleanup(s);
  delete s;
}

int32_t fsp_levels(const fsp_solver* s) { return s ? (int32_t)s->levels.size() : -1; }

fsp_status fsp_set_geometry(fsp_solver* S, const fsp_surface* surf, const float* eta, size_t n_eta,
                            const float* eta_roundoff, const uint8_t* solid, size_t n_solid, const float* open_u,
                            size_t n_open_u, const float* open_v, size_t n_open_v, const float* open_w,
                            size_t n_open_w) {
  if (!S || !surf) return FSP_INVALID;
  if (!eta || n_eta != S->ncols) return FSP_SHAPE;
  if ((solid && n_solid != S->n) || (open_u && n_open_u != S->nu) || (open_v && n_open_v != S->nv) ||
      (open_w && n_open_w != S->nw))
    return FSP_SHAPE;
  if (!std::isfinite(surf->rho) || !(surf->rho > 0.f) || !std::isfinite(surf->g_eff) || !std::isfinite(surf->rest))
    return FSP_NOT_FINITE;
  if (!all_finite(eta, S->ncols) || (eta_roundoff && !all_finite(eta_roundoff, S->ncols))) return FSP_NOT_FINITE;
  // apertures in [0, 1]; any other than 0 or 1 makes the geometry fractional
  const float* opens[3] = {open_u, open_v, open_w};
  const size_t lens[3] = {S->nu, S->nv, S->nw};
  int frac = 0;
  for (int a = 0; a < 3; ++a)
    if (opens[a])
      for (size_t q = 0; q < lens[a]; ++q) {
        const float x = opens[a][q];
        if (!std::isfinite(x)) return FSP_NOT_FINITE;
        if (!(x >= 0.f && x <= 1.f)) return FSP_DOMAIN;
        if (x != 0.f && x != 1.f) frac = 1;
      }
  // the Galerkin coarse operator is built for unit couplings (its compact
  // levels hold them in binary16); fractional apertures take the
  // re-discretised one, the reference's
  if (frac && S->opt.coarse_operator == FSP_COARSE_GALERKIN && S->opt.preconditioner != FSP_PRECOND_JACOBI) {
    std::fprintf(stderr, "fsp: fractional apertures need coarse_operator = FSP_COARSE_REDISCRETISED\n");
    return FSP_DOMAIN;
  }
  // face code: 0 closed, 1 open, 2 fractional (the value in au/av/aw)
  auto code = [](float x) -> uint8_t { return x == 0.f ? 0 : (x == 1.f ? 1 : 2); };
  const size_t nx = S->dom.nx, ny = S->dom.ny, nz = S->dom.nz;
  const int cut = (solid || open_u || open_v || open_w || surf->closed_lid) ? 1 : 0;
  std::vector<uint8_t> hs(S->n, 0), hu(S->nu, 1), hv(S->nv, 1), hw(S->nw, 1);
  if (solid) for (size_t q = 0; q < S->n; ++q) hs[q] = solid[q] ? 1 : 0;
  if (cut) {
    // Derived apertures: a face touching a solid cell or the box walls is
    // closed, the lid open unless closed_lid; explicit apertures win.
    auto Sd = [&](size_t i, size_t j, size_t k) { return hs[(k * ny + j) * nx + i] != 0; };
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i <= nx; ++i) {
          const size_t f = (k * ny + j) * (nx + 1) + i;
          uint8_t a = (i == 0 || i == nx || Sd(i - 1, j, k) || Sd(i, j, k)) ? 0 : 1;
          hu[f] = open_u ? code(open_u[f]) : a;
        }
    for (size_t k = 0; k < nz; ++k)
      for (size_t j = 0; j <= ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          const size_t f = (k * (ny + 1) + j) * nx + i;
          uint8_t a = (j == 0 || j == ny || Sd(i, j - 1, k) || Sd(i, j, k)) ? 0 : 1;
          hv[f] = open_v ? code(open_v[f]) : a;
        }
    for (size_t k = 0; k <= nz; ++k)
      for (size_t j = 0; j < ny; ++j)
        for (size_t i = 0; i < nx; ++i) {
          const size_t f = (k * ny + j) * nx + i;
          uint8_t a;
          if (k == 0) a = 0;
          else if (k == nz) a = (Sd(i, j, nz - 1) || surf->closed_lid) ? 0 : 1;
          else a = (Sd(i, j, k - 1) || Sd(i, j, k)) ? 0 : 1;
          uint8_t val = open_w ? code(open_w[f]) : a;
          if (k == nz && surf->closed_lid) val = 0;
          hw[f] = val;
        }
  }
  // column floors of the surface-step bounds (the reference's cut geometry
  // raises a column's floor to the top of the highest cell a solid touches,
  // delta3d_cut.rs:429-447)
  std::vector<float> floors(S->ncols, 0.f);
  for (size_t k = 0; k < nz; ++k)
    for (size_t c = 0; c < S->ncols; ++c)
      if (hs[k * S->ncols + c]) floors[c] = (float)(k + 1) * S->dom.dx;
  FSP_HIP(hipSetDevice(S->opt.device));
  if (frac) {
    // the fractional values, reserved here at configuration (set_geometry
    // precedes every projection) the first time a geometry has them
    if (!S->au) {
      fsp_status as;
      if ((as = dalloc(S, &S->au, S->nu)) || (as = dalloc(S, &S->av, S->nv)) || (as = dalloc(S, &S->aw, S->nw)))
        return as;
    }
    float* dst[3] = {S->au, S->av, S->aw};
    for (int a = 0; a < 3; ++a) {
      if (opens[a]) FSP_HIP(hipMemcpyAsync(dst[a], opens[a], lens[a] * sizeof(float), hipMemcpyHostToDevice, S->stream));
      else FSP_HIP(hipMemsetAsync(dst[a], 0, lens[a] * sizeof(float), S->stream));
    }
  }
  FSP_HIP(hipMemcpyAsync(S->col_floor, floors.data(), S->ncols * sizeof(float), hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->solid, hs.data(), S->n, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->ou, hu.data(), S->nu, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->ov, hv.data(), S->nv, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->ow, hw.data(), S->nw, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->eta, eta, S->ncols * sizeof(float), hipMemcpyHostToDevice, S->stream));
  if (eta_roundoff) {
    FSP_HIP(hipMemcpyAsync(S->eta_rt, eta_roundoff, S->ncols * sizeof(float), hipMemcpyHostToDevice, S->stream));
  } else {
    FSP_HIP(hipMemsetAsync(S->eta_rt, 0, S->ncols * sizeof(float), S->stream));
  }
  FSP_HIP(hipStreamSynchronize(S->stream));
  S->surf = *surf;
  S->cut = cut;
  // a coupled surface stays on over a new column geometry (the coupled
  // step sets its surface every step) but not over fractional apertures,
  // which the coupled step refuses (delta3d_coupling.rs:729)
  const bool recapture = S->frac != frac || S->levelset || (S->coupled && frac);
  S->frac = frac;
  S->levelset = 0;
  if (frac) S->coupled = 0;
  S->geom = true;
  if (recapture) return regraph(S);
  return FSP_OK;
}

fsp_status fsp_set_level_set(fsp_solver* S, const uint8_t* labels, size_t n_labels, const float* phi, size_t n_phi,
                             float theta_min) {
  if (!S || !labels || !phi) return FSP_INVALID;
  if (n_labels != S->n || n_phi != S->n) return FSP_SHAPE;
  if (!std::isfinite(theta_min)) return FSP_NOT_FINITE;
  if (!(theta_min > 0.f && theta_min <= 1.f)) return FSP_DOMAIN;
  if (!all_finite(phi, S->n)) return FSP_NOT_FINITE;
  for (size_t q = 0; q < S->n; ++q)
    if (labels[q] > FSP_LABEL_SOLID) return FSP_DOMAIN;
  const size_t nx = S->dom.nx, ny = S->dom.ny, nz = S->dom.nz;
  std::vector<uint8_t> hs(S->n), hwat(S->n), hu(S->nu), hv(S->nv), hw(S->nw);
  for (size_t q = 0; q < S->n; ++q) {
    hs[q] = labels[q] == FSP_LABEL_SOLID;
    hwat[q] = labels[q] == FSP_LABEL_WATER;
  }
  // the domain's boundary is a wall (apic3d.rs:786-806); every inner face
  // open: a face next to a solid carries the body's velocity into the
  // divergence and is skipped by the operator and the correction
  for (size_t k = 0; k < nz; ++k)
    for (size_t j = 0; j < ny; ++j)
      for (size_t i = 0; i <= nx; ++i) hu[(k * ny + j) * (nx + 1) + i] = (i == 0 || i == nx) ? 0 : 1;
  for (size_t k = 0; k < nz; ++k)
    for (size_t j = 0; j <= ny; ++j)
      for (size_t i = 0; i < nx; ++i) hv[(k * (ny + 1) + j) * nx + i] = (j == 0 || j == ny) ? 0 : 1;
  for (size_t k = 0; k <= nz; ++k)
    for (size_t j = 0; j < ny; ++j)
      for (size_t i = 0; i < nx; ++i) hw[(k * ny + j) * nx + i] = (k == 0 || k == nz) ? 0 : 1;
  FSP_HIP(hipSetDevice(S->opt.device));
  if (!S->water) {
    fsp_status as;
    if ((as = dalloc(S, &S->water, S->n)) || (as = dalloc(S, &S->phi, S->n))) return as;
  }
  FSP_HIP(hipMemcpyAsync(S->solid, hs.data(), S->n, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->water, hwat.data(), S->n, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->phi, phi, S->n * sizeof(float), hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->ou, hu.data(), S->nu, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->ov, hv.data(), S->nv, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->ow, hw.data(), S->nw, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemsetAsync(S->eta, 0, S->ncols * sizeof(float), S->stream));
  FSP_HIP(hipMemsetAsync(S->eta_rt, 0, S->ncols * sizeof(float), S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  fsp_surface surf{};
  surf.rho = S->surf.rho > 0.f ? S->surf.rho : 1.f;
  surf.homogeneous_ghost = 1;  // p = 0 on the zero level
  surf.closed_lid = 1;
  S->surf = surf;
  S->cut = 1;
  const bool recapture = !S->levelset || S->frac || S->coupled || S->theta_min != theta_min;
  S->frac = 0;
  S->levelset = 1;
  S->coupled = 0;
  S->theta_min = theta_min;
  S->geom = true;
  if (recapture) return regraph(S);
  return FSP_OK;
}

namespace {
// fsp_set_coupling and fsp_set_coupling_device: copies of kind `kind`.
fsp_status set_coupling(fsp_solver* S, const float* total, size_t n_total, const float* bg_up, size_t n_up,
                        const float* bg_x, size_t n_x, const float* bg_y, size_t n_y, hipMemcpyKind kind) {
  if (!S) return FSP_INVALID;
  if (!S->geom || S->levelset || S->frac) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  if (!total) {
    const bool was = S->coupled != 0;
    S->coupled = 0;
    return was ? regraph(S) : FSP_OK;
  }
  if (n_total != S->ncols || (bg_up && n_up != S->ncols) || (bg_x && n_x != S->nu) || (bg_y && n_y != S->nv))
    return FSP_SHAPE;
  if (kind == hipMemcpyHostToDevice &&
      (!all_finite(total, S->ncols) || (bg_up && !all_finite(bg_up, S->ncols)) ||
       (bg_x && !all_finite(bg_x, S->nu)) || (bg_y && !all_finite(bg_y, S->nv))))
    return FSP_NOT_FINITE;
  if (!S->total) {
    fsp_status as;
    if ((as = dalloc(S, &S->total, S->ncols)) || (as = dalloc(S, &S->bg_up, S->ncols)) ||
        (as = dalloc(S, &S->bg_x, S->nu)) || (as = dalloc(S, &S->bg_y, S->nv)))
      return as;
  }
  const float* src[4] = {total, bg_up, bg_x, bg_y};
  float* dst[4] = {S->total, S->bg_up, S->bg_x, S->bg_y};
  const size_t len[4] = {S->ncols, S->ncols, S->nu, S->nv};
  for (int a = 0; a < 4; ++a) {
    if (src[a]) FSP_HIP(hipMemcpyAsync(dst[a], src[a], len[a] * sizeof(float), kind, S->stream));
    else FSP_HIP(hipMemsetAsync(dst[a], 0, len[a] * sizeof(float), S->stream));
  }
  FSP_HIP(hipStreamSynchronize(S->stream));
  const bool recapture = !S->coupled;
  S->coupled = 1;
  if (recapture) return regraph(S);
  return FSP_OK;
}
}  // namespace

fsp_status fsp_set_coupling(fsp_solver* S, const float* total, size_t n_total, const float* bg_up, size_t n_up,
                            const float* bg_x, size_t n_x, const float* bg_y, size_t n_y) {
  return set_coupling(S, total, n_total, bg_up, n_up, bg_x, n_x, bg_y, n_y, hipMemcpyHostToDevice);
}

fsp_status fsp_set_coupling_device(fsp_solver* S, const float* d_total, const float* d_bg_up, const float* d_bg_x,
                                   const float* d_bg_y) {
  if (!S) return FSP_INVALID;
  return set_coupling(S, d_total, S->ncols, d_bg_up, S->ncols, d_bg_x, S->nu, d_bg_y, S->nv,
                      hipMemcpyDeviceToDevice);
}

fsp_status fsp_enable_graded(fsp_solver* S, const int32_t* nodes, size_t count) {
  if (!S) return FSP_INVALID;
  if (!nodes || count == 0) {
    S->gnodes.clear();
    return FSP_OK;
  }
  const int nz = S->dom.nz;
  // the nodes start at layer 0, end at the top layer and increase strictly
  // (delta3d_graded.rs:66-68)
  if (nodes[0] != 0 || nodes[count - 1] != nz - 1) return FSP_SHAPE;
  for (size_t q = 1; q < count; ++q)
    if (nodes[q] <= nodes[q - 1]) return FSP_SHAPE;
  // no cut geometry (:70-72)
  if (S->geom && (S->cut || S->levelset)) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  const size_t len = S->ncols * count;
  if (len > S->gcap) {
    float** vecs[7] = {&S->gx, &S->gr, &S->gd, &S->gq, &S->gb, &S->gm, &S->gz};
    for (float** v : vecs)
      if (fsp_status as = dalloc(S, v, len)) return as;
    S->gcap = len;
  }
  if (!S->g_node) {
    fsp_status as;
    if ((as = dalloc(S, &S->g_node, (size_t)nz)) || (as = dalloc(S, &S->g_seg, (size_t)nz)) ||
        (as = dalloc(S, &S->g_wa, (size_t)nz)) || (as = dalloc(S, &S->g_wb, (size_t)nz)) ||
        (as = dalloc(S, &S->g_weight, (size_t)nz)) || (as = dalloc(S, &S->gpart, (size_t)2 * MAX_PARTIALS)) ||
        (as = dalloc(S, &S->gout, (size_t)4)))
      return as;
    FSP_HIP(hipHostMalloc((void**)&S->hgout, 4 * sizeof(double), hipHostMallocDefault));
  }
  // per layer its segment and (1 - t, t), t = (k - a) / (b - a) in f32
  // (:48-53); per node its weight sum_k P_kj in the reference's order (:81-91)
  std::vector<int> seg((size_t)nz, 0);
  std::vector<float> wa((size_t)nz, 0.f), wb((size_t)nz, 0.f), weight(count, 0.f);
  for (size_t sg = 0; sg + 1 < count; ++sg) {
    const int a = nodes[sg], b = nodes[sg + 1];
    for (int k = a; k < b; ++k) {
      const float t = (float)(k - a) / (float)(b - a);
      wa[k] = 1.0f - t;
      wb[k] = t;
      seg[k] = (int)sg;
      weight[sg] += wa[k];
      weight[sg + 1] += wb[k];
    }
  }
  weight[count - 1] += 1.0f;
  FSP_HIP(hipMemcpy(S->g_node, nodes, count * sizeof(int32_t), hipMemcpyHostToDevice));
  FSP_HIP(hipMemcpy(S->g_seg, seg.data(), (size_t)nz * sizeof(int), hipMemcpyHostToDevice));
  FSP_HIP(hipMemcpy(S->g_wa, wa.data(), (size_t)nz * sizeof(float), hipMemcpyHostToDevice));
  FSP_HIP(hipMemcpy(S->g_wb, wb.data(), (size_t)nz * sizeof(float), hipMemcpyHostToDevice));
  FSP_HIP(hipMemcpy(S->g_weight, weight.data(), count * sizeof(float), hipMemcpyHostToDevice));
  S->gnodes.assign(nodes, nodes + count);
  return FSP_OK;
}

int32_t fsp_pressure_unknowns_per_column(const fsp_solver* S) {
  if (!S) return -1;
  return S->gnodes.empty() ? S->dom.nz : (int32_t)S->gnodes.size();
}

fsp_status fsp_refine_device(fsp_solver* S, float* d_p, float* d_u, float* d_v, float* d_w, float scale, float k1,
                             double rel_tol, uint32_t max_iters, fsp_report* report) {
  if (!S || !d_p || !d_u || !d_v || !d_w) return FSP_INVALID;
  if (!std::isfinite(scale) || !std::isfinite(k1) || !(rel_tol > 0.0)) return FSP_NOT_FINITE;
  if (!S->geom) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  // base pressure and the corrected faces as the new predicted faces
  // (delta3d_coupling.rs:233-237), in the solver's staging buffers
  FSP_HIP(hipMemcpyAsync(S->p, d_p, S->n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->us, d_u, S->nu * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->vs, d_v, S->nv * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->ws, d_w, S->nw * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
  // homogeneous ghosts from p = 0 (:238-240); the iteration graphs read no
  // ghost value, so they stay as captured
  const int saved = S->surf.homogeneous_ghost;
  S->surf.homogeneous_ghost = 1;
  fsp_report R{};
  const fsp_status st = run(S, S->us, S->vs, S->ws, d_p, 0, d_u, d_v, d_w, scale, k1, rel_tol, max_iters, &R);
  S->surf.homogeneous_ghost = saved;
  if (st != FSP_OK) {
    // the base pressure back in d_p
    FSP_HIP(hipMemcpyAsync(d_p, S->p, S->n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
    FSP_HIP(hipStreamSynchronize(S->stream));
    return st;
  }
  hipLaunchKernelGGL(k_add_base, dim3(grid_for(S->n)), dim3(BLOCK), 0, S->stream, d_p, (const float*)S->p, S->n);
  FSP_HIP(hipGetLastError());
  FSP_HIP(hipStreamSynchronize(S->stream));
  R.refinements = 1;
  if (report) *report = R;
  return FSP_OK;
}

fsp_status fsp_project_device(fsp_solver* S, const float* d_us, const float* d_vs, const float* d_ws, float* d_p,
                              int32_t p_warm, float* d_u, float* d_v, float* d_w, float scale, float k1,
                              double rel_tol, uint32_t max_iters, fsp_report* report) {
  if (!S || !d_us || !d_vs || !d_ws || !d_p || !d_u || !d_v || !d_w) return FSP_INVALID;
  if (!std::isfinite(scale) || !std::isfinite(k1) || !(rel_tol > 0.0)) return FSP_NOT_FINITE;
  FSP_HIP(hipSetDevice(S->opt.device));
  return run(S, d_us, d_vs, d_ws, d_p, p_warm, d_u, d_v, d_w, scale, k1, rel_tol, max_iters, report);
}

fsp_status fsp_project(fsp_solver* S, const float* us, size_t n_us, const float* vs, size_t n_vs, const float* ws,
                       size_t n_ws, float* p, size_t n_p, int32_t p_warm, float* u, size_t n_u, float* v, size_t n_v,
                       float* w, size_t n_w, float scale, float k1, double rel_tol, uint32_t max_iters,
                       fsp_report* report) {
  if (!S || !us || !vs || !ws || !p || !u || !v || !w) return FSP_INVALID;
  if (n_us != S->nu || n_vs != S->nv || n_ws != S->nw || n_p != S->n || n_u != S->nu || n_v != S->nv ||
      n_w != S->nw)
    return FSP_SHAPE;
  if (!std::isfinite(scale) || !std::isfinite(k1) || !(rel_tol > 0.0)) return FSP_NOT_FINITE;
  FSP_HIP(hipSetDevice(S->opt.device));
  auto t0 = std::chrono::steady_clock::now();
  FSP_HIP(hipMemcpyAsync(S->us, us, S->nu * 4, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->vs, vs, S->nv * 4, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipMemcpyAsync(S->ws, ws, S->nw * 4, hipMemcpyHostToDevice, S->stream));
  if (p_warm) FSP_HIP(hipMemcpyAsync(S->p, p, S->n * 4, hipMemcpyHostToDevice, S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  auto t1 = std::chrono::steady_clock::now();
  fsp_report R{};
  fsp_status st = run(S, S->us, S->vs, S->ws, S->p, p_warm, S->u, S->v, S->w, scale, k1, rel_tol, max_iters, &R);
  if (st != FSP_OK) return st;
  auto t2 = std::chrono::steady_clock::now();
  FSP_HIP(hipMemcpyAsync(p, S->p, S->n * 4, hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipMemcpyAsync(u, S->u, S->nu * 4, hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipMemcpyAsync(v, S->v, S->nv * 4, hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipMemcpyAsync(w, S->w, S->nw * 4, hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  auto t3 = std::chrono::steady_clock::now();
  R.ms_transfer = std::chrono::duration<double, std::milli>((t1 - t0) + (t3 - t2)).count();
  if (report) *report = R;
  return FSP_OK;
}

fsp_status fsp_apply_operator_device(fsp_solver* S, const float* d_p, float* d_out) {
  if (!S || !d_p || !d_out) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  build_geometry(S, false);
  Launch K(S);
  FSP_LAUNCH(S, k_apply, dim3(grid_for(K.G.n)), K.st, K.G, d_p, d_out);
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status fsp_assemble_device(fsp_solver* S, const float* d_div, float scale, float* d_rhs, float* d_prec) {
  if (!S || !d_div || !d_rhs || !d_prec) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  build_geometry(S, false);
  Launch K(S);
  if (S->quad && aligned16(d_div) && aligned16(d_rhs) && aligned16(d_prec)) {
    const dim3 gq((unsigned)(S->qfine.nbx * S->qfine.nby * S->qfine.planes));
    FSP_LAUNCH(S, k_assemble_q, gq, K.st, K.G, S->qfine, (const float*)nullptr, (const float*)nullptr,
               (const float*)nullptr, d_div, scale, d_rhs, d_prec, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_B2, gq, false, 0);
  } else {
    FSP_LAUNCH(S, k_assemble, dim3(grid_for(K.G.n)), K.st, K.G, (const float*)nullptr,
               (const float*)nullptr, (const float*)nullptr, d_div, scale, d_rhs, d_prec, S->ctl, S->partials);
    FSP_FINISH(S, K.st, OP_B2, dim3(grid_for(K.G.n)), false, 0);
  }
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status fsp_attach_net(fsp_solver* S, fsp_net* net) {
  if (!S) return FSP_INVALID;
  if (net && !net_shape_ok(S->dom.nx, S->dom.ny, S->dom.nz)) return FSP_SHAPE;
  FSP_HIP(hipSetDevice(S->opt.device));
  FSP_HIP(hipStreamSynchronize(S->stream));
  S->net = net;
  // graphs of the learned phase hold the previous network's buffers
  drop_graph(S, 1);
  if (!net) return FSP_OK;
  fsp_status st = net_reserve(net, S->dom.nx, S->dom.ny, S->dom.nz);
  if (st == FSP_OK && S->opt.preconditioner == FSP_PRECOND_LEARNED) st = prepare_graph(S, 1);
  return st;
}

fsp_status fsp_net_features_device(fsp_net* net, fsp_solver* S, const float* d_r, uint16_t* d_x, float sigma,
                                   float* sigma_out) {
  if (!net || !S || !d_r || !d_x) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  fsp_status st = net_reserve(net, S->dom.nx, S->dom.ny, S->dom.nz);
  if (st) return st;
  build_geometry(S, false);
  Launch K(S);
  const double* rr = nullptr;
  if (!(sigma > 0.f)) {
    // |r|^2 over wet cells into the control block (OP_GENERIC slot)
    FSP_LAUNCH(S, k_dot, dim3(grid_for(S->n)), S->stream, S->n, d_r, d_r, OP_GENERIC, 0, S->ctl, S->partials);
    FSP_FINISH(S, S->stream, OP_GENERIC, dim3(grid_for(S->n)), false, 0);
    rr = &S->ctl->red[OP_GENERIC][0];
  }
  if ((st = net_features(net, S->stream, d_r, S->mask, S->solid, rr, &S->ctl->wet_cells, sigma, d_x))) return st;
  float s_used = sigma;
  if (sigma_out) {
    FSP_HIP(hipStreamSynchronize(S->stream));
    if (!(sigma > 0.f)) {
      if ((st = sync_ctl(S))) return st;
      const double r2 = S->hctl->red[OP_GENERIC][0];
      const double nw = S->hctl->wet_cells > 0 ? (double)S->hctl->wet_cells : 1.0;
      s_used = r2 > 0.0 ? (float)std::sqrt(nw / r2) : 1.0f;
    }
    *sigma_out = s_used;
  }
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status fsp_precondition_device(fsp_solver* S, const float* d_r, float* d_z) {
  if (!S || !d_r || !d_z) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  build_geometry(S, true);
  // the fine kernels' form a projection on this geometry takes (the
  // table-only one when every ghost row lies in its column's table)
  if (const fsp_status st = sync_ctl(S)) return st;
  struct TableGuard {
    fsp_solver* S;
    ~TableGuard() { S->table_only = 0; }
  } table_guard{S};
  S->table_only = (!S->frac && !S->levelset && !S->hctl->table_miss) ? 1 : 0;
  Launch K(S);
  FSP_HIP(hipMemcpyAsync(S->r, d_r, S->n * 4, hipMemcpyDeviceToDevice, S->stream));
  K.set_pass(1, 0.0, 0.0, 0.0, 0, 1);
  K.precondition(OP_GENERIC);
  K.export_z(d_z);
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status fsp_divergence_device(fsp_solver* S, const float* d_u, const float* d_v, const float* d_w, float* d_out) {
  if (!S || !d_u || !d_v || !d_w || !d_out) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  Launch K(S);
  FSP_LAUNCH(S, k_divergence, dim3(grid_for(K.G.n)), K.st, K.G, d_u, d_v, d_w, d_out);
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

// ---------------------------------------------------------------- the surface step
namespace {

fsp_status step_reserve(fsp_solver* S) {
  if (S->step_sc) return FSP_OK;
  const size_t nx = S->dom.nx, ny = S->dom.ny;
  fsp_status st;
  if ((st = dalloc(S, &S->sv_u, S->nu)) || (st = dalloc(S, &S->sv_v, S->nv)) || (st = dalloc(S, &S->sv_w, S->nw)) ||
      (st = dalloc(S, &S->sv_p, S->n)) || (st = dalloc(S, &S->sv_eta, S->ncols)) ||
      (st = dalloc(S, &S->sv_rt, S->ncols)) || (st = dalloc(S, &S->flux_x, (nx + 1) * ny)) ||
      (st = dalloc(S, &S->flux_y, nx * (ny + 1))))
    return st;
  StepScalars* sc = nullptr;
  if ((st = dalloc(S, &sc, 1))) return st;
  FSP_HIP(hipHostMalloc((void**)&S->hstep, sizeof(StepScalars), hipHostMallocDefault));
  S->step_sc = sc;
  return FSP_OK;
}

// The step's scalars of the current state (flux arrays: also the wall sum),
// copied to S->hstep.
fsp_status step_scalars(fsp_solver* S, bool with_flux) {
  const Fine G = S->fine();
  k_step_reset<<<1, 1, 0, S->stream>>>(S->step_sc);
  k_step_scalars<<<1, STEP_THREADS, 0, S->stream>>>(G, S->col_floor, with_flux ? S->flux_x : nullptr,
                                                      with_flux ? S->flux_y : nullptr, S->step_sc);
  FSP_HIP(hipGetLastError());
  FSP_HIP(hipMemcpyAsync(S->hstep, S->step_sc, sizeof(StepScalars), hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status step_copy(fsp_solver* S, float* dst, const float* src, size_t n) {
  FSP_HIP(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, S->stream));
  return FSP_OK;
}

fsp_status advect(fsp_solver* S, const float* u, const float* v, const float* w, float dt, float* us, float* vs,
                  float* ws) {
  const Fine G = S->fine();
  const float h = 0.5f / G.dx;
  FSP_LAUNCH(S, k_advect_h<0>, dim3(grid_items(S->nu)), S->stream, G, u, v, w, dt, h, us);
  FSP_LAUNCH(S, k_advect_h<1>, dim3(grid_items(S->nv)), S->stream, G, u, v, w, dt, h, vs);
  FSP_LAUNCH(S, k_advect_w, dim3(grid_items(S->nw)), S->stream, G, u, v, w, dt, h, ws);
  FSP_HIP(hipGetLastError());
  return FSP_OK;
}

fsp_status extrapolate(fsp_solver* S, float* u, float* v, float* w) {
  const Fine G = S->fine();
  FSP_LAUNCH(S, k_extrapolate, dim3(grid_for(3 * S->ncols)), S->stream, G, u, v, w);
  FSP_HIP(hipGetLastError());
  return FSP_OK;
}

fsp_status transport(fsp_solver* S, const float* u, const float* v, float factor) {
  const Fine G = S->fine();
  const size_t nx = S->dom.nx, ny = S->dom.ny;
  FSP_LAUNCH(S, k_flux, dim3(grid_for((nx + 1) * ny + nx * (ny + 1))), S->stream, G, u, v, S->flux_x, S->flux_y);
  FSP_LAUNCH(S, k_heights, dim3(grid_for(S->ncols)), S->stream, G, (const float*)S->flux_x, (const float*)S->flux_y,
             factor, S->eta, S->eta_rt);
  FSP_HIP(hipGetLastError());
  return FSP_OK;
}

}  // namespace

fsp_status fsp_step_surface_mobile_device(fsp_solver* S, float* d_u, float* d_v, float* d_w, float* d_p,
                                          uint64_t duration_us, uint32_t max_iters, fsp_report* report,
                                          fsp_balance* balance) {
  if (!S || !d_u || !d_v || !d_w || !d_p) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  if (S->frac || S->levelset || S->coupled) return FSP_DOMAIN;  // built for uncoupled column geometries
  if (duration_us == 0 || duration_us > (1ull << 53)) return FSP_NOT_FINITE;
  const double dt = (double)duration_us * 1e-6;
  const float dx = S->dom.dx, rho = S->surf.rho, g = S->surf.g_eff;
  FSP_HIP(hipSetDevice(S->opt.device));
  fsp_status st = step_reserve(S);
  if (st) return st;
  if ((st = step_scalars(S, false))) return st;
  if (g <= 0.f || dt * dt * (double)g / (double)dx > 1.0 || !S->hstep->in_bounds) return FSP_DOMAIN;
  const float scale = (float)(-(double)rho / dt);
  const float correction = (float)(dt / (double)rho);
  const float factor = (float)(dt / (double)dx);
  const float advection = (float)dt;
  if (!std::isfinite(scale) || !std::isfinite(correction) || correction == 0.f || !std::isfinite(factor) ||
      factor == 0.f || !(advection > 0.f))
    return FSP_NOT_FINITE;
  const double volume_before = S->hstep->volume;
  // state a refusal restores
  if ((st = step_copy(S, S->sv_u, d_u, S->nu)) || (st = step_copy(S, S->sv_v, d_v, S->nv)) ||
      (st = step_copy(S, S->sv_w, d_w, S->nw)) || (st = step_copy(S, S->sv_p, d_p, S->n)) ||
      (st = step_copy(S, S->sv_eta, S->eta, S->ncols)) || (st = step_copy(S, S->sv_rt, S->eta_rt, S->ncols)))
    return st;
  auto restore = [&](fsp_status why) -> fsp_status {
    if (step_copy(S, d_u, S->sv_u, S->nu) || step_copy(S, d_v, S->sv_v, S->nv) || step_copy(S, d_w, S->sv_w, S->nw) ||
        step_copy(S, d_p, S->sv_p, S->n) || step_copy(S, S->eta, S->sv_eta, S->ncols) ||
        step_copy(S, S->eta_rt, S->sv_rt, S->ncols) || hipStreamSynchronize(S->stream) != hipSuccess)
      return FSP_DEVICE;
    return why;
  };
  if ((st = advect(S, d_u, d_v, d_w, advection, S->us, S->vs, S->ws))) return restore(st);
  fsp_report R{};
  st = run(S, S->us, S->vs, S->ws, d_p, 1, d_u, d_v, d_w, scale, correction, 1e-6, max_iters, &R);
  if (report) *report = R;
  if (st) return restore(st);
  if (R.degraded) return restore(FSP_CONVERGENCE);
  if ((st = extrapolate(S, d_u, d_v, d_w)) || (st = transport(S, d_u, d_v, factor))) return restore(st);
  // scalars after the step, and the finiteness of every field the step wrote
  const Fine G = S->fine();
  k_step_reset<<<1, 1, 0, S->stream>>>(S->step_sc);
  k_step_scalars<<<1, STEP_THREADS, 0, S->stream>>>(G, S->col_floor, S->flux_x, S->flux_y, S->step_sc);
  const FiniteSet F{{d_u, d_v, d_w, d_p, S->eta, S->eta_rt, S->us, S->vs, S->ws, S->b, S->r},
                    {S->nu, S->nv, S->nw, S->n, S->ncols, S->ncols, S->nu, S->nv, S->nw, S->n, S->n}};
  size_t longest = 0;
  for (size_t len : F.n) longest = std::max(longest, len);
  k_finite<<<dim3(grid_for(longest), FINITE_FIELDS), BLOCK, 0, S->stream>>>(F, S->step_sc);
  FSP_HIP(hipGetLastError());
  FSP_HIP(hipMemcpyAsync(S->hstep, S->step_sc, sizeof(StepScalars), hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  const StepScalars after = *S->hstep;
  if (!after.finite || !std::isfinite(R.residual) || !std::isfinite(R.divergence)) return restore(FSP_NOT_FINITE);
  if (!after.in_bounds) return restore(FSP_DOMAIN);
  if (balance) {
    const double perturbation_in = after.boundary * dt * (double)dx;
    const double delta = after.volume - volume_before;
    *balance = fsp_balance{after.volume, delta, 0.0, perturbation_in, 0.0, delta - perturbation_in, 0.0};
  }
  return FSP_OK;
}

fsp_status fsp_get_surface(fsp_solver* S, float* eta, float* eta_roundoff) {
  if (!S || !eta) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  FSP_HIP(hipMemcpyAsync(eta, S->eta, S->ncols * sizeof(float), hipMemcpyDeviceToHost, S->stream));
  if (eta_roundoff)
    FSP_HIP(hipMemcpyAsync(eta_roundoff, S->eta_rt, S->ncols * sizeof(float), hipMemcpyDeviceToHost, S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status fsp_advect_device(fsp_solver* S, const float* d_u, const float* d_v, const float* d_w, float dt,
                             float* d_us, float* d_vs, float* d_ws) {
  if (!S || !d_u || !d_v || !d_w || !d_us || !d_vs || !d_ws) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  if (S->frac || S->levelset || S->coupled) return FSP_DOMAIN;  // built for uncoupled column geometries
  if (!std::isfinite(dt)) return FSP_NOT_FINITE;
  FSP_HIP(hipSetDevice(S->opt.device));
  fsp_status st = advect(S, d_u, d_v, d_w, dt, d_us, d_vs, d_ws);
  if (st) return st;
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status fsp_extrapolate_device(fsp_solver* S, float* d_u, float* d_v, float* d_w) {
  if (!S || !d_u || !d_v || !d_w) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  if (S->frac || S->levelset || S->coupled) return FSP_DOMAIN;  // built for uncoupled column geometries
  FSP_HIP(hipSetDevice(S->opt.device));
  fsp_status st = extrapolate(S, d_u, d_v, d_w);
  if (st) return st;
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status fsp_transport_device(fsp_solver* S, const float* d_u, const float* d_v, float factor, float* d_flux_x,
                                float* d_flux_y) {
  if (!S || !d_u || !d_v) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  if (S->frac || S->levelset || S->coupled) return FSP_DOMAIN;  // built for uncoupled column geometries
  if (!std::isfinite(factor)) return FSP_NOT_FINITE;
  FSP_HIP(hipSetDevice(S->opt.device));
  fsp_status st = step_reserve(S);
  if (st || (st = transport(S, d_u, d_v, factor))) return st;
  const size_t nx = S->dom.nx, ny = S->dom.ny;
  if (d_flux_x) FSP_HIP(hipMemcpyAsync(d_flux_x, S->flux_x, (nx + 1) * ny * 4, hipMemcpyDeviceToDevice, S->stream));
  if (d_flux_y) FSP_HIP(hipMemcpyAsync(d_flux_y, S->flux_y, nx * (ny + 1) * 4, hipMemcpyDeviceToDevice, S->stream));
  FSP_HIP(hipStreamSynchronize(S->stream));
  return FSP_OK;
}

fsp_status fsp_surface_scalars(fsp_solver* S, double* volume, int32_t* in_bounds) {
  if (!S) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  FSP_HIP(hipSetDevice(S->opt.device));
  fsp_status st = step_reserve(S);
  if (st || (st = step_scalars(S, false))) return st;
  if (volume) *volume = S->hstep->volume;
  if (in_bounds) *in_bounds = S->hstep->in_bounds;
  return FSP_OK;
}

fsp_status fsp_project_batch(fsp_solver* const* solvers, const fsp_patch_io* io, int32_t n, double rel_tol,
                             uint32_t max_iters, int32_t concurrency, fsp_report* reports) {
  if (!solvers || !io || !reports || n < 0) return FSP_INVALID;
  if (!(rel_tol > 0.0)) return FSP_NOT_FINITE;
  for (int32_t i = 0; i < n; ++i) {
    const fsp_patch_io& q = io[i];
    if (!solvers[i] || !q.us || !q.vs || !q.ws || !q.p || !q.u || !q.v || !q.w) return FSP_INVALID;
    if (!std::isfinite(q.scale) || !std::isfinite(q.k1)) return FSP_NOT_FINITE;
    for (int32_t j = 0; j < i; ++j) {
      if (solvers[j] == solvers[i]) return FSP_INVALID;  // one patch per solver: they run concurrently
      // a network's activations belong to one forward pass at a time
      if (solvers[i]->net && solvers[j]->net == solvers[i]->net &&
          solvers[i]->opt.preconditioner == FSP_PRECOND_LEARNED &&
          solvers[j]->opt.preconditioner == FSP_PRECOND_LEARNED)
        return FSP_INVALID;
    }
  }
  const int workers = std::max(1, std::min<int>(n, concurrency > 0 ? concurrency : 4));
  std::atomic<int32_t> next{0};
  std::vector<fsp_status> status(n, FSP_OK);
  auto work = [&]() {
    for (;;) {
      const int32_t i = next.fetch_add(1);
      if (i >= n) return;
      fsp_solver* S = solvers[i];
      const fsp_patch_io& q = io[i];
      fsp_report r{};
      fsp_status st = hipSetDevice(S->opt.device) == hipSuccess ? FSP_OK : FSP_DEVICE;
      if (st == FSP_OK)
        st = run(S, q.us, q.vs, q.ws, q.p, q.p_warm, q.u, q.v, q.w, q.scale, q.k1, rel_tol, max_iters, &r);
      reports[i] = r;
      status[i] = st;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < workers; ++t) pool.emplace_back(work);
  work();
  for (std::thread& t : pool) t.join();
  for (fsp_status st : status)
    if (st != FSP_OK) return st;
  return FSP_OK;
}

fsp_status fsp_autotune(fsp_solver* S, int32_t* tile_x, double* us) {
  if (!S) return FSP_INVALID;
  if (!S->geom) return FSP_DOMAIN;
  if (!S->quad) {
    if (tile_x) *tile_x = 0;
    if (us) *us = 0.0;
    return FSP_OK;
  }
  FSP_HIP(hipSetDevice(S->opt.device));
  build_geometry(S, false);
  Launch K(S);
  const int nx = S->dom.nx, ny = S->dom.ny, nz = S->dom.nz, nq = nx / 4;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(S->n)), dim3(BLOCK), 0, K.st, S->z, S->n, 0.f);
  hipLaunchKernelGGL(k_fill, dim3(grid_for(S->n)), dim3(BLOCK), 0, K.st, S->r, S->n, 1.f);
  K.set_pass(1, 0.0, 0.0, 0.0, 0, 1);
  auto make_q = [](int nqv, int rows, int planes, int qx) {
    QuadGrid Q{};
    Q.nq = nqv; Q.QX = qx; Q.RY = QBLOCK / qx;
    Q.nbx = (nqv + qx - 1) / qx;
    Q.nby = (rows + Q.RY - 1) / Q.RY;
    Q.rows = rows; Q.planes = planes;
    return Q;
  };
  int widest = 1;
  while (widest < nq && widest < 64) widest <<= 1;
  int best = S->qfine.QX;
  double best_us = 1e30;
  const Level none_l{};
  const float* none = nullptr;
  for (int qx = widest; qx >= 4; qx >>= 1) {
    const QuadGrid Q = make_q(nq, ny, nz, qx);
    const size_t blocks = (size_t)Q.nbx * Q.nby * Q.planes;
    if (blocks > (size_t)MAX_PARTIALS) continue;
    const dim3 g((unsigned)blocks);
    for (int w = 0; w < 3; ++w)
      k_q_gs<0, ST_F32><<<g, dim3(BLOCK), 0, K.st>>>(K.G, Q, S->z, S->z, S->r, w & 1, none_l, none, -1, S->ctl,
                                                      S->partials);
    FSP_HIP(hipEventRecord(S->ev0, K.st));
    const int reps = 10;
    for (int w = 0; w < reps; ++w)
      k_q_gs<0, ST_F32><<<g, dim3(BLOCK), 0, K.st>>>(K.G, Q, S->z, S->z, S->r, w & 1, none_l, none, -1, S->ctl,
                                                      S->partials);
    FSP_HIP(hipEventRecord(S->ev1, K.st));
    FSP_HIP(hipEventSynchronize(S->ev1));
    float ms = 0.f;
    FSP_HIP(hipEventElapsedTime(&ms, S->ev0, S->ev1));
    const double t = 1000.0 * ms / reps;
    if (t < best_us) { best_us = t; best = qx; }
  }
  S->qfine = make_q(nq, ny, nz, best);
  S->qrestr = make_q(nq, ny / 2, nz / 2, best);
  // graphs hold the previous launch shapes: capture them again
  for (int gi = 0; gi < 2; ++gi) drop_graph(S, gi);
  fsp_status gst = prepare_graph(S, 0);
  if (gst == FSP_OK && S->net && S->opt.preconditioner == FSP_PRECOND_LEARNED) gst = prepare_graph(S, 1);
  if (gst) return gst;
  if (tile_x) *tile_x = best;
  if (us) *us = best_us;
  return FSP_OK;
}

fsp_status fsp_profile_begin(fsp_solver* S, const char* kernel, uint32_t max_launches) {
  if (!S || !kernel || max_launches == 0) return FSP_INVALID;
  FSP_HIP(hipSetDevice(S->opt.device));
  FSP_HIP(hipStreamSynchronize(S->stream));
  for (hipEvent_t e : S->prof_events) (void)hipEventDestroy(e);
  S->prof_events.clear();
  if (S->prof_active) (void)hipFree(S->prof_active);
  S->prof_active = nullptr;
  FSP_HIP(hipMalloc((void**)&S->prof_active, (size_t)max_launches * sizeof(int32_t)));
  S->prof_events.assign(2 * (size_t)max_launches, nullptr);
  for (auto& e : S->prof_events) FSP_HIP(hipEventCreate(&e));
  S->prof_name = kernel;
  S->prof_used = 0;
  S->prof_on = true;
  return FSP_OK;
}

fsp_status fsp_profile_end(fsp_solver* S, uint32_t* launches, double* total_ms, float* durations_ms,
                           int32_t* active, uint32_t max_durations) {
  if (!S) return FSP_INVALID;
  FSP_HIP(hipStreamSynchronize(S->stream));
  const size_t count = S->prof_used / 2;
  double total = 0.0;
  for (size_t q = 0; q < count; ++q) {
    float ms = 0.f;
    FSP_HIP(hipEventElapsedTime(&ms, S->prof_events[2 * q], S->prof_events[2 * q + 1]));
    total += ms;
    if (durations_ms && q < max_durations) durations_ms[q] = ms;
  }
  if (active && count > 0)
    FSP_HIP(hipMemcpy(active, S->prof_active, std::min<size_t>(count, max_durations) * sizeof(int32_t),
                      hipMemcpyDeviceToHost));
  if (launches) *launches = (uint32_t)count;
  if (total_ms) *total_ms = total;
  S->prof_on = false;
  S->prof_used = 0;
  return FSP_OK;
}

}  // extern "C"
