// CPU ORACLE — test infrastructure only (see oracle.hpp). C entry points for
// ctypes: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
#include "oracle.hpp"
#include "cut.hpp"
#include "apic.hpp"

#include <algorithm>

using namespace orc;

namespace {

template <class T>
Status setup(Volume<T>& vol, const Geometry* g, const Options* o) {
  Options opt = o ? *o : default_options();
  return vol.configure(*g, opt);
}

// The coupled surface's arrays (total heights and background ghost terms
// per column, u face and v face); a null term array is zeros.
struct Coupling {
  const float* total;
  const float* bg_up;
  const float* bg_x;
  const float* bg_y;
};
template <class T>
void couple(Volume<T>& vol, const Coupling* cpl) {
  const size_t cols = vol.nx * vol.ny;
  auto take = [](std::vector<T>& dst, const float* src, size_t n) {
    dst.assign(n, T(0));
    if (src) for (size_t q = 0; q < n; ++q) dst[q] = (T)src[q];
  };
  vol.coupled = true;
  take(vol.surface_total, cpl->total, cols);
  take(vol.ghost_bg_up, cpl->bg_up, cols);
  take(vol.ghost_bg_x, cpl->bg_x, vol.nu());
  take(vol.ghost_bg_y, cpl->bg_y, vol.nv());
}
// refine: 0 the projection alone; 1 step_perturbation_mobile's rule, a
// refinement after a degraded floor stop with the first solve's iterations
// added (delta3d_coupling.rs:835-840); 2 a refinement always. refine_iters
// bounds the refinement (the reference's max_iters when 0).
// The graded column's nodes (fine layers, bottom up), or none.
struct Nodes {
  const int32_t* layers;
  size_t count;
};
template <class T>
Status grade(Volume<T>& vol, const Nodes* nodes) {
  if (!nodes || !nodes->layers) return OK;
  std::vector<size_t> k(nodes->count);
  for (size_t q = 0; q < nodes->count; ++q) {
    if (nodes->layers[q] < 0) return ERR_SHAPE;
    k[q] = (size_t)nodes->layers[q];
  }
  return vol.enable_graded(k.data(), k.size());
}
template <class T>
int project(const Geometry* g, const Options* o, const Coupling* cpl, const Nodes* nodes, const float* us,
            const float* vs, const float* ws, const float* p_warm, float scale, float k1, uint32_t max_iters,
            int refine, uint32_t refine_iters, float* p_out,
            float* u_out, float* v_out, float* w_out, double* p_out64, Report* rep) {
  Volume<T> vol;
  Status st = setup(vol, g, o);
  if (st) return st;
  if (cpl && cpl->total) couple(vol, cpl);
  if ((st = grade(vol, nodes))) return st;
  for (size_t q = 0; q < vol.nu(); ++q) vol.us[q] = (T)us[q];
  for (size_t q = 0; q < vol.nv(); ++q) vol.vs[q] = (T)vs[q];
  for (size_t q = 0; q < vol.nw(); ++q) vol.ws[q] = (T)ws[q];
  if (p_warm) for (size_t q = 0; q < vol.cells(); ++q) vol.p[q] = (T)p_warm[q];
  Report r{};
  st = vol.project_mobile3((T)scale, (T)k1, max_iters, r);
  if (st) return st;
  if (refine == 2 || (refine == 1 && r.degraded && r.floor)) {
    const uint32_t first = r.iterations;
    st = vol.refine_mobile3((T)scale, (T)k1, refine_iters ? refine_iters : max_iters, r);
    if (st) return st;
    if (refine == 1) r.iterations = first + r.iterations < first ? 0xffffffffu : first + r.iterations;
  }
  if (rep) *rep = r;
  if (p_out) for (size_t q = 0; q < vol.cells(); ++q) p_out[q] = (float)vol.p[q];
  if (p_out64) for (size_t q = 0; q < vol.cells(); ++q) p_out64[q] = (double)vol.p[q];
  if (u_out) for (size_t q = 0; q < vol.nu(); ++q) u_out[q] = (float)vol.u[q];
  if (v_out) for (size_t q = 0; q < vol.nv(); ++q) v_out[q] = (float)vol.v[q];
  if (w_out) for (size_t q = 0; q < vol.nw(); ++q) w_out[q] = (float)vol.w[q];
  return OK;
}

template <class T>
int vcycle(const Geometry* g, const Options* o, const float* r, double* z) {
  Volume<T> vol;
  Options opt = o ? *o : default_options();
  opt.multigrid = 1;
  Status st = vol.configure(*g, opt);
  if (st) return st;
  std::fill(vol.rhs.begin(), vol.rhs.end(), T(0));
  vol.rhs_mobile3(T(1));
  vol.prepare_multigrid3();
  for (size_t q = 0; q < vol.cells(); ++q) vol.res[q] = (T)r[q];
  vol.v_cycle3();
  for (size_t q = 0; q < vol.cells(); ++q) z[q] = (double)vol.mz[q];
  return OK;
}

template <class T>
int step_surface(const Geometry* g, const Options* o, float* u, float* v, float* w, float* p, float* eta,
                 float* eta_roundoff, uint32_t steps, uint64_t duration_us, uint32_t max_iters, Report* reps,
                 double* balances, uint32_t* done) {
  Geometry g2 = *g;
  g2.eta = eta;
  g2.eta_roundoff = eta_roundoff;
  Volume<T> vol;
  Status st = setup(vol, &g2, o);
  if (st) return st;
  for (size_t q = 0; q < vol.nu(); ++q) vol.u[q] = (T)u[q];
  for (size_t q = 0; q < vol.nv(); ++q) vol.v[q] = (T)v[q];
  for (size_t q = 0; q < vol.nw(); ++q) vol.w[q] = (T)w[q];
  for (size_t q = 0; q < vol.cells(); ++q) vol.p[q] = (T)p[q];
  uint32_t n = 0;
  for (; n < steps; ++n) {
    Report r{};
    typename Volume<T>::Balance b{};
    st = vol.step_surface_mobile(duration_us, max_iters, r, b);
    if (reps) reps[n] = r;
    if (st) break;
    if (balances) {
      const double bb[7] = {b.volume, b.delta, b.band_in, b.perturbation_in, b.sponge_out, b.residual, b.outgoing};
      std::copy(bb, bb + 7, balances + 7 * (size_t)n);
    }
  }
  if (done) *done = n;
  for (size_t q = 0; q < vol.nu(); ++q) u[q] = (float)vol.u[q];
  for (size_t q = 0; q < vol.nv(); ++q) v[q] = (float)vol.v[q];
  for (size_t q = 0; q < vol.nw(); ++q) w[q] = (float)vol.w[q];
  for (size_t q = 0; q < vol.cells(); ++q) p[q] = (float)vol.p[q];
  for (size_t q = 0; q < vol.nx * vol.ny; ++q) {
    eta[q] = (float)vol.eta[q];
    eta_roundoff[q] = (float)vol.eta_roundoff[q];
  }
  return st;
}

template <class T>
int backward_error(const Geometry* g, const double* p, const double* rhs, const double* res, double* out) {
  Volume<T> vol;
  Options opt = default_options();
  opt.multigrid = 0;
  Status st = vol.configure(*g, opt);
  if (st) return st;
  for (size_t q = 0; q < vol.cells(); ++q) {
    vol.p[q] = (T)p[q];
    vol.rhs[q] = (T)rhs[q];
    vol.res[q] = (T)res[q];
  }
  *out = (double)vol.backward_error_mobile3();
  return OK;
}

template <class T>
int coarse_apply(const Geometry* g, const Options* o, int level, const double* x, double* out) {
  Volume<T> vol;
  Options opt = o ? *o : default_options();
  opt.multigrid = 1;
  Status st = vol.configure(*g, opt);
  if (st) return st;
  if (level < 0 || (size_t)level >= vol.levels.size()) return ERR_DOMAIN;
  vol.prepare_multigrid3();
  const Level3<T>& L = vol.levels[level];
  const size_t n = L.nx * L.ny * L.nz;
  std::vector<T> xx(n), oo(n);
  for (size_t q = 0; q < n; ++q) xx[q] = (T)x[q];
  Volume<T>::row_sums(L, &xx, oo);
  for (size_t q = 0; q < n; ++q) out[q] = (double)oo[q];
  return OK;
}

Options no_multigrid() {
  Options opt = default_options();
  opt.multigrid = 0;
  return opt;
}

}  // namespace

extern "C" {

// The APIC-3D pressure projection (apic3d.rs:940-1051) in f32 (precision 32)
// or f64 (64): u, v, w in (predicted, walls and body imposed) and out
// (corrected); p (cells, f64) out; iterations, relative residual and the
// divergence metric (apic3d.rs:1053-1080) out.
int orc_apic_project(int precision, int nx, int ny, int nz, float dx, float rho, float dt, const uint8_t* label,
                     const float* phi, float* u, float* v, float* w, double* p_out, uint32_t* iterations,
                     double* residual, double* divergence) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0.f) || !label || !phi) return ERR_DOMAIN;
  auto run = [&](auto proj) {
    proj.nx = nx; proj.ny = ny; proj.nz = nz; proj.dx = dx; proj.label = label; proj.phi = phi;
    const apic::Result r = proj.project(rho, dt, u, v, w);
    for (size_t c = 0; c < proj.p.size(); ++c) p_out[c] = (double)proj.p[c];
    *iterations = r.iterations;
    *residual = r.residual;
    *divergence = r.divergence;
  };
  if (precision == 64) run(apic::Projection<double>{});
  else run(apic::Projection<float>{});
  return OK;
}

// y = A x of the APIC operator (apic3d.rs:912-935), f64.
int orc_apic_apply(int nx, int ny, int nz, const uint8_t* label, const float* phi, const double* x, double* y) {
  apic::Projection<double> proj{};
  proj.nx = nx; proj.ny = ny; proj.nz = nz; proj.dx = 1.0; proj.label = label; proj.phi = phi;
  const size_t n = (size_t)nx * ny * nz;
  std::vector<double> xv(x, x + n), yv(n, 0.0);
  proj.apply(xv, yv);
  std::copy(yv.begin(), yv.end(), y);
  return OK;
}

// The APIC divergence metric (apic3d.rs:1053-1080) of given faces.
double orc_apic_divergence_metric(int nx, int ny, int nz, const uint8_t* label, const float* u, const float* v,
                                  const float* w) {
  apic::Projection<float> proj{};
  proj.nx = nx; proj.ny = ny; proj.nz = nz; proj.dx = 1.f; proj.label = label; proj.phi = nullptr;
  return proj.divergence_metric(u, v, w);
}

// The cut of a bottom height field (delta3d_cut.rs:172-258): frac (cells),
// open_u, open_v, open_w (faces) and floor (columns) from the nx*ny bottom
// heights at the column centres.
int orc_cut(int nx, int ny, int nz, float dx, const float* bottom, float* frac, float* open_u, float* open_v,
            float* open_w, float* floor) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0.f) || !bottom) return ERR_DOMAIN;
  const cut::Cut c = cut::cut(nx, ny, nz, dx, bottom);
  std::copy(c.frac.begin(), c.frac.end(), frac);
  std::copy(c.open_u.begin(), c.open_u.end(), open_u);
  std::copy(c.open_v.begin(), c.open_v.end(), open_v);
  std::copy(c.open_w.begin(), c.open_w.end(), open_w);
  std::copy(c.floor.begin(), c.floor.end(), floor);
  return OK;
}

// Surface steps (delta3d_mobile.rs:819-927), `steps` of them in a row; u, v,
// w, p, eta, eta_roundoff in and out (g's eta pointers are ignored). Returns
// the status of the last step taken; *done = steps completed; balances: 7
// doubles per step (Balance3 order: volume, delta, band_in,
// perturbation_in, sponge_out, residual, outgoing).
int orc_step_surface(int precision, const Geometry* g, const Options* o, float* u, float* v, float* w, float* p,
                     float* eta, float* eta_roundoff, uint32_t steps, uint64_t duration_us, uint32_t max_iters,
                     Report* reps, double* balances, uint32_t* done) {
  if (precision == 64)
    return step_surface<double>(g, o, u, v, w, p, eta, eta_roundoff, steps, duration_us, max_iters, reps, balances,
                                done);
  return step_surface<float>(g, o, u, v, w, p, eta, eta_roundoff, steps, duration_us, max_iters, reps, balances, done);
}

// Advection alone (delta3d_mobile.rs:615-732), f32.
int orc_advect(const Geometry* g, const float* u, const float* v, const float* w, float dt, float* us, float* vs,
               float* ws) {
  Volume<float> vol;
  const Options opt = no_multigrid();
  Status st = setup(vol, g, &opt);
  if (st) return st;
  std::copy(u, u + vol.nu(), vol.u.begin());
  std::copy(v, v + vol.nv(), vol.v.begin());
  std::copy(w, w + vol.nw(), vol.w.begin());
  vol.advect_mobile3(dt);
  std::copy(vol.us.begin(), vol.us.end(), us);
  std::copy(vol.vs.begin(), vol.vs.end(), vs);
  std::copy(vol.ws.begin(), vol.ws.end(), ws);
  return OK;
}

// Extrapolation alone (delta3d_mobile.rs:345-394), in place, f32.
int orc_extrapolate(const Geometry* g, float* u, float* v, float* w) {
  Volume<float> vol;
  const Options opt = no_multigrid();
  Status st = setup(vol, g, &opt);
  if (st) return st;
  std::copy(u, u + vol.nu(), vol.u.begin());
  std::copy(v, v + vol.nv(), vol.v.begin());
  std::copy(w, w + vol.nw(), vol.w.begin());
  vol.extrapolate_mobile3();
  std::copy(vol.u.begin(), vol.u.end(), u);
  std::copy(vol.v.begin(), vol.v.end(), v);
  std::copy(vol.w.begin(), vol.w.end(), w);
  return OK;
}

// Height transport alone (delta3d_mobile.rs:734-782), f32: new eta and
// eta_roundoff, the discharges ((nx+1)*ny and nx*(ny+1)).
int orc_transport(const Geometry* g, const float* u, const float* v, float transport, float* eta_out,
                  float* eta_roundoff_out, float* flux_x, float* flux_y) {
  Volume<float> vol;
  const Options opt = no_multigrid();
  Status st = setup(vol, g, &opt);
  if (st) return st;
  std::copy(u, u + vol.nu(), vol.u.begin());
  std::copy(v, v + vol.nv(), vol.v.begin());
  vol.transport_mobile3(transport);
  std::copy(vol.eta.begin(), vol.eta.end(), eta_out);
  std::copy(vol.eta_roundoff.begin(), vol.eta_roundoff.end(), eta_roundoff_out);
  if (flux_x) std::copy(vol.flux_x.begin(), vol.flux_x.end(), flux_x);
  if (flux_y) std::copy(vol.flux_y.begin(), vol.flux_y.end(), flux_y);
  return OK;
}

// delta3d_mobile.rs:801-813: 1 when the surface is within bounds.
int orc_mobile_in_bounds(const Geometry* g) {
  Volume<float> vol;
  const Options opt = no_multigrid();
  if (setup(vol, g, &opt)) return -1;
  return vol.mobile_in_bounds() ? 1 : 0;
}

// delta3d_balance.rs:105-113
double orc_perturbation_volume(const Geometry* g) {
  Volume<float> vol;
  const Options opt = no_multigrid();
  if (setup(vol, g, &opt)) return std::nan("");
  return vol.perturbation_volume();
}

int orc_project(int precision, const Geometry* g, const Options* o, const float* us, const float* vs, const float* ws,
                const float* p_warm, float scale, float k1, uint32_t max_iters, float* p_out, float* u_out,
                float* v_out, float* w_out, double* p_out64, Report* rep) {
  if (precision == 64)
    return project<double>(g, o, nullptr, nullptr, us, vs, ws, p_warm, scale, k1, max_iters, 0, 0, p_out, u_out,
                           v_out, w_out, p_out64, rep);
  return project<float>(g, o, nullptr, nullptr, us, vs, ws, p_warm, scale, k1, max_iters, 0, 0, p_out, u_out, v_out,
                        w_out, p_out64, rep);
}

// The projection path of step_perturbation_mobile (delta3d_coupling.rs:
// 720-864): the projection over a coupled surface, with refine_mobile3 as
// `refine` says.
int orc_project_coupled(int precision, const Geometry* g, const Options* o, const Coupling* cpl, const float* us,
                        const float* vs, const float* ws, const float* p_warm, float scale, float k1,
                        uint32_t max_iters, int refine, uint32_t refine_iters, float* p_out, float* u_out,
                        float* v_out, float* w_out, double* p_out64, Report* rep) {
  if (precision == 64)
    return project<double>(g, o, cpl, nullptr, us, vs, ws, p_warm, scale, k1, max_iters, refine, refine_iters, p_out,
                           u_out, v_out, w_out, p_out64, rep);
  return project<float>(g, o, cpl, nullptr, us, vs, ws, p_warm, scale, k1, max_iters, refine, refine_iters, p_out,
                        u_out, v_out, w_out, p_out64, rep);
}

// project_mobile3 over the graded column (delta3d_graded.rs:310-447), with an
// optional coupled surface.
int orc_project_graded(int precision, const Geometry* g, const Options* o, const Coupling* cpl, const Nodes* nodes,
                       const float* us, const float* vs, const float* ws, const float* p_warm, float scale, float k1,
                       uint32_t max_iters, float* p_out, float* u_out, float* v_out, float* w_out, double* p_out64,
                       Report* rep) {
  if (!nodes || !nodes->layers) return ERR_SHAPE;
  if (precision == 64)
    return project<double>(g, o, cpl, nodes, us, vs, ws, p_warm, scale, k1, max_iters, 0, 0, p_out, u_out, v_out,
                           w_out, p_out64, rep);
  return project<float>(g, o, cpl, nodes, us, vs, ws, p_warm, scale, k1, max_iters, 0, 0, p_out, u_out, v_out, w_out,
                        p_out64, rep);
}

// The graded column's operators (delta3d_graded.rs:109-162) in precision T:
// op 0 fine = P x, 1 x = P^T fine, 2 x = P^T A P x, 3 the nodes' weights,
// 4 the graded right-hand side P^T b of the predicted faces `in` (u, v, w
// one after the other) and `scale`.
int orc_graded_op(int precision, const Geometry* g, const Nodes* nodes, int op, const double* in, double* out,
                  double scale) {
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    Volume<T> vol;
    Options opt = default_options();
    opt.multigrid = 0;
    Status st = vol.configure(*g, opt);
    if (st) return st;
    if (!nodes || !nodes->layers) return ERR_SHAPE;
    if ((st = grade(vol, nodes))) return st;
    const size_t n = vol.graded_len(), cells = vol.cells();
    if (op == 0) {
      std::vector<T> x(in, in + n), fine(cells);
      vol.graded_expand(x, fine);
      std::copy(fine.begin(), fine.end(), out);
    } else if (op == 1) {
      std::vector<T> fine(in, in + cells), x(n);
      vol.graded_restrict(fine, x);
      std::copy(x.begin(), x.end(), out);
    } else if (op == 2) {
      std::vector<T> x(in, in + n), y(n);
      vol.graded_apply(x, y);
      std::copy(y.begin(), y.end(), out);
    } else if (op == 3) {
      std::copy(vol.gweight.begin(), vol.gweight.end(), out);
    } else if (op == 4) {
      const size_t nu = vol.nu(), nv = vol.nv(), nw = vol.nw();
      vol.us.assign(in, in + nu);
      vol.vs.assign(in + nu, in + nu + nv);
      vol.ws.assign(in + nu + nv, in + nu + nv + nw);
      vol.divergence(vol.us, vol.vs, vol.ws, vol.rhs);
      vol.rhs_mobile3((T)scale);
      std::vector<T> x(n);
      vol.graded_restrict(vol.rhs, x);
      std::copy(x.begin(), x.end(), out);
    } else {
      return ERR_DOMAIN;
    }
    return OK;
  };
  return precision == 64 ? run(double()) : run(float());
}

// Operator action (delta3d_mobile.rs:936-949).
int orc_apply(const Geometry* g, const float* x, float* out) {
  Volume<float> vol;
  Options o = default_options();
  o.multigrid = 0;
  Status st = vol.configure(*g, o);
  if (st) return st;
  std::vector<float> xx(x, x + vol.cells()), oo(vol.cells());
  vol.apply_mobile3(xx, oo);
  std::copy(oo.begin(), oo.end(), out);
  return OK;
}

// Operator action in f64 on f64 input.
int orc_apply64(const Geometry* g, const double* x, double* out) {
  Volume<double> vol;
  Options o = default_options();
  o.multigrid = 0;
  Status st = vol.configure(*g, o);
  if (st) return st;
  std::vector<double> xx(x, x + vol.cells()), oo(vol.cells());
  vol.apply_mobile3(xx, oo);
  std::copy(oo.begin(), oo.end(), out);
  return OK;
}

// Assembly (delta3d_mobile.rs:960-979): rhs and Jacobi prec from a divergence.
int orc_assemble(const Geometry* g, const float* div, float scale, float* rhs, float* prec) {
  Volume<float> vol;
  Options o = default_options();
  o.multigrid = 0;
  Status st = vol.configure(*g, o);
  if (st) return st;
  std::copy(div, div + vol.cells(), vol.rhs.begin());
  vol.rhs_mobile3(scale);
  std::copy(vol.rhs.begin(), vol.rhs.end(), rhs);
  std::copy(vol.prec.begin(), vol.prec.end(), prec);
  return OK;
}

// Divergence of faces (delta3d.rs:747-763, :806-823).
int orc_divergence(const Geometry* g, const float* u, const float* v, const float* w, float* out) {
  Volume<float> vol;
  Options o = default_options();
  o.multigrid = 0;
  Status st = vol.configure(*g, o);
  if (st) return st;
  std::vector<float> uu(u, u + vol.nu()), vv(v, v + vol.nv()), ww(w, w + vol.nw()), oo(vol.cells());
  vol.divergence(uu, vv, ww, oo);
  std::copy(oo.begin(), oo.end(), out);
  return OK;
}

// Oettli-Prager backward error (delta3d_mobile.rs:243-271) of a given
// pressure, right-hand side and residual, in f32 (32) or f64 (64).
int orc_backward_error(int precision, const Geometry* g, const double* p, const double* rhs, const double* res,
                       double* out) {
  if (precision == 64) return backward_error<double>(g, p, rhs, res, out);
  return backward_error<float>(g, p, rhs, res, out);
}

// row_sums (delta3d_multigrid.rs:216-255) of coarse level `level` applied to x,
// after the hierarchy is prepared from the geometry (options as given).
int orc_coarse_apply(int precision, const Geometry* g, const Options* o, int level, const double* x, double* out) {
  if (precision == 64) return coarse_apply<double>(g, o, level, x, out);
  return coarse_apply<float>(g, o, level, x, out);
}

// The acceptance metric (delta3d_mobile.rs:398-433) of given faces: the
// maximum |div| over wet rows and over plain rows, normalised by the largest
// face speed (f32, the reference's arithmetic).
int orc_divergence_metric(const Geometry* g, const float* u, const float* v, const float* w, double* div_all,
                          double* div_plain) {
  Volume<float> vol;
  Options o = default_options();
  o.multigrid = 0;
  Status st = vol.configure(*g, o);
  if (st) return st;
  std::copy(u, u + vol.nu(), vol.u.begin());
  std::copy(v, v + vol.nv(), vol.v.begin());
  std::copy(w, w + vol.nw(), vol.w.begin());
  vol.divergence_mobile3(*div_all, *div_plain);
  return OK;
}

// Gradient subtract (delta3d_mobile.rs:284-343) followed by the acceptance
// metric (delta3d_mobile.rs:398-433).
int orc_correct_coupled(const Geometry* g, const Coupling* cpl, const float* us, const float* vs, const float* ws,
                        const float* p, float k1, float* u, float* v, float* w, double* div_all, double* div_plain);

int orc_correct(const Geometry* g, const float* us, const float* vs, const float* ws, const float* p, float k1,
                float* u, float* v, float* w, double* div_all, double* div_plain) {
  return orc_correct_coupled(g, nullptr, us, vs, ws, p, k1, u, v, w, div_all, div_plain);
}

// The same over a coupled surface (cpl null: uncoupled).
int orc_correct_coupled(const Geometry* g, const Coupling* cpl, const float* us, const float* vs, const float* ws,
                        const float* p, float k1, float* u, float* v, float* w, double* div_all, double* div_plain) {
  Volume<float> vol;
  Options o = default_options();
  o.multigrid = 0;
  Status st = vol.configure(*g, o);
  if (st) return st;
  if (cpl && cpl->total) couple(vol, cpl);
  std::copy(us, us + vol.nu(), vol.us.begin());
  std::copy(vs, vs + vol.nv(), vol.vs.begin());
  std::copy(ws, ws + vol.nw(), vol.ws.begin());
  std::copy(p, p + vol.cells(), vol.p.begin());
  vol.correct_mobile3(k1);
  double a, b;
  vol.divergence_mobile3(a, b);
  std::copy(vol.u.begin(), vol.u.end(), u);
  std::copy(vol.v.begin(), vol.v.end(), v);
  std::copy(vol.w.begin(), vol.w.end(), w);
  if (div_all) *div_all = a;
  if (div_plain) *div_plain = b;
  return OK;
}

// One V-cycle z = M^-1 r on the geometry (delta3d_multigrid.rs:393-429).
int orc_vcycle(int precision, const Geometry* g, const Options* o, const float* r, double* z) {
  if (precision == 64) return vcycle<double>(g, o, r, z);
  return vcycle<float>(g, o, r, z);
}

// Per-cell classification: 1 wet, 0 solid, -1 air (delta3d_multigrid.rs:357-367),
// and the ghost-row flag used by the plain-row metric.
int orc_classify(const Geometry* g, int8_t* kind, uint8_t* ghost_row) {
  Volume<float> vol;
  Options o = default_options();
  o.multigrid = 0;
  Status st = vol.configure(*g, o);
  if (st) return st;
  for (size_t k = 0; k < vol.nz; ++k)
    for (size_t j = 0; j < vol.ny; ++j)
      for (size_t i = 0; i < vol.nx; ++i) {
        size_t cc = vol.c(i, j, k);
        kind[cc] = (int8_t)vol.fine_kind3(cc);
        uint8_t gr = 0;
        if (vol.wet3(i, j, k)) {
          auto row = vol.mobile_row(i, j, k);
          for (int f = 0; f < 6; ++f) if (row.n[f] < 0 && row.a[f] > 0.f) gr = 1;
        }
        if (ghost_row) ghost_row[cc] = gr;
      }
  return OK;
}

// Coarse level geometry after prepare_multigrid3 (delta3d_multigrid.rs:370-387):
// kind and diag of level `level` (0 = first coarse level).
int orc_coarse_level(const Geometry* g, const Options* o, int level, float* kind, float* diag) {
  Volume<float> vol;
  Options opt = o ? *o : default_options();
  opt.multigrid = 1;
  Status st = vol.configure(*g, opt);
  if (st) return st;
  vol.prepare_multigrid3();
  if (level < 0 || (size_t)level >= vol.levels.size()) return ERR_DOMAIN;
  auto& L = vol.levels[level];
  std::copy(L.kind.begin(), L.kind.end(), kind);
  std::copy(L.diag.begin(), L.diag.end(), diag);
  return OK;
}

size_t orc_level_count3(size_t nx, size_t ny, size_t nz) { return level_count3(nx, ny, nz); }
size_t orc_hierarchy_floats3(size_t nx, size_t ny, size_t nz) { return hierarchy_floats3(nx, ny, nz); }
uint64_t orc_fingerprint(const float* v, size_t n) { return fingerprint_f32(v, n); }

// FNV-1a 64 over raw bytes, with the constants fingerprint_f32 uses.
uint64_t orc_fnv_bytes(const uint8_t* b, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= (uint64_t)b[i];
    h *= 0x100000001b3ull;
  }
  return h;
}
float orc_roundoff_3d() { return roundoff_3d<float>(); }
// Storage rounding of the V-cycle vectors (Options::storage), in place.
void orc_round_storage(int storage, float* x, size_t n) {
  for (size_t i = 0; i < n; ++i) x[i] = round_storage(storage, x[i]);
}

}  // extern "C"
