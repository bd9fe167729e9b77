// CPU ORACLE — test infrastructure only (see oracle.hpp). C entry points for
// ctypes: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
#include "oracle.hpp"

#include <algorithm>

using namespace orc;

namespace {

template <class T>
Status setup(Volume<T>& vol, const Geometry* g, const Options* o) {
  Options opt = o ? *o : default_options();
  return vol.configure(*g, opt);
}

template <class T>
int project(const Geometry* g, const Options* o, const float* us, const float* vs, const float* ws, const float* p_warm,
            float scale, float k1, uint32_t max_iters, float* p_out, float* u_out, float* v_out, float* w_out,
            double* p_out64, Report* rep) {
  Volume<T> vol;
  Status st = setup(vol, g, o);
  if (st) return st;
  for (size_t q = 0; q < vol.nu(); ++q) vol.us[q] = (T)us[q];
  for (size_t q = 0; q < vol.nv(); ++q) vol.vs[q] = (T)vs[q];
  for (size_t q = 0; q < vol.nw(); ++q) vol.ws[q] = (T)ws[q];
  if (p_warm) for (size_t q = 0; q < vol.cells(); ++q) vol.p[q] = (T)p_warm[q];
  Report r{};
  st = vol.project_mobile3((T)scale, (T)k1, max_iters, r);
  if (st) return st;
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

}  // namespace

extern "C" {

int orc_project(int precision, const Geometry* g, const Options* o, const float* us, const float* vs, const float* ws,
                const float* p_warm, float scale, float k1, uint32_t max_iters, float* p_out, float* u_out,
                float* v_out, float* w_out, double* p_out64, Report* rep) {
  if (precision == 64)
    return project<double>(g, o, us, vs, ws, p_warm, scale, k1, max_iters, p_out, u_out, v_out, w_out, p_out64, rep);
  return project<float>(g, o, us, vs, ws, p_warm, scale, k1, max_iters, p_out, u_out, v_out, w_out, p_out64, rep);
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

// Gradient subtract (delta3d_mobile.rs:284-343) followed by the acceptance
// metric (delta3d_mobile.rs:398-433).
int orc_correct(const Geometry* g, const float* us, const float* vs, const float* ws, const float* p, float k1,
                float* u, float* v, float* w, double* div_all, double* div_plain) {
  Volume<float> vol;
  Options o = default_options();
  o.multigrid = 0;
  Status st = vol.configure(*g, o);
  if (st) return st;
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
