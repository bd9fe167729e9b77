// Seeded synthetic inputs for the projection workloads (host code).
//
// RNG: the reference's test generator `noise` (code/water-core/src/
// tests_delta3d.rs:66-74), a xorshift64 seeded with seed*0x9e3779b97f4a7c15|1
// whose draw (s>>40)/2^24*2-1 lies in [-1, 1). Each purpose draws from its own
// stream, seeded with seed*1000 + purpose.
//
// cfg 0 — closed box: every cell fluid, faces i.i.d. U[-1,1] from the noise
//         streams u (1), v (2), w (3); every boundary-normal face zeroed.
// cfg 1 — free-surface tank (sizes in reference orientation, nz vertical,
//         cell units dx = 1, s = nz/128 scales amplitudes and radii):
//   surface h(i,j) = nz/2 + sum_{m<4} a_m sin(f_m x + g_m y + phi_m), x = i+1/2,
//     y = j+1/2, a ~ U[2,6]*s, f,g ~ U[0.02,0.15], phi ~ U[0,2pi) (stream 11);
//   8 spheres: centre x ~ U[0,nx), y ~ U[0,ny), z ~ U[0, h(x,y)),
//     radius ~ U[6,14]*s (stream 12); a cell is solid iff its centre is inside;
//   velocity: 16 modes (stream 13), wavevector components uniform in
//     {-4..-1, 1..4}, amplitude 1/|k|^2, unit polarisation uniform on the
//     sphere, phase ~ U[0,2pi); face components evaluated at face centres
//     with phase 2pi(kx X/nx + ky Y/ny + kz Z/nz); faces on the box walls or
//     touching a solid are 0; the field is scaled to RMS 1 over the others.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/fsp.h"

namespace {

struct Noise {
  uint64_t s;
  explicit Noise(uint64_t seed) : s(seed * 0x9e3779b97f4a7c15ull | 1ull) {}
  float next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return ((float)(s >> 40) / 16777216.0f) * 2.0f - 1.0f;
  }
  // U[a, b) from one draw.
  double uni(double a, double b) { return a + (b - a) * ((double)next() + 1.0) * 0.5; }
};

uint64_t stream_seed(uint64_t seed, uint64_t purpose) { return seed * 1000ull + purpose; }

template <class F>
void parallel_slices(int n, F f) {
  unsigned hw = std::thread::hardware_concurrency();
  int threads = (int)(hw == 0 ? 1 : (hw > 16 ? 16 : hw));
  if (threads > n) threads = n;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      for (int k = t; k < n; k += threads) f(k);
    });
  for (auto& th : pool) th.join();
}

struct Mode {
  double kx, ky, kz, amp, px, py, pz, phase;
};

}  // namespace

extern "C" fsp_status fsp_generate(int32_t cfg, int32_t nx, int32_t ny, int32_t nz, uint64_t seed, float* eta,
                                   uint8_t* solid, float* us, float* vs, float* ws) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return FSP_DOMAIN;
  const size_t NX = nx, NY = ny, NZ = nz;
  const size_t cells = NX * NY * NZ;
  const size_t nu = (NX + 1) * NY * NZ, nv = NX * (NY + 1) * NZ, nw = NX * NY * (NZ + 1);
  auto fu = [&](size_t i, size_t j, size_t k) { return (k * NY + j) * (NX + 1) + i; };
  auto fv = [&](size_t i, size_t j, size_t k) { return (k * (NY + 1) + j) * NX + i; };
  auto fw = [&](size_t i, size_t j, size_t k) { return (k * NY + j) * NX + i; };

  if (cfg == 0) {
    if (eta) for (size_t q = 0; q < NX * NY; ++q) eta[q] = (float)nz;
    if (solid) std::memset(solid, 0, cells);
    float* arrays[3] = {us, vs, ws};
    size_t lens[3] = {nu, nv, nw};
    for (int a = 0; a < 3; ++a) {
      if (!arrays[a]) continue;
      Noise n(stream_seed(seed, 1 + a));
      for (size_t q = 0; q < lens[a]; ++q) arrays[a][q] = n.next();
    }
    if (us) for (size_t k = 0; k < NZ; ++k) for (size_t j = 0; j < NY; ++j) { us[fu(0, j, k)] = 0.f; us[fu(NX, j, k)] = 0.f; }
    if (vs) for (size_t k = 0; k < NZ; ++k) for (size_t i = 0; i < NX; ++i) { vs[fv(i, 0, k)] = 0.f; vs[fv(i, NY, k)] = 0.f; }
    if (ws) for (size_t j = 0; j < NY; ++j) for (size_t i = 0; i < NX; ++i) { ws[fw(i, j, 0)] = 0.f; ws[fw(i, j, NZ)] = 0.f; }
    return FSP_OK;
  }
  if (cfg != 1) return FSP_INVALID;

  const double s = (double)nz / 128.0;
  const double two_pi = 6.283185307179586;
  // surface
  double a[4], f[4], g[4], ph[4];
  {
    Noise n(stream_seed(seed, 11));
    for (int m = 0; m < 4; ++m) {
      a[m] = n.uni(2.0, 6.0) * s;
      f[m] = n.uni(0.02, 0.15);
      g[m] = n.uni(0.02, 0.15);
      ph[m] = n.uni(0.0, two_pi);
    }
  }
  auto height = [&](double x, double y) {
    double h = 0.5 * (double)nz;
    for (int m = 0; m < 4; ++m) h += a[m] * std::sin(f[m] * x + g[m] * y + ph[m]);
    return h;
  };
  std::vector<float> h_col(NX * NY);
  for (size_t j = 0; j < NY; ++j)
    for (size_t i = 0; i < NX; ++i) h_col[j * NX + i] = (float)height((double)i + 0.5, (double)j + 0.5);
  if (eta) std::memcpy(eta, h_col.data(), NX * NY * sizeof(float));

  // spheres
  double sc[8][4];
  {
    Noise n(stream_seed(seed, 12));
    for (int q = 0; q < 8; ++q) {
      double cx = n.uni(0.0, (double)nx), cy = n.uni(0.0, (double)ny);
      double cz = n.uni(0.0, height(cx, cy));
      double r = n.uni(6.0, 14.0) * s;
      sc[q][0] = cx; sc[q][1] = cy; sc[q][2] = cz; sc[q][3] = r;
    }
  }
  std::vector<uint8_t> sol(cells, 0);
  parallel_slices(nz, [&](int k) {
    for (size_t j = 0; j < NY; ++j)
      for (size_t i = 0; i < NX; ++i) {
        double x = i + 0.5, y = j + 0.5, z = k + 0.5;
        uint8_t in = 0;
        for (int q = 0; q < 8; ++q) {
          double dx = x - sc[q][0], dy = y - sc[q][1], dz = z - sc[q][2];
          if (dx * dx + dy * dy + dz * dz < sc[q][3] * sc[q][3]) in = 1;
        }
        sol[((size_t)k * NY + j) * NX + i] = in;
      }
  });
  if (solid) std::memcpy(solid, sol.data(), cells);

  // velocity modes
  Mode modes[16];
  {
    Noise n(stream_seed(seed, 13));
    auto comp = [&]() {
      int v = (int)std::floor(n.uni(0.0, 8.0));  // 0..7
      if (v > 7) v = 7;
      return (double)(v < 4 ? v - 4 : v - 3);    // -4..-1, 1..4
    };
    for (int m = 0; m < 16; ++m) {
      Mode& M = modes[m];
      M.kx = comp(); M.ky = comp(); M.kz = comp();
      M.amp = 1.0 / (M.kx * M.kx + M.ky * M.ky + M.kz * M.kz);
      double z = n.uni(-1.0, 1.0), t = n.uni(0.0, two_pi);
      double rxy = std::sqrt(std::fmax(0.0, 1.0 - z * z));
      M.px = rxy * std::cos(t); M.py = rxy * std::sin(t); M.pz = z;
      M.phase = n.uni(0.0, two_pi);
    }
  }
  auto S = [&](size_t i, size_t j, size_t k) { return sol[(k * NY + j) * NX + i] != 0; };
  // Phase tables: e^{i 2pi kx X / nx} at integer and half-integer X.
  auto table = [&](int axis, size_t n_pts, bool half) {
    std::vector<double> re(16 * n_pts), im(16 * n_pts);
    size_t len = axis == 0 ? NX : axis == 1 ? NY : NZ;
    for (int m = 0; m < 16; ++m) {
      double kk = axis == 0 ? modes[m].kx : axis == 1 ? modes[m].ky : modes[m].kz;
      for (size_t q = 0; q < n_pts; ++q) {
        double X = (double)q + (half ? 0.5 : 0.0);
        double th = two_pi * kk * X / (double)len;
        re[m * n_pts + q] = std::cos(th);
        im[m * n_pts + q] = std::sin(th);
      }
    }
    return std::make_pair(re, im);
  };
  auto xi = table(0, NX + 1, false), xh = table(0, NX, true);
  auto yi = table(1, NY + 1, false), yh = table(1, NY, true);
  auto zi = table(2, NZ + 1, false), zh = table(2, NZ, true);
  double cph[16], sph[16];
  for (int m = 0; m < 16; ++m) { cph[m] = std::cos(modes[m].phase); sph[m] = std::sin(modes[m].phase); }
  // value of component `comp` at a point given the three axis table entries
  auto eval = [&](int comp, const std::pair<std::vector<double>, std::vector<double>>& X, size_t nxp, size_t ix,
                  const std::pair<std::vector<double>, std::vector<double>>& Y, size_t nyp, size_t iy,
                  const std::pair<std::vector<double>, std::vector<double>>& Z, size_t nzp, size_t iz) {
    double acc = 0.0;
    for (int m = 0; m < 16; ++m) {
      double ar = X.first[m * nxp + ix], ai = X.second[m * nxp + ix];
      double br = Y.first[m * nyp + iy], bi = Y.second[m * nyp + iy];
      double cr = Z.first[m * nzp + iz], ci = Z.second[m * nzp + iz];
      double r1 = ar * br - ai * bi, i1 = ar * bi + ai * br;
      double r2 = r1 * cr - i1 * ci, i2 = r1 * ci + i1 * cr;
      double re = r2 * cph[m] - i2 * sph[m];  // cos(theta + phase)
      double pol = comp == 0 ? modes[m].px : comp == 1 ? modes[m].py : modes[m].pz;
      acc += modes[m].amp * pol * re;
    }
    return acc;
  };
  std::vector<float> U(nu, 0.f), V(nv, 0.f), W(nw, 0.f);
  std::vector<double> sumsq(NZ + 1, 0.0);
  std::vector<size_t> count(NZ + 1, 0);
  parallel_slices(nz + 1, [&](int kk) {
    size_t k = (size_t)kk;
    double ss = 0.0;
    size_t cnt = 0;
    if (k < NZ) {
      for (size_t j = 0; j < NY; ++j)
        for (size_t i = 0; i <= NX; ++i) {
          if (i == 0 || i == NX || S(i - 1, j, k) || S(i, j, k)) continue;
          double val = eval(0, xi, NX + 1, i, yh, NY, j, zh, NZ, k);
          U[fu(i, j, k)] = (float)val;
          ss += val * val; ++cnt;
        }
      for (size_t j = 0; j <= NY; ++j)
        for (size_t i = 0; i < NX; ++i) {
          if (j == 0 || j == NY || S(i, j - 1, k) || S(i, j, k)) continue;
          double val = eval(1, xh, NX, i, yi, NY + 1, j, zh, NZ, k);
          V[fv(i, j, k)] = (float)val;
          ss += val * val; ++cnt;
        }
    }
    if (k > 0 && k < NZ) {
      for (size_t j = 0; j < NY; ++j)
        for (size_t i = 0; i < NX; ++i) {
          if (S(i, j, k - 1) || S(i, j, k)) continue;
          double val = eval(2, xh, NX, i, yh, NY, j, zi, NZ + 1, k);
          W[fw(i, j, k)] = (float)val;
          ss += val * val; ++cnt;
        }
    }
    sumsq[k] = ss;
    count[k] = cnt;
  });
  double ss = 0.0;
  size_t cnt = 0;
  for (size_t k = 0; k <= NZ; ++k) { ss += sumsq[k]; cnt += count[k]; }
  double rms = cnt ? std::sqrt(ss / (double)cnt) : 1.0;
  float inv_rms = rms > 0 ? (float)(1.0 / rms) : 1.f;
  if (us) for (size_t q = 0; q < nu; ++q) us[q] = U[q] * inv_rms;
  if (vs) for (size_t q = 0; q < nv; ++q) vs[q] = V[q] * inv_rms;
  if (ws) for (size_t q = 0; q < nw; ++q) ws[q] = W[q] * inv_rms;
  return FSP_OK;
}
