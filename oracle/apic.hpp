// CPU ORACLE — test infrastructure only (see oracle.hpp).
//
// The pressure projection of the reference's APIC-3D solver
// (code/water-core/src/apic3d.rs:887-1080), restated: cells labelled AIR,
// WATER or SOLID, a level set phi at the cell centres, the ghost fraction
// theta = max(phi_c / (phi_c - phi_a), 0.01) towards an air neighbour, the
// operator without 1/dx^2, the right-hand side -(rho dx^2/dt) div u* / dx,
// Jacobi-preconditioned CG with f64 dot products to ||r||^2 <= 1e-12 ||b||^2
// (at most 4000 iterations), the correction u -= dt/(rho dx) grad p on every
// face that touches water (never on a face next to a solid), and the
// divergence metric over the water cells whose six neighbours are water.
// T = float follows the reference's arithmetic; T = double is the FP64
// oracle of the same system.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace orc {
namespace apic {

enum : uint8_t { AIR = 0, WATER = 1, SOLID = 2 };
constexpr float THETA_MIN = 0.01f;          // apic3d.rs:724
constexpr double TOLERANCE2 = 1e-12;        // apic3d.rs:716
constexpr uint32_t MAX_ITERATIONS = 4000;   // apic3d.rs:717

struct Result {
  uint32_t iterations;
  double residual;    // ||r|| / ||b|| at the stop
  double divergence;  // the metric after the correction
};

template <class T>
struct Projection {
  int nx, ny, nz;
  T dx;
  const uint8_t* label;
  const float* phi;
  std::vector<T> p, rhs, diag, r, z, d, q;

  size_t cell(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }
  size_t fu(int i, int j, int k) const { return ((size_t)k * ny + j) * (nx + 1) + i; }
  size_t fv(int i, int j, int k) const { return ((size_t)k * (ny + 1) + j) * nx + i; }
  size_t fw(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }

  // apic3d.rs:887-900: (neighbour cell, shared face, axis), or none at the
  // domain's boundary, in the order x-, x+, y-, y+, z-, z+
  struct Nb {
    bool has;
    size_t n, f;
    int axis;
  };
  void neighbours(int i, int j, int k, Nb out[6]) const {
    out[0] = {i > 0, i > 0 ? cell(i - 1, j, k) : 0, fu(i, j, k), 0};
    out[1] = {i + 1 < nx, i + 1 < nx ? cell(i + 1, j, k) : 0, fu(i + 1, j, k), 0};
    out[2] = {j > 0, j > 0 ? cell(i, j - 1, k) : 0, fv(i, j, k), 1};
    out[3] = {j + 1 < ny, j + 1 < ny ? cell(i, j + 1, k) : 0, fv(i, j + 1, k), 1};
    out[4] = {k > 0, k > 0 ? cell(i, j, k - 1) : 0, fw(i, j, k), 2};
    out[5] = {k + 1 < nz, k + 1 < nz ? cell(i, j, k + 1) : 0, fw(i, j, k + 1), 2};
  }
  // apic3d.rs:902-907
  T theta(size_t c, size_t a) const {
    const T fc = (T)phi[c], fa = (T)phi[a];
    const T t = fc / (fc - fa);
    return t > (T)THETA_MIN ? t : (T)THETA_MIN;
  }
  static double dot(const std::vector<T>& a, const std::vector<T>& b) {
    double s = 0.0;
    for (size_t c = 0; c < a.size(); ++c) s += (double)a[c] * (double)b[c];
    return s;
  }
  // apic3d.rs:912-935: y = A x on water cells (no 1/dx^2)
  void apply(const std::vector<T>& x, std::vector<T>& y) const {
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t c = cell(i, j, k);
          if (label[c] != WATER) { y[c] = T(0); continue; }
          Nb nb[6];
          neighbours(i, j, k, nb);
          T s = T(0);
          for (int m = 0; m < 6; ++m) {
            if (!nb[m].has) continue;
            const uint8_t l = label[nb[m].n];
            if (l == WATER) s += x[c] - x[nb[m].n];
            else if (l == AIR) s += x[c] / theta(c, nb[m].n);
          }
          y[c] = s;
        }
  }
  // apic3d.rs:940-1051
  Result project(float rho, float dt, float* u, float* v, float* w) {
    const size_t n = (size_t)nx * ny * nz;
    p.assign(n, T(0)); rhs.assign(n, T(0)); diag.assign(n, T(0));
    r.assign(n, T(0)); z.assign(n, T(0)); d.assign(n, T(0)); q.assign(n, T(0));
    const T scale = -(T)rho * dx * dx / (T)dt;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t c = cell(i, j, k);
          if (label[c] != WATER) continue;
          Nb nb[6];
          neighbours(i, j, k, nb);
          T div = T(0), dg = T(0);
          for (int m = 0; m < 6; ++m) {
            const T sign = m % 2 == 0 ? T(-1) : T(1);
            const T face = (T)(m == 0 ? u[fu(i, j, k)] : m == 1 ? u[fu(i + 1, j, k)] : m == 2 ? v[fv(i, j, k)]
                               : m == 3 ? v[fv(i, j + 1, k)] : m == 4 ? w[fw(i, j, k)] : w[fw(i, j, k + 1)]);
            div += sign * face;
            if (nb[m].has) {
              const uint8_t l = label[nb[m].n];
              if (l == WATER) dg += T(1);
              else if (l == AIR) dg += T(1) / theta(c, nb[m].n);
            }
          }
          rhs[c] = scale * div / dx;
          diag[c] = dg;
        }
    const double b2 = dot(rhs, rhs);
    r = rhs;
    for (size_t c = 0; c < n; ++c) z[c] = diag[c] > T(0) ? r[c] / diag[c] : T(0);
    d = z;
    double rz = dot(r, z), rr = b2;
    uint32_t it = 0;
    while (b2 > 0.0 && rr > TOLERANCE2 * b2 && it < MAX_ITERATIONS) {
      apply(d, q);
      const double dq = dot(d, q);
      if (!(dq > 0.0)) break;
      const T alpha = (T)(rz / dq);
      for (size_t c = 0; c < n; ++c) {
        p[c] += alpha * d[c];
        r[c] -= alpha * q[c];
        z[c] = diag[c] > T(0) ? r[c] / diag[c] : T(0);
      }
      const double zn = dot(r, z);
      const T beta = (T)(zn / rz);
      for (size_t c = 0; c < n; ++c) d[c] = z[c] + beta * d[c];
      rz = zn;
      rr = dot(r, r);
      ++it;
    }
    // correction of the faces that touch water, each from the cell on its
    // negative side; a face next to a solid is never corrected
    const T k1 = (T)dt / ((T)rho * dx);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t c = cell(i, j, k);
          Nb nb[6];
          neighbours(i, j, k, nb);
          for (int m = 1; m < 6; m += 2) {
            if (!nb[m].has) continue;
            const size_t nn = nb[m].n;
            if (label[c] == SOLID || label[nn] == SOLID) continue;
            const bool wc = label[c] == WATER, wn = label[nn] == WATER;
            T grad;
            if (wc && wn) grad = p[nn] - p[c];
            else if (wc) grad = -p[c] / theta(c, nn);
            else if (wn) grad = p[nn] / theta(nn, c);
            else continue;
            float* field = nb[m].axis == 0 ? u : nb[m].axis == 1 ? v : w;
            field[nb[m].f] = (float)((T)field[nb[m].f] - k1 * grad);
          }
        }
    Result res{};
    res.iterations = it;
    res.residual = b2 > 0.0 ? std::sqrt(rr / b2) : 0.0;
    res.divergence = divergence_metric(u, v, w);
    return res;
  }
  // apic3d.rs:1053-1080
  double divergence_metric(const float* u, const float* v, const float* w) const {
    float worst = 0.f;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const size_t c = cell(i, j, k);
          if (label[c] != WATER) continue;
          Nb nb[6];
          neighbours(i, j, k, nb);
          bool inner = true;
          for (int m = 0; m < 6; ++m)
            if (nb[m].has && label[nb[m].n] != WATER) inner = false;
          if (!inner) continue;
          const float div = u[fu(i + 1, j, k)] - u[fu(i, j, k)] + v[fv(i, j + 1, k)] - v[fv(i, j, k)] +
                            w[fw(i, j, k + 1)] - w[fw(i, j, k)];
          worst = std::fmax(worst, std::fabs(div));
        }
    const size_t nu = (size_t)(nx + 1) * ny * nz, nv = (size_t)nx * (ny + 1) * nz, nw = (size_t)nx * ny * (nz + 1);
    float umax = 0.f;
    for (size_t q2 = 0; q2 < nu; ++q2) umax = std::fmax(umax, std::fabs(u[q2]));
    for (size_t q2 = 0; q2 < nv; ++q2) umax = std::fmax(umax, std::fabs(v[q2]));
    for (size_t q2 = 0; q2 < nw; ++q2) umax = std::fmax(umax, std::fabs(w[q2]));
    return umax > 0.f ? (double)(worst / umax) : 0.0;
  }
};

}  // namespace apic
}  // namespace orc
