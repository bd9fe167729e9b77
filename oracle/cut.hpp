// CPU ORACLE — test infrastructure only (see oracle.hpp).
//
// The reference's cut of a bottom height field (code/water-core/src/
// delta3d_cut.rs:80-258, with delta_projection.rs:795-827): volume fractions
// and the apertures of the three face families of a grid whose bottom is
// given at the column centres. Restated in f32 with the reference's
// expression order (the build keeps every operation separately rounded,
// -ffp-contract=off), since the projection pins (the bump-bottom S385
// iteration counts, MULTIGRILLE-3D-S385.md:59-60) depend on these values.
//
// Model: the bottom at the footprint corners by nested averages (along x as
// the 2D edge, then along y); each footprint cut into four triangles around
// its centre (the mean of its corners), the bottom linear on each; side
// apertures the 1-D cut formula along the face's edge; horizontal apertures
// the part of the footprint where the bottom lies below the face; volume
// fractions the exact integral of clamp((z - bottom)/dx) over the footprint.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace orc {
namespace cut {

inline float clamp01(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }

// Volume::cut_fraction (delta_projection.rs:816-827): the wet part of a
// side face of top `top` whose bottom edge runs linearly from a to b.
inline float cut_fraction(float a, float b, float top, float dx) {
  const float low = std::min(a, b), high = std::max(a, b);
  if (low >= top) return 0.f;
  if (high <= top - dx) return 1.f;
  if (low == high) return clamp01((top - low) / dx);
  const float span = high - low;
  const float full = clamp01((top - dx - low) / span);
  const float wet = clamp01((top - low) / span);
  const float h0 = clamp01((top - low) / dx);
  const float h1 = clamp01((top - high) / dx);
  return clamp01(full + (wet - full) * (0.5f * (h0 + h1)));
}

// Volume::open_below (delta_projection.rs:795-805): the part of a segment
// from a to b lying below z.
inline float open_below(float a, float b, float z) {
  float open;
  if (a < z && b < z) {
    open = 1.f;
  } else if (a >= z && b >= z) {
    open = 0.f;
  } else {
    const float t = (z - a) / (b - a);
    open = a < z ? t : 1.f - t;
  }
  return clamp01(open);
}

inline void sort3(const float w[3], float v[3]) {
  v[0] = w[0]; v[1] = w[1]; v[2] = w[2];
  std::sort(v, v + 3);
}

// Normalised area of a triangle where the linear field with vertex values w
// lies strictly below s (delta3d_cut.rs:120-136).
inline float triangle_below(const float w[3], float s) {
  float v[3];
  sort3(w, v);
  const float v0 = v[0], v1 = v[1], v2 = v[2];
  if (s <= v0) return 0.f;
  if (s >= v2) return 1.f;
  if (s <= v1) return (s - v0) * (s - v0) / ((v1 - v0) * (v2 - v0));
  return 1.f - (v2 - s) * (v2 - s) / ((v2 - v0) * (v2 - v1));
}

// Mean over [0, d] of the area below s, w relative to the cell's bottom:
// the fluid part of a slab of height d over a linear bottom
// (delta3d_cut.rs:138-170), from the piecewise-cubic primitive.
inline float triangle_fraction(const float w[3], float d) {
  float v[3];
  sort3(w, v);
  const float v0 = v[0], v1 = v[1], v2 = v[2];
  if (v2 <= 0.f) return 1.f;
  if (v0 >= d) return 0.f;
  const float g1 = v2 > v0 ? (v1 - v0) * (v1 - v0) / (3.f * (v2 - v0)) : 0.f;
  auto primitive = [&](float s) -> float {
    if (s <= v0) return 0.f;
    if (v2 == v0) return s - v0;
    if (s <= v1) return (s - v0) * (s - v0) * (s - v0) / (3.f * (v1 - v0) * (v2 - v0));
    if (s < v2) {
      const float a = v2 - v1, b = v2 - s;
      return g1 + (s - v1) - (a * a * a - b * b * b) / (3.f * (v2 - v0) * a);
    }
    const float a = v2 - v1;
    const float g2 = g1 + a - a * a / (3.f * (v2 - v0));
    return g2 + (s - v2);
  };
  return clamp01((primitive(d) - primitive(0.f)) / d);
}

struct Cut {
  std::vector<float> frac, open_u, open_v, open_w, floor;
};

// cut (delta3d_cut.rs:172-258) for bottom heights at the nx*ny column
// centres.
inline Cut cut(int nx, int ny, int nz, float dx, const float* bottom) {
  // corners (delta3d_cut.rs:86-111): along x the mean of two columns (the
  // 2D edge), along y the mean of two of these edges; the borders copy the
  // neighbouring column
  auto b = [&](int i, int j) { return bottom[(size_t)j * nx + i]; };
  auto edge = [&](int i, int j) -> float {
    if (i == 0) return b(0, j);
    if (i == nx) return b(nx - 1, j);
    return 0.5f * (b(i - 1, j) + b(i, j));
  };
  std::vector<float> k0((size_t)(nx + 1) * (ny + 1));
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i)
      k0[(size_t)j * (nx + 1) + i] =
          j == 0 ? edge(i, 0) : (j == ny ? edge(i, ny - 1) : 0.5f * (edge(i, j - 1) + edge(i, j)));
  auto corner = [&](int i, int j) { return k0[(size_t)j * (nx + 1) + i]; };
  // footprint corners [bottom-left, bottom-right, top-right, top-left]
  auto footprint = [&](int i, int j, float c[4]) {
    c[0] = corner(i, j); c[1] = corner(i + 1, j); c[2] = corner(i + 1, j + 1); c[3] = corner(i, j + 1);
  };
  auto plane_in_y = [](const float c[4]) { return c[0] == c[3] && c[1] == c[2]; };
  // the four triangles around the footprint's centre (delta3d_cut.rs:164-168)
  auto triangles = [](const float c[4], float t[4][3]) {
    const float m = 0.25f * (c[0] + c[1] + c[2] + c[3]);
    const float tt[4][3] = {{c[0], c[1], m}, {c[1], c[2], m}, {c[2], c[3], m}, {c[3], c[0], m}};
    for (int a = 0; a < 4; ++a)
      for (int q = 0; q < 3; ++q) t[a][q] = tt[a][q];
  };
  Cut out;
  out.frac.assign((size_t)nx * ny * nz, 0.f);
  out.open_u.assign((size_t)(nx + 1) * ny * nz, 0.f);
  out.open_v.assign((size_t)nx * (ny + 1) * nz, 0.f);
  out.open_w.assign((size_t)nx * ny * (nz + 1), 0.f);
  for (int k = 0; k < nz; ++k) {
    const float bas = (float)k * dx, top = (float)(k + 1) * dx;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        float c[4];
        footprint(i, j, c);
        float f;
        if (plane_in_y(c)) {
          f = cut_fraction(c[0], c[1], top, dx);
        } else {
          float t[4][3];
          triangles(c, t);
          auto part = [&](const float w[3]) {
            const float r[3] = {w[0] - bas, w[1] - bas, w[2] - bas};
            return triangle_fraction(r, dx);
          };
          f = 0.25f * (part(t[0]) + part(t[1]) + part(t[2]) + part(t[3]));
        }
        out.frac[((size_t)k * ny + j) * nx + i] = f;
      }
    // side faces: the 1-D formula along the edge; a bottom constant along
    // the edge keeps the 2D expression itself
    auto side = [&](float a, float bb) { return a == bb ? clamp01((top - a) / dx) : cut_fraction(a, bb, top, dx); };
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i <= nx; ++i)
        out.open_u[((size_t)k * ny + j) * (nx + 1) + i] = (i == 0 || i == nx) ? 0.f : side(corner(i, j), corner(i, j + 1));
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i < nx; ++i)
        out.open_v[((size_t)k * (ny + 1) + j) * nx + i] = (j == 0 || j == ny) ? 0.f : side(corner(i, j), corner(i + 1, j));
  }
  // horizontal faces: the part of the footprint where the bottom passes
  // below; the domain's bottom closed; the lid open where water is below
  for (int k = 1; k <= nz; ++k) {
    const float z = (float)k * dx;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        float c[4];
        footprint(i, j, c);
        float a;
        if (plane_in_y(c)) {
          a = open_below(c[0], c[1], z);
        } else {
          float t[4][3];
          triangles(c, t);
          a = 0.25f * (triangle_below(t[0], z) + triangle_below(t[1], z) + triangle_below(t[2], z) +
                       triangle_below(t[3], z));
        }
        out.open_w[((size_t)k * ny + j) * nx + i] = a;
      }
  }
  // a fluid cell with no open face would make the operator singular: solid
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t c = ((size_t)k * ny + j) * nx + i;
        if (out.frac[c] == 0.f) continue;
        const float a = out.open_u[((size_t)k * ny + j) * (nx + 1) + i] + out.open_u[((size_t)k * ny + j) * (nx + 1) + i + 1] +
                        out.open_v[((size_t)k * (ny + 1) + j) * nx + i] + out.open_v[((size_t)k * (ny + 1) + j + 1) * nx + i] +
                        out.open_w[((size_t)k * ny + j) * nx + i] + out.open_w[((size_t)(k + 1) * ny + j) * nx + i];
        if (a == 0.f) out.frac[c] = 0.f;
      }
  out.floor.resize((size_t)nx * ny);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      float c[4];
      footprint(i, j, c);
      out.floor[(size_t)j * nx + i] = std::max(std::max(c[0], c[1]), std::max(c[2], c[3]));
    }
  return out;
}

}  // namespace cut
}  // namespace orc
