// Compact Galerkin coarse levels on quads.
//
// A Galerkin coarse level (P^T A P / 16, P piecewise constant) couples a cell
// to a neighbour only through a face whose coupling is non-zero, and such a
// face always joins two active cells; an inactive cell has no coupling and no
// Dirichlet part. The row of the operator is therefore a function of the
// couplings and the Dirichlet part alone, with no cell kinds:
//
//   row(x)_c = inv * ( sum over faces f in order x-, x+, y-, y+, z-, z+ with
//                      a_f != 0 of a_f (x_c - x_f)  +  dd_c x_c )
//   diag_c   = inv * ( sum of those a_f  +  dd_c )
//
// which is row_sums (delta3d_multigrid.rs:216-255) on such a level, term for
// term. The couplings are stored cell-aligned (a cell holds its x-, y- and z-
// faces; its + faces are the next cell's - faces) in binary16: on level l
// (1-based) they are multiples of 4^-l no larger than 1, exact in binary16 up
// to l = 5. Every load of a pass is independent of the others (no kind test
// before a neighbour load), so a pass costs one round of memory latency.
#pragma once
#include "kernels_fast.h"

namespace fsp {

typedef _Float16 gc_h4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 gc_ld4(const uint16_t* a, size_t c) {
  const gc_h4 w = *reinterpret_cast<const gc_h4*>(a + c);
  return make_float4((float)w.x, (float)w.y, (float)w.z, (float)w.w);
}
__device__ __forceinline__ float gc_ld(const uint16_t* a, size_t c) {
  return (float)*reinterpret_cast<const _Float16*>(a + c);
}

__device__ __forceinline__ float gc_diag(const float a[6], float dd, float inv);

// Cell-aligned binary16 couplings of a Galerkin level from its face arrays,
// and each row's diagonal from those binary16 values (what the passes compute).
__global__ void __launch_bounds__(BLOCK) k_gc_build(Level L) {
  FSP_FOR(c, L.n) {
    int i, j, k;
    ldecode(L, c, i, j, k);
    const size_t ux = ((size_t)k * L.ny + j) * (L.nx + 1) + i;
    const size_t vy = ((size_t)k * (L.ny + 1) + j) * L.nx + i;
    const size_t wz = ((size_t)k * L.ny + j) * L.nx + i;
    const float ax = i > 0 ? L.ou[ux] : 0.f;
    const float ay = j > 0 ? L.ov[vy] : 0.f;
    const float az = k > 0 ? L.ow[wz] : 0.f;
    const float bx = i + 1 < L.nx ? L.ou[ux + 1] : 0.f;
    const float by = j + 1 < L.ny ? L.ov[vy + L.nx] : 0.f;
    const float bz = k + 1 < L.nz ? L.ow[wz + (size_t)L.nx * L.ny] : 0.f;
    reinterpret_cast<_Float16*>(L.cx)[c] = (_Float16)ax;
    reinterpret_cast<_Float16*>(L.cy)[c] = (_Float16)ay;
    reinterpret_cast<_Float16*>(L.cz)[c] = (_Float16)az;
    const float a[6] = {(float)(_Float16)ax, (float)(_Float16)bx, (float)(_Float16)ay,
                        (float)(_Float16)by, (float)(_Float16)az, (float)(_Float16)bz};
    L.dg[c] = gc_diag(a, L.dd[c], L.inv);
  }
}

// The six couplings of the four cells of a quad, and their Dirichlet parts.
struct GcQuad {
  float4 xm, ym, yp, zm, zp, dd;
  float xp_last;  // x+ face of the last cell (x- face of the next quad's first)
};

__device__ __forceinline__ GcQuad gc_quad(const Level& L, size_t c0, int x0, int y, int k) {
  const size_t plane = (size_t)L.nx * L.ny;
  GcQuad g;
  g.xm = gc_ld4(L.cx, c0);
  g.xp_last = (x0 + 4 < L.nx) ? gc_ld(L.cx, c0 + 4) : 0.f;
  g.ym = gc_ld4(L.cy, c0);
  g.yp = (y + 1 < L.ny) ? gc_ld4(L.cy, c0 + L.nx) : f4zero();
  g.zm = gc_ld4(L.cz, c0);
  g.zp = (k + 1 < L.nz) ? gc_ld4(L.cz, c0 + plane) : f4zero();
  g.dd = ld4(L.dd, c0);
  return g;
}

__device__ __forceinline__ void gc_faces(const GcQuad& g, int e, float a[6]) {
  a[0] = q_get(g.xm, e);
  a[1] = e < 3 ? q_get(g.xm, e + 1) : g.xp_last;
  a[2] = q_get(g.ym, e);
  a[3] = q_get(g.yp, e);
  a[4] = q_get(g.zm, e);
  a[5] = q_get(g.zp, e);
}

__device__ __forceinline__ float gc_diag(const float a[6], float dd, float inv) {
#pragma clang fp contract(off)
  float acc = 0.f;
#pragma unroll
  for (int f = 0; f < 6; ++f)
    if (a[f] != 0.f) acc += a[f];
  acc += dd;
  return acc * inv;
}

__device__ __forceinline__ float gc_row(const float a[6], float dd, float inv, float pc, const float nb[6]) {
#pragma clang fp contract(off)
  float acc = 0.f;
#pragma unroll
  for (int f = 0; f < 6; ++f)
    if (a[f] != 0.f) acc += a[f] * (pc - nb[f]);
  acc += dd * pc;
  return acc * inv;
}

// One colour of a red-black pass on a compact level, quads of four x-cells.
// MODE 0: in place (x == xout).
// MODE 1: first pass from x = 0: colour cells r/diag, the others 0.
// MODE 2: x is first prolonged from the coarser level C (x += parent, as
//         prolong_add3 stores it), then the colour is updated; out of place.
// x, xout, the parents xp and L.r are held in storage ST.
template <int MODE, int ST>
__global__ void __launch_bounds__(QBLOCK) k_gc_gs(Level L, QuadGrid Q, const typename Store<ST>::T* x,
                                                  typename Store<ST>::T* xout, int colour, Level C,
                                                  const typename Store<ST>::T* xp, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  using T = typename SV::T;
  if (!ctl->active) return;
  int qx, y, k;
  const bool valid = quad_coords(Q, qx, y, k);
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t c0 = valid ? ((size_t)k * L.ny + y) * L.nx + 4 * (size_t)qx : 0;
  const T* lr = reinterpret_cast<const T*>(L.r);
  GcQuad g{};
  float4 rq = f4zero();
  if (valid) {
    g = gc_quad(L, c0, 4 * qx, y, k);
    rq = SV::ld4(lr, c0);
  }
  if (MODE == 1) {
    if (!valid) return;
    float4 out = f4zero();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (((4 * qx + e + y + k) & 1) != colour) continue;
      float a[6];
      gc_faces(g, e, a);
      const float dg = gc_diag(a, q_get(g.dd, e), L.inv);
      if (dg > 0.f) q_set(out, e, SV::rnd(q_get(rq, e) / dg));
    }
    SV::st4(xout, c0, out);
    return;
  }
  auto parent = [&](int yy, int kk) -> float2 { return SV::ld2(xp, lidx(C, 2 * qx, yy >> 1, kk >> 1)); };
  auto prolong = [&](float4& v, float2 a) {
    v.x = SV::rnd(v.x + a.x); v.y = SV::rnd(v.y + a.x); v.z = SV::rnd(v.z + a.y); v.w = SV::rnd(v.w + a.y);
  };
  float4 xc = f4zero(), ym = f4zero(), yp = f4zero(), zm = f4zero(), zp = f4zero();
  if (valid) {
    xc = SV::ld4(x, c0);
    if (y > 0) ym = SV::ld4(x, c0 - L.nx);
    if (y + 1 < L.ny) yp = SV::ld4(x, c0 + L.nx);
    if (k > 0) zm = SV::ld4(x, c0 - plane);
    if (k + 1 < L.nz) zp = SV::ld4(x, c0 + plane);
    if (MODE == 2) {
      prolong(xc, parent(y, k));
      if (y > 0) prolong(ym, parent(y - 1, k));
      if (y + 1 < L.ny) prolong(yp, parent(y + 1, k));
      if (k > 0) prolong(zm, parent(y, k - 1));
      if (k + 1 < L.nz) prolong(zp, parent(y, k + 1));
    }
  }
  float left, right;
  x_neighbours(Q, valid, qx, xc, left, right, [&](int off) {
    float v = SV::ld(x, c0 + off);
    if (MODE == 2) v = SV::rnd(v + SV::ld(xp, lidx(C, (4 * qx + off) >> 1, y >> 1, k >> 1)));
    return v;
  });
  if (!valid) return;
  float4 out = xc;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (((4 * qx + e + y + k) & 1) != colour) continue;
    float a[6];
    gc_faces(g, e, a);
    const float dd = q_get(g.dd, e);
    const float dg = gc_diag(a, dd, L.inv);
    if (!(dg > 0.f)) continue;
    const float pc = q_get(xc, e);
    const float nb[6] = {e == 0 ? left : q_get(xc, e - 1), e == 3 ? right : q_get(xc, e + 1),
                         q_get(ym, e), q_get(yp, e), q_get(zm, e), q_get(zp, e)};
    const float row = gc_row(a, dd, L.inv, pc, nb);
    q_set(out, e, SV::rnd(pc + (q_get(rq, e) - row) / dg));
  }
  SV::st4(xout, c0, out);
}

// The first two passes of a compact level's pre-smoothing from x = 0 in one
// launch (k_gc_gs<1> colour 0, then k_gc_gs<0> colour 1): the first pass is
// local (colour-0 cells r/diag, the others 0), so the second recomputes its
// neighbours' values from their r and diagonal (L.dg) instead of reading them
// back. Same values, same order: bit-identical to the two launches.
template <int ST>
__global__ void __launch_bounds__(QBLOCK) k_gc_gs_first(Level L, QuadGrid Q, typename Store<ST>::T* xout,
                                                        Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  using T = typename SV::T;
  if (!ctl->active) return;
  int qx, y, k;
  const bool valid = quad_coords(Q, qx, y, k);
  const size_t plane = (size_t)L.nx * L.ny;
  const size_t c0 = valid ? ((size_t)k * L.ny + y) * L.nx + 4 * (size_t)qx : 0;
  const T* lr = reinterpret_cast<const T*>(L.r);
  auto first = [&](int x, int yy, int kk, float dg, float rv) -> float {
    if (((x + yy + kk) & 1) != 0) return 0.f;
    return dg > 0.f ? SV::rnd(rv / dg) : 0.f;
  };
  auto first4 = [&](const float4& dq, const float4& rq4, int yy, int kk) {
    float4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) q_set(v, e, first(4 * qx + e, yy, kk, q_get(dq, e), q_get(rq4, e)));
    return v;
  };
  GcQuad g{};
  float4 rq = f4zero(), xc = f4zero(), ym = f4zero(), yp = f4zero(), zm = f4zero(), zp = f4zero();
  if (valid) {
    g = gc_quad(L, c0, 4 * qx, y, k);
    rq = SV::ld4(lr, c0);
    const float4 dq = ld4(L.dg, c0);
    const bool hym = y > 0, hyp = y + 1 < L.ny, hzm = k > 0, hzp = k + 1 < L.nz;
    const size_t cym = c0 - L.nx, cyp = c0 + L.nx, czm = c0 - plane, czp = c0 + plane;
    const float4 r_ym = hym ? SV::ld4(lr, cym) : f4zero(), d_ym = hym ? ld4(L.dg, cym) : f4zero();
    const float4 r_yp = hyp ? SV::ld4(lr, cyp) : f4zero(), d_yp = hyp ? ld4(L.dg, cyp) : f4zero();
    const float4 r_zm = hzm ? SV::ld4(lr, czm) : f4zero(), d_zm = hzm ? ld4(L.dg, czm) : f4zero();
    const float4 r_zp = hzp ? SV::ld4(lr, czp) : f4zero(), d_zp = hzp ? ld4(L.dg, czp) : f4zero();
    xc = first4(dq, rq, y, k);
    if (hym) ym = first4(d_ym, r_ym, y - 1, k);
    if (hyp) yp = first4(d_yp, r_yp, y + 1, k);
    if (hzm) zm = first4(d_zm, r_zm, y, k - 1);
    if (hzp) zp = first4(d_zp, r_zp, y, k + 1);
  }
  float left, right;
  x_neighbours(Q, valid, qx, xc, left, right, [&](int off) {
    const size_t c = c0 + off;
    return first(4 * qx + off, y, k, L.dg[c], SV::ld(lr, c));
  });
  if (!valid) return;
  float4 out = xc;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (((4 * qx + e + y + k) & 1) != 1) continue;
    float a[6];
    gc_faces(g, e, a);
    const float dd = q_get(g.dd, e);
    const float dg = gc_diag(a, dd, L.inv);
    if (!(dg > 0.f)) continue;
    const float pc = q_get(xc, e);
    const float nb[6] = {e == 0 ? left : q_get(xc, e - 1), e == 3 ? right : q_get(xc, e + 1),
                         q_get(ym, e), q_get(yp, e), q_get(zm, e), q_get(zp, e)};
    q_set(out, e, SV::rnd(pc + (q_get(rq, e) - gc_row(a, dd, L.inv, pc, nb)) / dg));
  }
  SV::st4(xout, c0, out);
}

// Residual r - row(x) of a compact level restricted to the next level
// (delta3d_multigrid.rs:312-321, restrict3 :258-272): per (quad of the fine
// level, coarse row J, coarse plane K), 16 fine cells -> 2 coarse cells,
// children summed in the reference's order, the work split over two threads
// (fine plane 2K in the first 128 threads of a workgroup, 2K + 1 in the last
// 128) with the first half's partial sums carried to the second through LDS
// and continued in child order: bit-identical to one thread per pair, with
// twice the waves (10 -> 6 us on the 128x128x64 level). H: the restriction's
// quad grid for 128-thread halves.
template <int ST>
__global__ void __launch_bounds__(QBLOCK) k_gc_restrict2(Level F, QuadGrid H, const typename Store<ST>::T* xf,
                                                         Level C, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  using T = typename SV::T;
  __shared__ float2 carry[QBLOCK / 2];
  if (!ctl->active) return;
  const int local = (int)(threadIdx.x & (QBLOCK / 2 - 1)), dz = (int)(threadIdx.x / (QBLOCK / 2));
  const unsigned Lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned bx = Lb % (unsigned)H.nbx, tb = Lb / (unsigned)H.nbx;
  const int qx = (int)bx * H.QX + (local & (H.QX - 1));
  const int J = (int)(tb % (unsigned)H.nby) * H.RY + local / H.QX;
  const int K = (int)(tb / (unsigned)H.nby);
  const bool valid = qx < H.nq && J < H.rows && K < H.planes;
  const size_t plane = (size_t)F.nx * F.ny;
  const T* fr = reinterpret_cast<const T*>(F.r);
  const int k = 2 * K + dz;
  float t[2][4];  // [dy][e]
#pragma unroll
  for (int dy = 0; dy < 2; ++dy) {
    const int y = 2 * J + dy;
    const size_t c0 = valid ? ((size_t)k * F.ny + y) * F.nx + 4 * (size_t)qx : 0;
    GcQuad g{};
    float4 xc = f4zero(), rq = f4zero(), ym = f4zero(), yp = f4zero(), zm = f4zero(), zp = f4zero();
    if (valid) {
      g = gc_quad(F, c0, 4 * qx, y, k);
      rq = SV::ld4(fr, c0);
      xc = SV::ld4(xf, c0);
      if (y > 0) ym = SV::ld4(xf, c0 - F.nx);
      if (y + 1 < F.ny) yp = SV::ld4(xf, c0 + F.nx);
      if (k > 0) zm = SV::ld4(xf, c0 - plane);
      if (k + 1 < F.nz) zp = SV::ld4(xf, c0 + plane);
    }
    float left, right;
    x_neighbours(H, valid, qx, xc, left, right, [&](int off) { return SV::ld(xf, c0 + off); });
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float a[6];
      gc_faces(g, e, a);
      const float pc = q_get(xc, e);
      const float nb[6] = {e == 0 ? left : q_get(xc, e - 1), e == 3 ? right : q_get(xc, e + 1),
                           q_get(ym, e), q_get(yp, e), q_get(zm, e), q_get(zp, e)};
      t[dy][e] = q_get(rq, e) - gc_row(a, q_get(g.dd, e), F.inv, pc, nb);
    }
  }
  float2 sum = make_float2(0.f, 0.f);
  if (dz == 0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float s = 0.f;
      s += t[0][2 * h]; s += t[0][2 * h + 1];
      s += t[1][2 * h]; s += t[1][2 * h + 1];
      if (h == 0) sum.x = s; else sum.y = s;
    }
    carry[local] = sum;
  }
  __syncthreads();
  if (dz == 0 || !valid) return;
  sum = carry[local];
  float2 out;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    float s = h == 0 ? sum.x : sum.y;
    s += t[0][2 * h]; s += t[0][2 * h + 1];
    s += t[1][2 * h]; s += t[1][2 * h + 1];
    if (h == 0) out.x = 0.125f * s; else out.y = 0.125f * s;
  }
  SV::st2(reinterpret_cast<T*>(C.r), lidx(C, 2 * qx, J, K), out);
}

// ---------------------------------------------------------------- tiled compact levels
// On a compact level small enough that its one-pass launches are bound by
// launch latency rather than by memory, one workgroup takes an 8x8x8 tile
// with a halo of 4 cells and runs every colour pass of the pre-smoothing and
// the restriction of the residual (k_gb_pre), or the prolongation and every
// pass of the post-smoothing (k_gb_post), in LDS. The halo's values are
// recomputed by each workgroup: pass m of n runs on the tile grown by n - m
// cells (n - m + 1 before a restriction), which is exactly the region whose
// next-pass neighbours it feeds, so every interior value is computed from the
// operands k_gc_gs / k_gc_restrict2 use, in their order, and is bit-identical.
// Four halo cells cover two sweeps of two colours. Box cells outside the
// level hold zero couplings, Dirichlet part, r and x, which the row excludes
// exactly as the edge tests of the quad kernels do.
#ifndef FSP_GB_MAX_CELLS
#define FSP_GB_MAX_CELLS (1 << 18)  // largest compact level smoothed in tiles
#endif
constexpr int GB_T = 8;                    // tile interior per dimension (even: colours and children align)
constexpr int GB_H = 4;                    // halo
constexpr int GB_B = GB_T + 2 * GB_H + 1;  // box per dimension (+1: the + faces of the outermost halo cells)
constexpr int GB_B2 = GB_B * GB_B;
constexpr int GB_N = GB_B2 * GB_B;
constexpr int GB_I = GB_T * GB_T * GB_T;   // interior cells
constexpr int GB_THREADS = 1024;
constexpr int GB_MAX_SWEEPS = GB_H / 2;
static_assert(GB_THREADS >= GB_I, "one thread per interior cell");

struct GbLds {
  float x[GB_N], r[GB_N], dd[GB_N], dg[GB_N];
  _Float16 fx[GB_N], fy[GB_N], fz[GB_N];
  float res[GB_I];
};

struct GbTile {
  int i0, j0, k0;  // level coordinates of box cell (0, 0, 0)
};

__device__ __forceinline__ GbTile gb_tile(const Level& L) {
  const int tx = (L.nx + GB_T - 1) / GB_T, ty = (L.ny + GB_T - 1) / GB_T;
  const int b = (int)blockIdx.x;
  GbTile t;
  t.i0 = (b % tx) * GB_T - GB_H;
  t.j0 = ((b / tx) % ty) * GB_T - GB_H;
  t.k0 = (b / (tx * ty)) * GB_T - GB_H;
  return t;
}

__device__ __forceinline__ bool gb_inside(const Level& L, int i, int j, int k) {
  return (unsigned)i < (unsigned)L.nx && (unsigned)j < (unsigned)L.ny && (unsigned)k < (unsigned)L.nz;
}

// Couplings, Dirichlet parts, diagonals (L.dg, gc_diag of the same binary16
// couplings) and r of the box, and its starting x: pre-smoothing (POST
// false): the first pass from x = 0 (k_gc_gs<1>, colour 0 cells r/diag, the
// others 0); post-smoothing: the level's x prolonged from its parents in C as
// k_gc_gs<2> prolongs it. One round of loads: cells outside the level (or
// past the box) load cell 0 and are replaced by zeros.
template <int ST, bool POST>
__device__ __forceinline__ void gb_load(GbLds& s, const Level& L, const GbTile& t, const typename Store<ST>::T* x,
                                        const Level& C, const typename Store<ST>::T* xp) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  const typename SV::T* lr = reinterpret_cast<const typename SV::T*>(L.r);
  constexpr int ITER = (GB_N + GB_THREADS - 1) / GB_THREADS;
  float ax[ITER], ay[ITER], az[ITER], dgv[ITER], dd[ITER], r[ITER], v[ITER];
  bool in[ITER], colour0[ITER];
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int q = (int)threadIdx.x + it * GB_THREADS;
    const int bz = q / GB_B2, by = (q / GB_B) % GB_B, bx = q % GB_B;
    const int i = t.i0 + bx, j = t.j0 + by, k = t.k0 + bz;
    in[it] = q < GB_N && gb_inside(L, i, j, k);
    colour0[it] = ((bx + by + bz) & 1) == 0;  // box and level parities agree: i0, j0, k0 are even
    const size_t c = in[it] ? lidx(L, i, j, k) : 0;
    ax[it] = gc_ld(L.cx, c);
    ay[it] = gc_ld(L.cy, c);
    az[it] = gc_ld(L.cz, c);
    dd[it] = L.dd[c];
    if (!POST) dgv[it] = L.dg[c];
    r[it] = SV::ld(lr, c);
    if (POST) v[it] = SV::rnd(SV::ld(x, c) + SV::ld(xp, in[it] ? lidx(C, i >> 1, j >> 1, k >> 1) : 0));
  }
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int q = (int)threadIdx.x + it * GB_THREADS;
    if (q >= GB_N) break;
    const bool o = in[it];
    const float d = o ? dd[it] : 0.f, rr = o ? r[it] : 0.f, dg = (o && !POST) ? dgv[it] : 0.f;
    s.fx[q] = (_Float16)(o ? ax[it] : 0.f);
    s.fy[q] = (_Float16)(o ? ay[it] : 0.f);
    s.fz[q] = (_Float16)(o ? az[it] : 0.f);
    s.dd[q] = d;
    if (!POST) s.dg[q] = dg;
    s.r[q] = rr;
    float x0 = 0.f;
    if (POST) {
      if (o) x0 = v[it];
    } else if (o && colour0[it]) {
      if (dg > 0.f) x0 = SV::rnd(rr / dg);
    }
    s.x[q] = x0;
  }
  __syncthreads();
}

__device__ __forceinline__ void gb_faces(const GbLds& s, int c, float a[6]) {
  a[0] = (float)s.fx[c];
  a[1] = (float)s.fx[c + 1];
  a[2] = (float)s.fy[c];
  a[3] = (float)s.fy[c + GB_B];
  a[4] = (float)s.fz[c];
  a[5] = (float)s.fz[c + GB_B2];
}

__device__ __forceinline__ float gb_row(const GbLds& s, int c, const float a[6], float inv) {
  const float nb[6] = {s.x[c - 1], s.x[c + 1], s.x[c - GB_B], s.x[c + GB_B], s.x[c - GB_B2], s.x[c + GB_B2]};
  return gc_row(a, s.dd[c], inv, s.x[c], nb);
}

// One colour (k_gc_gs<0>) on the tile grown by GROW cells (1 <= GROW < GB_H:
// every neighbour read stays in the box). DG: the diagonals were loaded into
// LDS (pre-smoothing: faster); else they are formed from the couplings
// (post-smoothing: the extra load cost more than the arithmetic it saves).
template <int ST, int GROW, bool DG>
__device__ __forceinline__ void gb_pass(GbLds& s, float inv, int colour) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  static_assert(GROW >= 0 && GROW < GB_H, "pass region inside the box");
  constexpr int e = GB_T + 2 * GROW, o = GB_H - GROW, half = e / 2, cnt = half * e * e;
#pragma unroll
  for (int it = 0; it < (cnt + GB_THREADS - 1) / GB_THREADS; ++it) {
    const int q = (int)threadIdx.x + it * GB_THREADS;
    if (q >= cnt) break;
    const int u = q % half, row = q / half;
    const int by = o + row % e, bz = o + row / e;
    const int bx = o + 2 * u + ((colour + by + bz + o) & 1);
    const int c = (bz * GB_B + by) * GB_B + bx;
    float a[6];
    gb_faces(s, c, a);
    const float dg = DG ? s.dg[c] : gc_diag(a, s.dd[c], inv);
    if (!(dg > 0.f)) continue;
    const float pc = s.x[c];
    s.x[c] = SV::rnd(pc + (s.r[c] - gb_row(s, c, a, inv)) / dg);
  }
  __syncthreads();
}

// Passes 2 .. 2 SWEEPS of the pre-smoothing (colours 1, 0, 1, ...), pass m on
// the tile grown by 2 SWEEPS - m + 1.
template <int ST, int SWEEPS>
__device__ __forceinline__ void gb_pre_passes(GbLds& s, float inv) {
  if (SWEEPS >= 2) {
    gb_pass<ST, 2 * SWEEPS - 1, true>(s, inv, 1);
    gb_pass<ST, 2 * SWEEPS - 2, true>(s, inv, 0);
  }
  gb_pass<ST, 1, true>(s, inv, 1);
}

// Passes 1 .. 2 SWEEPS of the post-smoothing (colours 1, 0, 1, 0), pass m on
// the tile grown by 2 SWEEPS - m.
template <int ST, int SWEEPS>
__device__ __forceinline__ void gb_post_passes(GbLds& s, float inv) {
  if (SWEEPS >= 2) {
    gb_pass<ST, 2 * SWEEPS - 1, false>(s, inv, 1);
    gb_pass<ST, 2 * SWEEPS - 2, false>(s, inv, 0);
  }
  gb_pass<ST, 1, false>(s, inv, 1);
  gb_pass<ST, 0, false>(s, inv, 0);
}

// Pre-smoothing from zero (SWEEPS forward red-black sweeps, as gc_sweeps
// launches them) and the restriction of the residual to C.r (k_gc_restrict2).
// Writes the smoothed x of the tile's interior to L.x.
template <int ST, int SWEEPS>
__global__ void __launch_bounds__(GB_THREADS) k_gb_pre(Level L, Level C, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  using T = typename SV::T;
  static_assert(SWEEPS >= 1 && SWEEPS <= GB_MAX_SWEEPS, "halo");
  if (!ctl->active) return;
  __shared__ GbLds s;
  const GbTile t = gb_tile(L);
  gb_load<ST, false>(s, L, t, nullptr, C, nullptr);
  gb_pre_passes<ST, SWEEPS>(s, L.inv);
  const int tid = (int)threadIdx.x;
  if (tid < GB_I) {
    const int ix = tid % GB_T, iy = (tid / GB_T) % GB_T, iz = tid / (GB_T * GB_T);
    const int c = ((iz + GB_H) * GB_B + iy + GB_H) * GB_B + ix + GB_H;
    const int i = t.i0 + GB_H + ix, j = t.j0 + GB_H + iy, k = t.k0 + GB_H + iz;
    float a[6];
    gb_faces(s, c, a);
    s.res[tid] = s.r[c] - gb_row(s, c, a, L.inv);
    if (gb_inside(L, i, j, k)) SV::st(reinterpret_cast<T*>(L.x), lidx(L, i, j, k), s.x[c]);
  }
  __syncthreads();
  constexpr int HT = GB_T / 2;
  if (tid >= HT * HT * HT) return;
  const int I = tid % HT, J = (tid / HT) % HT, K = tid / (HT * HT);
  const int ci = (t.i0 + GB_H) / 2 + I, cj = (t.j0 + GB_H) / 2 + J, ck = (t.k0 + GB_H) / 2 + K;
  if (!gb_inside(C, ci, cj, ck)) return;
  float sum = 0.f;
#pragma unroll
  for (int ch = 0; ch < 8; ++ch) {  // children in z, y, x order, x fastest
    const int fx = 2 * I + (ch & 1), fy = 2 * J + ((ch >> 1) & 1), fz = 2 * K + (ch >> 2);
    sum += s.res[(fz * GB_T + fy) * GB_T + fx];
  }
  SV::st(reinterpret_cast<T*>(C.r), lidx(C, ci, cj, ck), 0.125f * sum);
}

// Prolongation of the parents xp (level C) onto x, then SWEEPS backward
// red-black sweeps (colour 1 first, as gc_ascend launches them); the tile's
// interior goes to xout.
template <int ST, int SWEEPS>
__global__ void __launch_bounds__(GB_THREADS) k_gb_post(Level L, Level C, const typename Store<ST>::T* x,
                                                        const typename Store<ST>::T* xp,
                                                        typename Store<ST>::T* xout, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  static_assert(SWEEPS >= 1 && SWEEPS <= GB_MAX_SWEEPS, "halo");
  if (!ctl->active) return;
  __shared__ GbLds s;
  const GbTile t = gb_tile(L);
  gb_load<ST, true>(s, L, t, x, C, xp);
  gb_post_passes<ST, SWEEPS>(s, L.inv);
  const int tid = (int)threadIdx.x;
  if (tid >= GB_I) return;
  const int ix = tid % GB_T, iy = (tid / GB_T) % GB_T, iz = tid / (GB_T * GB_T);
  const int i = t.i0 + GB_H + ix, j = t.j0 + GB_H + iy, k = t.k0 + GB_H + iz;
  if (gb_inside(L, i, j, k))
    SV::st(xout, lidx(L, i, j, k), s.x[((iz + GB_H) * GB_B + iy + GB_H) * GB_B + ix + GB_H]);
}

inline unsigned gb_blocks(const Level& L) {
  return (unsigned)(((L.nx + GB_T - 1) / GB_T) * ((L.ny + GB_T - 1) / GB_T) * ((L.nz + GB_T - 1) / GB_T));
}

}  // namespace fsp

namespace fsp {

// ---------------------------------------------------------------- compact tail
// Every Galerkin level below the compact ones in one workgroup, in LDS:
// per cell the x-, y-, z- couplings in f32 and dd, then x and r (values
// rounded to storage ST): 6 floats. Same arithmetic as k_gc_gs /
// k_gc_restrict2. The couplings of all levels lie first, level after level
// ([fx, fy, fz, dd] blocks of n floats each), exactly as k_gct_pack lays
// them out in global memory, so that the kernel loads them as one
// contiguous copy.
struct GcTailLvl {
  int nx, ny, nz, n;
  float inv;
  float rnx, rny, rhx;  // 1/nx, 1/ny, 1/(nx/2) for the cell decodes
  int o_x, o_r, o_fx, o_fy, o_fz, o_dd;
};

// Float offset of level l's coupling block in the pack (and in LDS).
__host__ __device__ inline size_t gct_pack_offset(const TailSet& T, int l) {
  size_t off = 0;
  for (int m = 0; m < l; ++m) off += 4 * T.lv[m].n;
  return off;
}

// The couplings of the tail levels gathered from their face arrays (run with
// the hierarchy). Blocks are handed out level by level, so a block's level
// is uniform.
__global__ void __launch_bounds__(BLOCK) k_gct_pack(TailSet T, float* pack) {
  int b = (int)blockIdx.x, l = 0;
  size_t off = 0;
  for (; l < T.count; ++l) {
    const int nb = (int)((T.lv[l].n + BLOCK - 1) / BLOCK);
    if (b < nb) break;
    b -= nb;
    off += 4 * T.lv[l].n;
  }
  if (l >= T.count) return;
  const Level& g = T.lv[l];
  const size_t c = (size_t)b * BLOCK + threadIdx.x;
  if (c >= g.n) return;
  int i, j, k;
  ldecode(g, c, i, j, k);
  const size_t n = g.n;
  pack[off + c] = i > 0 ? g.ou[((size_t)k * g.ny + j) * (g.nx + 1) + i] : 0.f;
  pack[off + n + c] = j > 0 ? g.ov[((size_t)k * (g.ny + 1) + j) * g.nx + i] : 0.f;
  pack[off + 2 * n + c] = k > 0 ? g.ow[((size_t)k * g.ny + j) * g.nx + i] : 0.f;
  pack[off + 3 * n + c] = g.dd[c];
}

inline unsigned gct_pack_blocks(const TailSet& T) {
  unsigned b = 0;
  for (int l = 0; l < T.count; ++l) b += (unsigned)((T.lv[l].n + BLOCK - 1) / BLOCK);
  return b;
}

// Cell c -> (i, j, k) without integer division: a/b = floor((a + 1/2)/b),
// exact in f32 for a < 2^22 (the error of the product stays below the 1/(2b)
// margin); tail levels have at most 4096 cells.
__device__ __forceinline__ void gct_decode(const GcTailLvl& L, int c, int& i, int& j, int& k) {
  const int t = (int)(((float)c + 0.5f) * L.rnx);
  i = c - t * L.nx;
  k = (int)(((float)t + 0.5f) * L.rny);
  j = t - k * L.ny;
}

__device__ __forceinline__ void gct_faces(const float* lds, const GcTailLvl& L, int i, int j, int k, int c,
                                          float a[6]) {
  const int plane = L.nx * L.ny;
  a[0] = lds[L.o_fx + c];
  a[1] = i + 1 < L.nx ? lds[L.o_fx + c + 1] : 0.f;
  a[2] = lds[L.o_fy + c];
  a[3] = j + 1 < L.ny ? lds[L.o_fy + c + L.nx] : 0.f;
  a[4] = lds[L.o_fz + c];
  a[5] = k + 1 < L.nz ? lds[L.o_fz + c + plane] : 0.f;
}

__device__ __forceinline__ float gct_row(const float* lds, const GcTailLvl& L, int i, int j, int k, int c,
                                         const float a[6]) {
  const float* x = lds + L.o_x;
  const int plane = L.nx * L.ny;
  const float nb[6] = {i > 0 ? x[c - 1] : 0.f, i + 1 < L.nx ? x[c + 1] : 0.f,
                       j > 0 ? x[c - L.nx] : 0.f, j + 1 < L.ny ? x[c + L.nx] : 0.f,
                       k > 0 ? x[c - plane] : 0.f, k + 1 < L.nz ? x[c + plane] : 0.f};
  return gc_row(a, lds[L.o_dd + c], L.inv, x[c], nb);
}

// One red-black colour on a tail level (first: x starts at zero).
template <int ST>
__device__ __forceinline__ void gct_pass(float* lds, const GcTailLvl& L, int col, bool first) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  float* x = lds + L.o_x;
  const float* r = lds + L.o_r;
  if ((L.nx & 1) == 0) {
    // even rows: every pair (2u, 2u + 1) holds one cell of each colour, so
    // the threads walk the colour's cells only
    const int hx = L.nx >> 1, m = L.n >> 1;
    for (int q = (int)threadIdx.x; q < m; q += TAIL_THREADS) {
      const int row = (int)(((float)q + 0.5f) * L.rhx);
      const int u = q - row * hx;
      const int k = (int)(((float)row + 0.5f) * L.rny);
      const int j = row - k * L.ny;
      const int i = 2 * u + ((col + j + k) & 1);
      const int c = row * L.nx + i;
      if (first) x[c ^ 1] = 0.f;  // the pair's other cell (row * nx is even)
      float a[6];
      gct_faces(lds, L, i, j, k, c, a);
      const float dg = gc_diag(a, lds[L.o_dd + c], L.inv);
      if (first) { x[c] = dg > 0.f ? SV::rnd(r[c] / dg) : 0.f; continue; }
      if (dg > 0.f) x[c] = SV::rnd(x[c] + (r[c] - gct_row(lds, L, i, j, k, c, a)) / dg);
    }
    __syncthreads();
    return;
  }
  for (int c = (int)threadIdx.x; c < L.n; c += TAIL_THREADS) {
    int i, j, k;
    gct_decode(L, c, i, j, k);
    if (((i + j + k) & 1) != col) {
      if (first) x[c] = 0.f;
      continue;
    }
    float a[6];
    gct_faces(lds, L, i, j, k, c, a);
    const float dg = gc_diag(a, lds[L.o_dd + c], L.inv);
    if (first) { x[c] = dg > 0.f ? SV::rnd(r[c] / dg) : 0.f; continue; }
    if (dg > 0.f) x[c] = SV::rnd(x[c] + (r[c] - gct_row(lds, L, i, j, k, c, a)) / dg);
  }
  __syncthreads();
}

// The palindromic sweeps of a coarsest level of at most 64 cells in the
// first wavefront: each lane holds one cell's x in a register and reads its
// neighbours' by lane shuffle, so the passes need no workgroup barrier.
// Per cell the arithmetic of gct_pass, in the same order.
template <int ST>
__device__ __forceinline__ void gct_coarsest_wave(float* lds, const GcTailLvl& L, int count, bool from_zero) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (threadIdx.x < 64) {
    const int c = (int)threadIdx.x;
    const bool in = c < L.n;
    int i = 0, j = 0, k = 0;
    if (in) gct_decode(L, c, i, j, k);
    const int plane = L.nx * L.ny;
    float a[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float dd = 0.f, r = 0.f, x = 0.f;
    if (in) {
      gct_faces(lds, L, i, j, k, c, a);
      dd = lds[L.o_dd + c];
      r = lds[L.o_r + c];
      if (!from_zero) x = lds[L.o_x + c];
    }
    const bool has[6] = {in && i > 0, in && i + 1 < L.nx, in && j > 0, in && j + 1 < L.ny, in && k > 0,
                         in && k + 1 < L.nz};
    const int src[6] = {has[0] ? c - 1 : c, has[1] ? c + 1 : c, has[2] ? c - L.nx : c,
                        has[3] ? c + L.nx : c, has[4] ? c - plane : c, has[5] ? c + plane : c};
    const float dg = gc_diag(a, dd, L.inv);
    const int mine = (i + j + k) & 1;
    auto pass = [&](int col, bool first) {
      float nb[6];
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const float v = __shfl(x, src[f]);
        nb[f] = has[f] ? v : 0.f;
      }
      if (!in) return;
      if (mine != col) {
        if (first) x = 0.f;
        return;
      }
      if (first) { x = dg > 0.f ? SV::rnd(r / dg) : 0.f; return; }
      if (dg > 0.f) x = SV::rnd(x + (r - gc_row(a, dd, L.inv, x, nb)) / dg);
    };
    bool fz = from_zero;
    const int half = count / 2;
    for (int s = 0; s < half; ++s) { pass(0, fz); fz = false; pass(1, false); }
    if (count & 1) { pass(0, fz); fz = false; }
    for (int s = 0; s < half; ++s) { pass(1, fz); fz = false; pass(0, false); }
    if (in) lds[L.o_x + c] = x;
  }
  __syncthreads();
}

template <int ST>
__device__ __forceinline__ void gct_sweeps(float* lds, const GcTailLvl& L, int count, bool forward, bool coarsest,
                                           bool from_zero) {
  if (count == 0) {
    if (from_zero) {
      for (int c = (int)threadIdx.x; c < L.n; c += TAIL_THREADS) lds[L.o_x + c] = 0.f;
      __syncthreads();
    }
    return;
  }
  bool fz = from_zero;
  if (coarsest && L.n <= 64) {
    gct_coarsest_wave<ST>(lds, L, count, from_zero);
    return;
  }
  if (coarsest) {
    const int half = count / 2;
    for (int s = 0; s < half; ++s) { gct_pass<ST>(lds, L, 0, fz); fz = false; gct_pass<ST>(lds, L, 1, false); }
    if (count & 1) { gct_pass<ST>(lds, L, 0, fz); fz = false; }
    for (int s = 0; s < half; ++s) { gct_pass<ST>(lds, L, 1, fz); fz = false; gct_pass<ST>(lds, L, 0, false); }
    return;
  }
  for (int s = 0; s < count; ++s) {
    const int c0 = forward ? 0 : 1;
    gct_pass<ST>(lds, L, c0, fz);
    fz = false;
    gct_pass<ST>(lds, L, 1 - c0, false);
  }
}

// r of level C = restriction of the residual of level L. Eight consecutive
// lanes take the eight children of one coarse cell, one residual each, and
// every lane of the group sums them in child order (z, y, x; x fastest).
// Groups never straddle the loop's end (8 C.n threads), so the lanes read by
// a shuffle are active.
template <int ST>
__device__ __forceinline__ void gct_restrict(float* lds, const GcTailLvl& L, const GcTailLvl& C) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  const int lane = (int)(threadIdx.x & 63u), base = lane & ~7;
  for (int q = (int)threadIdx.x; q < 8 * C.n; q += TAIL_THREADS) {
    const int cc = q >> 3, ch = q & 7;
    int i, j, k;
    gct_decode(C, cc, i, j, k);
    const int fi = 2 * i + (ch & 1), fj = 2 * j + ((ch >> 1) & 1), fk = 2 * k + (ch >> 2);
    const int f = (fk * L.ny + fj) * L.nx + fi;
    float a[6];
    gct_faces(lds, L, fi, fj, fk, f, a);
    const float t = lds[L.o_r + f] - gct_row(lds, L, fi, fj, fk, f, a);
    float s = 0.f;
#pragma unroll
    for (int m = 0; m < 8; ++m) s += __shfl(t, base + m);
    if (ch == 0) lds[C.o_r + cc] = SV::rnd(0.125f * s);
  }
  __syncthreads();
}

// x of level L += x of its parent in C.
template <int ST>
__device__ __forceinline__ void gct_prolong(float* lds, const GcTailLvl& L, const GcTailLvl& C) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  for (int c = (int)threadIdx.x; c < L.n; c += TAIL_THREADS) {
    int i, j, k;
    gct_decode(L, c, i, j, k);
    lds[L.o_x + c] = SV::rnd(lds[L.o_x + c] + lds[C.o_x + ((k >> 1) * C.ny + (j >> 1)) * C.nx + (i >> 1)]);
  }
  __syncthreads();
}

// coarse_cycle3 (delta3d_multigrid.rs:303-332) over the tail levels, red-black
// smoothing, Galerkin levels; the first level's r is restricted, its x is
// written back.
template <int ST>
__global__ void __launch_bounds__(TAIL_THREADS) k_gc_tail(TailSet T, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (!ctl->active) return;
  extern __shared__ __attribute__((aligned(16))) float lds[];
  __shared__ GcTailLvl tl[TAIL_MAX_LEVELS];
  int total = 0;  // cells of all tail levels
  for (int l = 0; l < T.count; ++l) total += (int)T.lv[l].n;
  if (threadIdx.x == 0) {
    int off = 0, xr = 4 * total;
    for (int l = 0; l < T.count; ++l) {
      const Level& g = T.lv[l];
      GcTailLvl& L = tl[l];
      L.nx = g.nx; L.ny = g.ny; L.nz = g.nz; L.n = (int)g.n; L.inv = g.inv;
      L.rnx = 1.0f / (float)g.nx; L.rny = 1.0f / (float)g.ny; L.rhx = 1.0f / (float)(g.nx >> 1 > 0 ? g.nx >> 1 : 1);
      L.o_fx = off; off += L.n;
      L.o_fy = off; off += L.n;
      L.o_fz = off; off += L.n;
      L.o_dd = off; off += L.n;
      L.o_x = xr; xr += L.n;
      L.o_r = xr; xr += L.n;
    }
  }
  // the couplings of every level: one contiguous copy of 4 * total floats
  // (a multiple of four, 16-byte aligned at both ends), one round of loads
  {
    const float4* src = reinterpret_cast<const float4*>(T.pack);
    float4* dst = reinterpret_cast<float4*>(lds);
    constexpr int MAXQ = (TAIL_LDS_FLOATS / 6 + TAIL_THREADS - 1) / TAIL_THREADS;  // float4s per thread
    float4 v[MAXQ];
#pragma unroll
    for (int u = 0; u < MAXQ; ++u) {
      const int q = (int)threadIdx.x + u * TAIL_THREADS;
      v[u] = q < total ? src[q] : f4zero();
    }
    const int c0 = (int)threadIdx.x;
    const float r0 = c0 < (int)T.lv[0].n ? SV::ld(reinterpret_cast<const typename SV::T*>(T.lv[0].r), c0) : 0.f;
    const float r1 = c0 + TAIL_THREADS < (int)T.lv[0].n
                         ? SV::ld(reinterpret_cast<const typename SV::T*>(T.lv[0].r), c0 + TAIL_THREADS) : 0.f;
    const float r2 = c0 + 2 * TAIL_THREADS < (int)T.lv[0].n
                         ? SV::ld(reinterpret_cast<const typename SV::T*>(T.lv[0].r), c0 + 2 * TAIL_THREADS) : 0.f;
#pragma unroll
    for (int u = 0; u < MAXQ; ++u) {
      const int q = (int)threadIdx.x + u * TAIL_THREADS;
      if (q < total) dst[q] = v[u];
    }
    __syncthreads();  // tl
    const int orr = tl[0].o_r, n0 = tl[0].n;
    if (c0 < n0) lds[orr + c0] = r0;
    if (c0 + TAIL_THREADS < n0) lds[orr + c0 + TAIL_THREADS] = r1;
    if (c0 + 2 * TAIL_THREADS < n0) lds[orr + c0 + 2 * TAIL_THREADS] = r2;
  }
  __syncthreads();
  const int last = T.count - 1;
  for (int l = 0; l <= last; ++l) {
    // level descriptors in registers: the passes do not reload them after each barrier
    const GcTailLvl L = tl[l];
    gct_sweeps<ST>(lds, L, l == last ? T.coarse : T.pre, true, l == last, true);
    if (l < last) {
      const GcTailLvl C = tl[l + 1];
      gct_restrict<ST>(lds, L, C);
    }
  }
  for (int l = last - 1; l >= 0; --l) {
    const GcTailLvl L = tl[l], C = tl[l + 1];
    gct_prolong<ST>(lds, L, C);
    gct_sweeps<ST>(lds, L, T.post, false, false, false);
  }
  const GcTailLvl L0 = tl[0];
  for (int q = threadIdx.x; q < L0.n; q += TAIL_THREADS)
    SV::st(reinterpret_cast<typename SV::T*>(T.lv[0].x), q, lds[L0.o_x + q]);
}

inline size_t gc_tail_lds_bytes(const Level* lv, int count) {
  size_t floats = 0;
  for (int l = 0; l < count; ++l) floats += 6 * lv[l].n;
  return floats * sizeof(float);
}
static_assert(TAIL_LDS_FLOATS / 6 <= 3 * TAIL_THREADS, "the first tail level's r in three loads per thread");

}  // namespace fsp
