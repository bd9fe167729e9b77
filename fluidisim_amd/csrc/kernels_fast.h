// Fast path of the fine level and of the small coarse levels.
//
// Fine level ("quad" kernels, nx % 4 == 0): one thread owns four consecutive
// cells along x, so every lane of a wavefront is busy in both colours of a
// red-black pass and every access is a 16-byte vector access; x neighbours
// across quads come from the adjacent lane by wavefront shuffle, y and z
// neighbours are vector loads served mostly by L2. The 1-D grid is mapped to
// (x-segment, y-block, z) tiles in dispatch order (see xcd_block). Per-cell arithmetic
// is the fine operator's (face order x-, x+, y-, y+, z-, z+, no contraction),
// so results equal the generic kernels bit for bit.
//
// Small coarse levels: one workgroup runs every level below a size threshold
// with the vectors and coefficients in LDS, replacing dozens of tiny launches
// per V-cycle by one.
#pragma once
#include "kernels.h"

namespace fsp {

constexpr int QBLOCK = 256;
// Minimum workgroups per CU the fine red-black kernel is compiled for (the
// register budget follows from it).
#ifndef FSP_QGS_MIN_BLOCKS
#define FSP_QGS_MIN_BLOCKS 1
#endif
// Widest x extent of a workgroup tile, in quads (tile shapes for tuning).
#ifndef FSP_QX_MAX
#define FSP_QX_MAX 64
#endif

struct QuadGrid {
  int nq;   // quads per row (nx / 4)
  int QX;   // threads along x per block (power of two, <= 64)
  int RY;   // rows per block (QBLOCK / QX)
  int nbx;  // blocks along x
  int nby;  // blocks along y (over `rows`)
  int rows; // rows covered (ny, or ny/2 for the restriction)
  int planes;
};

// Tile coordinates of this thread; returns false for lanes past the edge
// (they still take part in shuffles).
__device__ __forceinline__ bool quad_coords(const QuadGrid& Q, int& qx, int& y, int& z) {
  const unsigned G = gridDim.x, b = blockIdx.x;
  const unsigned L = xcd_block(b, G);
  const unsigned bx = L % (unsigned)Q.nbx;
  const unsigned t = L / (unsigned)Q.nbx;
  const unsigned by = t % (unsigned)Q.nby;
  z = (int)(t / (unsigned)Q.nby);
  qx = (int)bx * Q.QX + (int)(threadIdx.x & (unsigned)(Q.QX - 1));
  y = (int)by * Q.RY + (int)(threadIdx.x / (unsigned)Q.QX);
  return qx < Q.nq && y < Q.rows && z < Q.planes;
}

__device__ __forceinline__ float q_get(const float4& v, int e) { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }
__device__ __forceinline__ void q_set(float4& v, int e, float x) {
  if (e == 0) v.x = x; else if (e == 1) v.y = x; else if (e == 2) v.z = x; else v.w = x;
}
__device__ __forceinline__ float4 ld4(const float* p, size_t c) { return *reinterpret_cast<const float4*>(p + c); }
__device__ __forceinline__ void st4(float* p, size_t c, const float4& v) { *reinterpret_cast<float4*>(p + c) = v; }
__device__ __forceinline__ float4 f4zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }

// The sum of a free-surface row (about one row in a hundred); GEN as ghost_row.
template <bool GEN = true>
__device__ __forceinline__ float ghost_row_sum(const Fine& G, int i, int j, int k, uint8_t m, float pc, const float nb[6],
                                            float& prec) {
#pragma clang fp contract(off)
  uint8_t type[6];
  float it[6], a[6];
  ghost_row<GEN>(G, i, j, k, m, type, it, a);
  float acc = 0.f, diag = 0.f;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    if (type[f] == 1) { acc += a[f] * (pc - nb[f]); diag += a[f]; }
    else if (type[f] == 2) { acc += a[f] * pc * it[f]; diag += a[f] * it[f]; }
  }
  prec = diag > 0.f ? 1.0f / (diag * G.inv) : 0.f;
  return acc;
}

// Row of the fine operator from neighbour values already in registers, in the
// reference's face order; prec as rhs_mobile3 computes it (GEN as ghost_row).
template <bool GEN = true>
__device__ __forceinline__ float row_local(const Fine& G, int i, int j, int k, uint8_t m, float pc, const float nb[6],
                                          float& prec) {
#pragma clang fp contract(off)
  prec = 0.f;
  if (!(m & M_WET)) return 0.f;
  float acc = 0.f;
  if (!(m & M_GHOST)) {
#pragma unroll
    for (int f = 0; f < 6; ++f)
      if (m & (1u << f)) acc += pc - nb[f];
    prec = G.prec_tab[__popc(m & 0x3Fu)];
  } else {
    acc = ghost_row_sum<GEN>(G, i, j, k, m, pc, nb, prec);
  }
  return acc * G.inv;
}

// The four cells' x neighbours: within the quad, from the adjacent lanes, or
// (at the edge of a lane segment) from memory through `edge`.
// Every lane takes part in the shuffles; only valid lanes read memory.
template <class Edge>
__device__ __forceinline__ void x_neighbours(const QuadGrid& Q, bool valid, int qx, const float4& v, float& left,
                                             float& right, Edge edge) {
  const float l_sh = __shfl_up(v.w, 1, 64);
  const float r_sh = __shfl_down(v.x, 1, 64);
  const int lane_in_seg = (int)(threadIdx.x & (unsigned)(Q.QX - 1));
  left = 0.f;
  right = 0.f;
  if (!valid) return;
  if (qx > 0) left = lane_in_seg > 0 ? l_sh : edge(-1);
  if (qx + 1 < Q.nq) right = (lane_in_seg < Q.QX - 1) ? r_sh : edge(4);
}

// ---------------------------------------------------------------- red-black pass
// One colour of a red-black Gauss-Seidel sweep on quads, in place (the first
// pass from z = 0 is fused with the second in k_q_gs_first).
// MODE 2: z is first prolonged from the first coarse level (z += x_c of the
//         parent, 0 on rows with prec == 0), then the colour is updated;
//         neighbours are prolonged on the fly, so this mode reads z and
//         writes zout (a different buffer).
// dot_op >= 0: also reduce <r, z> over every cell into that slot.
// z, zout and the parents xc are held in storage ST (r is the f32 CG
// residual); every value entering the arithmetic is a stored value.
template <int MODE, int ST, bool GEN = true>
__global__ void __launch_bounds__(QBLOCK, FSP_QGS_MIN_BLOCKS) k_q_gs(Fine G, QuadGrid Q, const typename Store<ST>::T* z,
                                                 typename Store<ST>::T* zout, const float* r, int colour, Level L0,
                                                 const typename Store<ST>::T* xc, int dot_op, Ctl* ctl,
                                                 double* partials) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (!ctl->active) return;
  int qx, y, k;
  const bool valid = quad_coords(Q, qx, y, k);
  const size_t plane = (size_t)G.nx * G.ny;
  const size_t c0 = valid ? ((size_t)k * G.ny + y) * G.nx + 4 * (size_t)qx : 0;
  double s = 0.0, unused = 0.0;
  static_assert(MODE == 0 || MODE == 2, "in place, or prolongation then update");
  // parent values of the prolongation
  auto parent = [&](int yy, int kk, int xq) -> float2 { return SV::ld2(xc, lidx(L0, 2 * xq, yy >> 1, kk >> 1)); };
  // prolongation z + parent, rounded to storage (prolong_add3 stores it)
  auto prolong = [&](float4& v, float2 a) {
    v.x = SV::rnd(v.x + a.x); v.y = SV::rnd(v.y + a.x); v.z = SV::rnd(v.z + a.y); v.w = SV::rnd(v.w + a.y);
  };
  float4 zc = f4zero(), rq = f4zero();
  uint32_t mq = 0;
  float4 ym = f4zero(), yp = f4zero(), zm = f4zero(), zp = f4zero();
  if (valid) {
    mq = *reinterpret_cast<const uint32_t*>(G.mask + c0);
    zc = SV::ld4(z, c0);
    rq = ld4(r, c0);
    if (y > 0) ym = SV::ld4(z, c0 - G.nx);
    if (y + 1 < G.ny) yp = SV::ld4(z, c0 + G.nx);
    if (k > 0) zm = SV::ld4(z, c0 - plane);
    if (k + 1 < G.nz) zp = SV::ld4(z, c0 + plane);
    if (MODE == 2) {
      prolong(zc, parent(y, k, qx));
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (!prec_positive((uint8_t)(mq >> (8 * e)))) q_set(zc, e, 0.f);
      if (y > 0) prolong(ym, parent(y - 1, k, qx));
      if (y + 1 < G.ny) prolong(yp, parent(y + 1, k, qx));
      if (k > 0) prolong(zm, parent(y, k - 1, qx));
      if (k + 1 < G.nz) prolong(zp, parent(y, k + 1, qx));
    }
  }
  float left, right;
  x_neighbours(Q, valid, qx, zc, left, right, [&](int off) {
    float v = SV::ld(z, c0 + off);
    if (MODE == 2) {
      const int xx = 4 * qx + off;
      v = SV::rnd(v + SV::ld(xc, lidx(L0, xx >> 1, y >> 1, k >> 1)));
    }
    return v;
  });
  if (valid) {
    float4 out = zc;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (((e + y + k) & 1) != colour) continue;
      const uint8_t m = (uint8_t)(mq >> (8 * e));
      const float pc = q_get(zc, e);
      const float xl = e == 0 ? left : q_get(zc, e - 1);
      const float xr = e == 3 ? right : q_get(zc, e + 1);
      const float nb[6] = {xl, xr, q_get(ym, e), q_get(yp, e), q_get(zm, e), q_get(zp, e)};
      float prec;
      const float az = row_local<GEN>(G, 4 * qx + e, y, k, m, pc, nb, prec);
      if (prec > 0.f) q_set(out, e, SV::rnd(pc + prec * (q_get(rq, e) - az)));
    }
    SV::st4(zout, c0, out);
    if (dot_op >= 0) s = (double)rq.x * out.x + (double)rq.y * out.y + (double)rq.z * out.z + (double)rq.w * out.w;
  }
  if (dot_op >= 0) reduce_finish<false>(s, unused, dot_op, ctl, partials);
}

// The first two passes of the pre-smoothing from z = 0 in one launch: the
// colour-0 pass from zero is local (colour-0 cells row_prec * r, colour-1
// cells 0), so the colour-1 pass recomputes those values for its neighbours
// from r and the per-cell row_prec (k_prec) instead of reading them back.
// Every value is the one the two single passes store, so the result is
// theirs bit for bit, for one pass's traffic plus the row_prec reads (a
// variant deriving plain rows' factor from the row code ran slower).
template <int ST, bool GEN = true>
__global__ void __launch_bounds__(QBLOCK, FSP_QGS_MIN_BLOCKS) k_q_gs_first(Fine G, QuadGrid Q,
                                                                         typename Store<ST>::T* zout, const float* r,
                                                                         Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (!ctl->active) return;
  int qx, y, k;
  const bool valid = quad_coords(Q, qx, y, k);
  const size_t plane = (size_t)G.nx * G.ny;
  const size_t c0 = valid ? ((size_t)k * G.ny + y) * G.nx + 4 * (size_t)qx : 0;
  // the first pass's value of cell (x, yy, kk): row_prec * r on colour 0
  auto first = [&](int x, int yy, int kk, float prec, float rv) -> float {
    if (((x + yy + kk) & 1) != 0) return 0.f;
    return prec > 0.f ? SV::rnd(prec * rv) : 0.f;
  };
  auto first4 = [&](const float4& pq, const float4& rq4, int yy, int kk) {
    float4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) q_set(v, e, first(4 * qx + e, yy, kk, q_get(pq, e), q_get(rq4, e)));
    return v;
  };
  float4 rq = f4zero(), zc = f4zero(), ym = f4zero(), yp = f4zero(), zm = f4zero(), zp = f4zero();
  uint32_t mq = 0;
  if (valid) {
    // every load first, then the values
    mq = *reinterpret_cast<const uint32_t*>(G.mask + c0);
    rq = ld4(r, c0);
    const float4 pc = ld4(G.precv, c0);
    const bool hym = y > 0, hyp = y + 1 < G.ny, hzm = k > 0, hzp = k + 1 < G.nz;
    const size_t cym = c0 - G.nx, cyp = c0 + G.nx, czm = c0 - plane, czp = c0 + plane;
    const float4 r_ym = hym ? ld4(r, cym) : f4zero(), p_ym = hym ? ld4(G.precv, cym) : f4zero();
    const float4 r_yp = hyp ? ld4(r, cyp) : f4zero(), p_yp = hyp ? ld4(G.precv, cyp) : f4zero();
    const float4 r_zm = hzm ? ld4(r, czm) : f4zero(), p_zm = hzm ? ld4(G.precv, czm) : f4zero();
    const float4 r_zp = hzp ? ld4(r, czp) : f4zero(), p_zp = hzp ? ld4(G.precv, czp) : f4zero();
    zc = first4(pc, rq, y, k);
    if (hym) ym = first4(p_ym, r_ym, y - 1, k);
    if (hyp) yp = first4(p_yp, r_yp, y + 1, k);
    if (hzm) zm = first4(p_zm, r_zm, y, k - 1);
    if (hzp) zp = first4(p_zp, r_zp, y, k + 1);
  }
  float left, right;
  x_neighbours(Q, valid, qx, zc, left, right, [&](int off) {
    const size_t c = c0 + off;
    return first(4 * qx + off, y, k, G.precv[c], r[c]);
  });
  if (!valid) return;
  float4 out = zc;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (((e + y + k) & 1) != 1) continue;
    const uint8_t m = (uint8_t)(mq >> (8 * e));
    const float pc = q_get(zc, e);
    const float xl = e == 0 ? left : q_get(zc, e - 1);
    const float xr = e == 3 ? right : q_get(zc, e + 1);
    const float nb[6] = {xl, xr, q_get(ym, e), q_get(yp, e), q_get(zm, e), q_get(zp, e)};
    float prec;
    const float az = row_local<GEN>(G, 4 * qx + e, y, k, m, pc, nb, prec);
    if (prec > 0.f) q_set(out, e, SV::rnd(pc + prec * (q_get(rq, e) - az)));
  }
  SV::st4(zout, c0, out);
}

// ---------------------------------------------------------------- fused red-black sweep on bricks
// Both colour passes of a red-black sweep in one launch, one brick per
// workgroup: z of the brick (BR_QX quads x BR_Y rows x BR_Z planes) and a
// halo of one quad in x and two cells in y and z is loaded into LDS at once;
// pass c runs over the brick and a ring of one cell (its neighbours are
// colour 1 - c values, untouched by it) back into LDS; after a barrier pass
// 1 - c runs over the brick (its neighbours are the colour-c values, all
// updated) and is stored to zout. Bricks are independent (zout must not
// alias zin), so many of them per CU overlap their loads with each other's
// arithmetic. Every cell's update uses the operands the single-colour passes
// use, in the same order: the sweep equals k_q_gs twice bit for bit.
// MODE 2: every value read from zin is first prolonged from the first coarse
// level (z + parent, rounded to storage), rows with prec == 0 become 0
// (prolong_add3 then the dry reset), as k_q_gs<2> does before its pass.
// dot_op >= 0: <r, z> of the result over every cell into that slot.
// Traffic per cell: mask, z, r in, z out (13 B in f32 storage), against
// twice that for two single passes.
#ifndef FSP_BRICK_XCD
#define FSP_BRICK_XCD 1
#endif
#ifndef FSP_BRICK_QX
#define FSP_BRICK_QX 8
#endif
#ifndef FSP_BRICK_Y
#define FSP_BRICK_Y 8
#endif
#ifndef FSP_BRICK_Z
#define FSP_BRICK_Z 8
#endif
constexpr int BR_QX = FSP_BRICK_QX, BR_Y = FSP_BRICK_Y, BR_Z = FSP_BRICK_Z;
#ifndef FSP_BRICK_THREADS
#define FSP_BRICK_THREADS 512
#endif
constexpr int BR_THREADS = FSP_BRICK_THREADS;
constexpr int BR_LQ = BR_QX + 2;                   // quads per LDS row
constexpr int BR_LX = 4 * BR_LQ;                   // floats per LDS row
constexpr int BR_LY = BR_Y + 4, BR_LZ = BR_Z + 4;  // LDS rows, planes
constexpr int BR_PAD = 4;                          // floats before and after (halo cells' unused reads)
constexpr int BR_FLOATS = BR_LX * BR_LY * BR_LZ + 2 * BR_PAD;
constexpr int BR_NLOAD = BR_LQ * BR_LY * BR_LZ;
constexpr int BR_LOADS = (BR_NLOAD + BR_THREADS - 1) / BR_THREADS;
constexpr int BR_NA = BR_LQ * (BR_Y + 2) * (BR_Z + 2);  // first-pass quads
constexpr int BR_AITEMS = (BR_NA + BR_THREADS - 1) / BR_THREADS;
constexpr int BR_NB = BR_QX * BR_Y * BR_Z;  // second-pass quads
constexpr int BR_BITEMS = (BR_NB + BR_THREADS - 1) / BR_THREADS;

struct BrickGrid {
  int bx, by, bz;  // bricks along x, y, z
};


template <int MODE, int ST, bool GEN = true>
__global__ void __launch_bounds__(BR_THREADS) k_q_brick(Fine G, BrickGrid B, const typename Store<ST>::T* zin,
                                                        typename Store<ST>::T* zout, const float* r, int colour,
                                                        Level L0, const typename Store<ST>::T* xc, int dot_op,
                                                        Ctl* ctl, double* partials) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  static_assert(MODE == 0 || MODE == 2, "in place, or prolongation then the sweep");
  __shared__ __attribute__((aligned(16))) float zs_all[BR_FLOATS];
  __shared__ float s_prec[8];
  __shared__ double s_red[BR_THREADS / 64];
  float* zs = zs_all + BR_PAD;
  if (!ctl->active) return;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const int tid = (int)threadIdx.x;
  // Bricks in x, then z, then y order, each XCD taking a contiguous range of
  // them (dispatch deals workgroups to the 8 XCDs in turn): bricks that share
  // halo planes and rows run on one XCD close in time, so their halos come
  // from that XCD's L2.
  const unsigned G8 = gridDim.x / 8u;
  const unsigned bl = (FSP_BRICK_XCD && gridDim.x % 8u == 0u) ? (blockIdx.x % 8u) * G8 + blockIdx.x / 8u : blockIdx.x;
  const int bxi = (int)(bl % (unsigned)B.bx), t = (int)(bl / (unsigned)B.bx);
  const int bzi = t % B.bz, byi = t / B.bz;
  const int qx0 = bxi * BR_QX - 1, y0 = byi * BR_Y - 2, z0 = bzi * BR_Z - 2;  // LDS origin (quad, row, plane)
  const int nq = nx / 4;
  if (tid < 8) s_prec[tid] = tid < 7 ? G.prec_tab[tid] : 0.f;
  // every load of the workgroup in flight first: z of the brick and its halo
  // (prolonged in MODE 2), then r and the row codes of both passes' quads
  float4 rqa[BR_AITEMS], rqb[BR_BITEMS];
  uint32_t mqa[BR_AITEMS], mqb[BR_BITEMS];
  bool ona[BR_AITEMS], onb[BR_BITEMS];
  {
    float4 v[BR_LOADS];
#pragma unroll
    for (int m = 0; m < BR_LOADS; ++m) {
      const int idx = tid + BR_THREADS * m;
      const int lq = idx % BR_LQ, rest = idx / BR_LQ, ly = rest % BR_LY, lz = rest / BR_LY;
      const int q = qx0 + lq, y = y0 + ly, k = z0 + lz;
      v[m] = f4zero();
      if (idx < BR_NLOAD && q >= 0 && q < nq && y >= 0 && y < ny && k >= 0 && k < nz) {
        const size_t c = ((size_t)k * ny + y) * nx + 4 * (size_t)q;
        v[m] = SV::ld4(zin, c);
        if (MODE == 2) {
          const float2 a = SV::ld2(xc, lidx(L0, 2 * q, y >> 1, k >> 1));
          v[m].x = SV::rnd(v[m].x + a.x); v[m].y = SV::rnd(v[m].y + a.x);
          v[m].z = SV::rnd(v[m].z + a.y); v[m].w = SV::rnd(v[m].w + a.y);
        }
      }
    }
#pragma unroll
    for (int m = 0; m < BR_AITEMS; ++m) {
      const int idx = tid + BR_THREADS * m;
      const int lq = idx % BR_LQ, rest = idx / BR_LQ, ly = 1 + rest % (BR_Y + 2), lz = 1 + rest / (BR_Y + 2);
      const int q = qx0 + lq, y = y0 + ly, k = z0 + lz;
      ona[m] = idx < BR_NA && q >= 0 && q < nq && y >= 0 && y < ny && k >= 0 && k < nz;
      rqa[m] = f4zero();
      mqa[m] = 0;
      if (ona[m]) {
        const size_t c = ((size_t)k * ny + y) * nx + 4 * (size_t)q;
        rqa[m] = ld4(r, c);
        mqa[m] = *reinterpret_cast<const uint32_t*>(G.mask + c);
      }
    }
#pragma unroll
    for (int m = 0; m < BR_BITEMS; ++m) {
      const int idx = tid + BR_THREADS * m;
      const int lq = 1 + idx % BR_QX, rest = idx / BR_QX, ly = 2 + rest % BR_Y, lz = 2 + rest / BR_Y;
      const int q = qx0 + lq, y = y0 + ly, k = z0 + lz;
      onb[m] = idx < BR_NB && q < nq && y < ny && k < nz;
      rqb[m] = f4zero();
      mqb[m] = 0;
      if (onb[m]) {
        const size_t c = ((size_t)k * ny + y) * nx + 4 * (size_t)q;
        rqb[m] = ld4(r, c);
        mqb[m] = *reinterpret_cast<const uint32_t*>(G.mask + c);
      }
    }
#pragma unroll
    for (int m = 0; m < BR_LOADS; ++m) {
      const int idx = tid + BR_THREADS * m;
      if (idx < BR_NLOAD) *reinterpret_cast<float4*>(zs + 4 * idx) = v[m];
    }
  }
  // One colour on the quad at LDS (lq, ly, lz): the colour's two cells
  // (x-cells par and par + 2) without branches on plain rows; free-surface
  // cells are left as they are and flagged in `late` (bits e) for
  // relax_late. Returns the quad with the colour's plain cells updated.
  auto relax = [&](int lq, int ly, int lz, int col, const float4 rq, uint32_t mq, uint32_t& late) -> float4 {
    const int q = qx0 + lq, y = y0 + ly, k = z0 + lz;
    const float* sc = zs + ((lz * BR_LY + ly) * BR_LX + 4 * lq);
    const float4 zc = *reinterpret_cast<const float4*>(sc);
    const float4 ym = *reinterpret_cast<const float4*>(sc - BR_LX);
    const float4 yp = *reinterpret_cast<const float4*>(sc + BR_LX);
    const float4 zm = *reinterpret_cast<const float4*>(sc - BR_LX * BR_LY);
    const float4 zp = *reinterpret_cast<const float4*>(sc + BR_LX * BR_LY);
    // the x neighbours' quads as whole 16-byte reads (scalar reads 16 bytes
    // apart would conflict in the LDS banks)
    const float4 xl4 = *reinterpret_cast<const float4*>(sc - 4);
    const float4 xr4 = *reinterpret_cast<const float4*>(sc + 4);
    const float left = q > 0 ? xl4.w : 0.f, right = q + 1 < nq ? xr4.x : 0.f;
    const bool odd = ((col + y + k) & 1) != 0;
    auto sel = [odd](float o, float e) -> float { return odd ? o : e; };
    auto cell = [&](float pc, float xl, float xr, float vym, float vyp, float vzm, float vzp, float rc,
                    uint32_t mc) -> float {
      const float nb[6] = {xl, xr, vym, vyp, vzm, vzp};
      float acc = 0.f;
#pragma unroll
      for (int f = 0; f < 6; ++f) acc = (mc & (1u << f)) ? acc + (pc - nb[f]) : acc;
      const float prec = s_prec[(mc & M_WET) ? __popc(mc & 0x3Fu) : 0];
      const float v = prec > 0.f ? SV::rnd(pc + prec * (rc - acc * G.inv)) : (MODE == 2 ? 0.f : pc);
      return (mc & M_GHOST) ? pc : v;
    };
    const uint32_t ma = (mq >> (odd ? 8 : 0)) & 0xFFu, mb = (mq >> (odd ? 24 : 16)) & 0xFFu;
    const float a = cell(sel(zc.y, zc.x), sel(zc.x, left), sel(zc.z, zc.y), sel(ym.y, ym.x), sel(yp.y, yp.x),
                         sel(zm.y, zm.x), sel(zp.y, zp.x), sel(rq.y, rq.x), ma);
    const float b = cell(sel(zc.w, zc.z), sel(zc.z, zc.y), sel(right, zc.w), sel(ym.w, ym.z), sel(yp.w, yp.z),
                         sel(zm.w, zm.z), sel(zp.w, zp.z), sel(rq.w, rq.z), mb);
    late = ((ma & M_GHOST) ? (odd ? 2u : 1u) : 0u) | ((mb & M_GHOST) ? (odd ? 8u : 4u) : 0u);
    return make_float4(sel(zc.x, a), sel(a, zc.y), sel(zc.z, b), sel(b, zc.w));
  };
  // A free-surface cell e of that quad: row_local's ghost form (per-column
  // table, else the geometry), neighbours from LDS.
  auto relax_late = [&](int lq, int ly, int lz, int e) -> float {
    const int q = qx0 + lq, y = y0 + ly, k = z0 + lz;
    const size_t c = ((size_t)k * ny + y) * nx + 4 * (size_t)q + e;
    const float* sc = zs + ((lz * BR_LY + ly) * BR_LX + 4 * lq + e);
    const float nb[6] = {(q > 0 || e > 0) ? sc[-1] : 0.f, (q + 1 < nq || e < 3) ? sc[1] : 0.f, sc[-BR_LX],
                         sc[BR_LX], sc[-BR_LX * BR_LY], sc[BR_LX * BR_LY]};
    const float pc = sc[0];
    float prec;
    const float az = row_local<GEN>(G, 4 * q + e, y, k, G.mask[c], pc, nb, prec);
    return prec > 0.f ? SV::rnd(pc + prec * (r[c] - az)) : (MODE == 2 ? 0.f : pc);
  };
  __syncthreads();
  // pass c over the brick and its ring (LDS rows 1..BR_Y+2, planes 1..BR_Z+2, every LDS quad)
  {
    uint32_t late = 0;
#pragma unroll
    for (int m = 0; m < BR_AITEMS; ++m) {
      if (!ona[m]) continue;
      const int idx = tid + BR_THREADS * m;
      const int lq = idx % BR_LQ, rest = idx / BR_LQ, ly = 1 + rest % (BR_Y + 2), lz = 1 + rest / (BR_Y + 2);
      uint32_t lt;
      const float4 o = relax(lq, ly, lz, colour, rqa[m], mqa[m], lt);
      *reinterpret_cast<float4*>(zs + ((lz * BR_LY + ly) * BR_LX + 4 * lq)) = o;
      late |= lt << (4 * m);
    }
    for (; late; late &= late - 1) {
      const int bit = __ffs(late) - 1, m = bit >> 2, e = bit & 3;
      const int idx = tid + BR_THREADS * m;
      const int lq = idx % BR_LQ, rest = idx / BR_LQ, ly = 1 + rest % (BR_Y + 2), lz = 1 + rest / (BR_Y + 2);
      zs[(lz * BR_LY + ly) * BR_LX + 4 * lq + e] = relax_late(lq, ly, lz, e);
    }
  }
  __syncthreads();
  // pass 1 - c over the brick (LDS quads 1..BR_QX, rows 2..BR_Y+1, planes 2..BR_Z+1), to zout:
  // the plain cells of every quad, then the free-surface cells (their
  // neighbours are colour c, final in LDS, and their own value before the
  // pass is still there), then the stores. Kept apart so that the rare
  // free-surface path is not scheduled into the items' registers.
  double s = 0.0;
  {
    float4 ob[BR_BITEMS];
    uint32_t late = 0;
#pragma unroll
    for (int m = 0; m < BR_BITEMS; ++m) {
      ob[m] = f4zero();
      if (!onb[m]) continue;
      const int idx = tid + BR_THREADS * m;
      const int lq = 1 + idx % BR_QX, rest = idx / BR_QX, ly = 2 + rest % BR_Y, lz = 2 + rest / BR_Y;
      uint32_t lt;
      ob[m] = relax(lq, ly, lz, 1 - colour, rqb[m], mqb[m], lt);
      late |= lt << (4 * m);
    }
    for (; late; late &= late - 1) {
      const int bit = __ffs(late) - 1, m = bit >> 2, e = bit & 3;
      const int idx = tid + BR_THREADS * m;
      const int lq = 1 + idx % BR_QX, rest = idx / BR_QX, ly = 2 + rest % BR_Y, lz = 2 + rest / BR_Y;
      const float x = relax_late(lq, ly, lz, e);
#pragma unroll
      for (int mm = 0; mm < BR_BITEMS; ++mm)
        if (mm == m) q_set(ob[mm], e, x);
    }
#pragma unroll
    for (int m = 0; m < BR_BITEMS; ++m) {
      if (!onb[m]) continue;
      const int idx = tid + BR_THREADS * m;
      const int lq = 1 + idx % BR_QX, rest = idx / BR_QX, ly = 2 + rest % BR_Y, lz = 2 + rest / BR_Y;
      const int q = qx0 + lq, y = y0 + ly, k = z0 + lz;
      const size_t c = ((size_t)k * ny + y) * nx + 4 * (size_t)q;
      SV::st4(zout, c, ob[m]);
      if (dot_op >= 0)
        s += (double)rqb[m].x * ob[m].x + (double)rqb[m].y * ob[m].y + (double)rqb[m].z * ob[m].z +
             (double)rqb[m].w * ob[m].w;
    }
  }
  if (dot_op >= 0) {
    const int lane = tid & 63, wid = tid >> 6;
    s = wave_sum(s);
    if (lane == 0) s_red[wid] = s;
    __syncthreads();
    if (tid == 0) {
      double tt = s_red[0];
      for (int w = 1; w < BR_THREADS / 64; ++w) tt += s_red[w];
      double* part = partials + (size_t)dot_op * MAX_PARTIALS * 2;
      part[2 * blockIdx.x] = tt;
      part[2 * blockIdx.x + 1] = 0.0;
    }
  }
}

// Row codes on quads (k_mask with four x-cells per thread: one decode, and
// the solid flags, apertures and surface heights of the quad and of its y and
// z neighbours as 4-byte and 16-byte loads). Same row_code per cell.
__global__ void __launch_bounds__(QBLOCK) k_mask_q(Fine G, QuadGrid Q, uint8_t* mask, Ctl* ctl, double* partials) {
  double wet = 0.0, ghosts = 0.0;
  int qx, j, k;
  const bool valid = quad_coords(Q, qx, j, k);
  if (valid) {
    const int nx = G.nx, ny = G.ny, nz = G.nz, i0 = 4 * qx;
    const int jm = j > 0 ? j - 1 : j, jp = j + 1 < ny ? j + 1 : j;
    const int km = k > 0 ? k - 1 : k, kp = k + 1 < nz ? k + 1 : k;
    const int im0 = i0 > 0 ? i0 - 1 : i0, ip3 = i0 + 4 < nx ? i0 + 4 : i0 + 3;
    auto u32 = [](const uint8_t* p, size_t c) { return *reinterpret_cast<const uint32_t*>(p + c); };
    const uint32_t sc = u32(G.solid, cidx(G, i0, j, k));
    const uint32_t sym = u32(G.solid, cidx(G, i0, jm, k)), syp = u32(G.solid, cidx(G, i0, jp, k));
    const uint32_t szm = u32(G.solid, cidx(G, i0, j, km)), szp = u32(G.solid, cidx(G, i0, j, kp));
    const uint8_t sxm = G.solid[cidx(G, im0, j, k)], sxp = G.solid[cidx(G, ip3, j, k)];
    const size_t fu = fuidx(G, i0, j, k);
    uint8_t au[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) au[e] = G.ou[fu + e];
    const uint32_t av0 = u32(G.ov, fvidx(G, i0, j, k)), av1 = u32(G.ov, fvidx(G, i0, j + 1, k));
    const uint32_t aw0 = u32(G.ow, fwidx(G, i0, j, k)), aw1 = u32(G.ow, fwidx(G, i0, j, k + 1));
    const float4 h4 = ld4(G.eta, (size_t)j * nx + i0);
    const float4 hym = ld4(G.eta, (size_t)jm * nx + i0), hyp = ld4(G.eta, (size_t)jp * nx + i0);
    const float hxm = G.eta[(size_t)j * nx + im0], hxp = G.eta[(size_t)j * nx + ip3];
    auto byte = [](uint32_t w, int e) { return (uint8_t)(w >> (8 * e)); };
    uint32_t out = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint8_t ap[6] = {au[e], au[e + 1], byte(av0, e), byte(av1, e), byte(aw0, e), byte(aw1, e)};
      const uint8_t sn[6] = {e > 0 ? byte(sc, e - 1) : sxm, e < 3 ? byte(sc, e + 1) : sxp, byte(sym, e),
                             byte(syp, e), byte(szm, e), byte(szp, e)};
      const float hn[4] = {e > 0 ? q_get(h4, e - 1) : hxm, e < 3 ? q_get(h4, e + 1) : hxp, q_get(hym, e),
                           q_get(hyp, e)};
      out |= (uint32_t)row_code(G, i0 + e, j, k, byte(sc, e), sn, ap, q_get(h4, e), hn, wet, ghosts) << (8 * e);
    }
    *reinterpret_cast<uint32_t*>(mask + cidx(G, i0, j, k)) = out;
  }
  reduce_finish<false>(wet, ghosts, OP_COUNT, ctl, partials);
}

// r = b - A p, |r|^2 on quads (k_residual with four x-cells per thread: the
// pressures of the quad and of its y and z neighbours as 16-byte loads with
// indices clamped into the grid, the row's case decided after; apply_row_t's
// arithmetic per cell).
// BACKWARD: the same pass also folds the backward error of the r it stores
// into OP_BACKWARD, k_backward_q's arithmetic from the same pressures (|A||p|
// beside A p), so that the wrap-up reads p once for both.
template <bool GEN = true, bool BACKWARD = false>
__global__ void __launch_bounds__(QBLOCK) k_residual_q(Fine G, QuadGrid Q, const double* p, const float* b, float* r,
                                                       int op, Ctl* ctl, double* partials, int gated) {
#pragma clang fp contract(off)
  if (gated && !replacing(ctl)) return;
  double s = 0.0, unused = 0.0, worst = 0.0;
  int qx, y, k;
  const bool valid = quad_coords(Q, qx, y, k);
  if (valid) {
    const size_t plane = (size_t)G.nx * G.ny;
    const size_t c0 = ((size_t)k * G.ny + y) * G.nx + 4 * (size_t)qx;
    const uint32_t mq = *reinterpret_cast<const uint32_t*>(G.mask + c0);
    const float4 bq = ld4(b, c0);
    auto quad = [&](size_t c, double v[4]) {
      const double2 a = *reinterpret_cast<const double2*>(p + c), e = *reinterpret_cast<const double2*>(p + c + 2);
      v[0] = a.x; v[1] = a.y; v[2] = e.x; v[3] = e.y;
    };
    double pc[4], pym[4], pyp[4], pzm[4], pzp[4];
    quad(c0, pc);
    quad(y > 0 ? c0 - G.nx : c0, pym);
    quad(y + 1 < G.ny ? c0 + G.nx : c0, pyp);
    quad(k > 0 ? c0 - plane : c0, pzm);
    quad(k + 1 < G.nz ? c0 + plane : c0, pzp);
    const double pxm = p[qx > 0 ? c0 - 1 : c0], pxp = p[4 * qx + 4 < G.nx ? c0 + 4 : c0 + 3];
    float4 out;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint8_t m = (uint8_t)(mq >> (8 * e));
      double acc = 0.0, aabs = 0.0;
      if (m & M_WET) {
        const double pe = pc[e], pa = fabs(pc[e]);
        const double nb[6] = {e > 0 ? pc[e - 1] : pxm, e < 3 ? pc[e + 1] : pxp, pym[e], pyp[e], pzm[e], pzp[e]};
        if (!(m & M_GHOST)) {
#pragma unroll
          for (int f = 0; f < 6; ++f)
            if (m & (1u << f)) {
              acc += pe - nb[f];
              if (BACKWARD) aabs += pa + fabs(nb[f]);
            }
        } else {
          uint8_t type[6];
          float it[6], a[6];
          ghost_row<GEN>(G, 4 * qx + e, y, k, m, type, it, a);
#pragma unroll
          for (int f = 0; f < 6; ++f) {
            if (type[f] == 1) {
              acc += (double)a[f] * (pe - nb[f]);
              if (BACKWARD) aabs += (double)a[f] * (pa + fabs(nb[f]));
            } else if (type[f] == 2) {
              acc += (double)a[f] * pe * (double)it[f];
              if (BACKWARD) aabs += (double)a[f] * pa * (double)it[f];
            }
          }
        }
        acc = acc * (double)G.inv;
      }
      const double rc = (double)q_get(bq, e) - acc;
      const float rf = (float)rc;
      q_set(out, e, rf);
      s += rc * rc;
      if (BACKWARD && (m & M_WET)) {
        const double scale = fabs((double)q_get(bq, e)) + aabs * (double)G.inv;
        if (scale > 0.0) worst = fmax(worst, fabs((double)rf) / scale);
      }
    }
    st4(r, c0, out);
  }
  if (op >= 0) reduce_finish<false>(s, unused, op, ctl, partials);
  if (BACKWARD) reduce_finish<true>(worst, unused, OP_BACKWARD, ctl, partials);
}

// The aperture codes and solid flags of a quad's faces and cells for the
// aperture-weighted divergence: the y and z faces and the solid flags as one
// 4-byte load each (rows of nx bytes, nx % 4 == 0), the five x faces as bytes.
struct ApQuad {
  uint8_t u[5];
  uint32_t v0, v1, w0, w1, solid;
};
__device__ __forceinline__ ApQuad ap_quad(const Fine& G, size_t c0, size_t fu, size_t fv0, size_t fv1, size_t fw0,
                                          size_t fw1) {
  auto u32 = [](const uint8_t* a, size_t c) { return *reinterpret_cast<const uint32_t*>(a + c); };
  ApQuad A;
#pragma unroll
  for (int e = 0; e < 5; ++e) A.u[e] = G.ou[fu + e];
  A.v0 = u32(G.ov, fv0); A.v1 = u32(G.ov, fv1); A.w0 = u32(G.ow, fw0); A.w1 = u32(G.ow, fw1);
  A.solid = u32(G.solid, c0);
  return A;
}
__device__ __forceinline__ uint8_t ap_byte(uint32_t word, int e) { return (uint8_t)(word >> (8 * e)); }

// b = scale * div + ghost terms, prec (optional) and |b|^2 on quads
// (k_assemble with four x-cells per thread: the faces of the quad as 16-byte
// loads, its x faces as five scalars, or the given divergence; the row's
// arithmetic is k_assemble's, dry cells get 0).
__global__ void __launch_bounds__(QBLOCK) k_assemble_q(Fine G, QuadGrid Q, const float* us, const float* vs,
                                                       const float* ws, const float* div_in, float scale, float* b,
                                                       float* prec_out, Ctl* ctl, double* partials) {
#pragma clang fp contract(off)
  double s2 = 0.0, unused = 0.0;
  int qx, j, k;
  const bool valid = quad_coords(Q, qx, j, k);
  if (valid) {
    const int i0 = 4 * qx;
    const size_t c0 = cidx(G, i0, j, k);
    const uint32_t mq = *reinterpret_cast<const uint32_t*>(G.mask + c0);
    const size_t fu = fuidx(G, i0, j, k), fv0 = fvidx(G, i0, j, k), fv1 = fvidx(G, i0, j + 1, k);
    const size_t fw0 = fwidx(G, i0, j, k), fw1 = fwidx(G, i0, j, k + 1);
    float u5[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    float4 v0 = f4zero(), v1 = f4zero(), w0 = f4zero(), w1 = f4zero(), dq = f4zero();
    ApQuad A{};
    if (div_in) {
      dq = ld4(div_in, c0);
    } else {
#pragma unroll
      for (int e = 0; e < 5; ++e) u5[e] = us[fu + e];
      v0 = ld4(vs, fv0); v1 = ld4(vs, fv1); w0 = ld4(ws, fw0); w1 = ld4(ws, fw1);
      if (G.cut) A = ap_quad(G, c0, fu, fv0, fv1, fw0, fw1);
    }
    float4 out, pout = f4zero();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint8_t m = (uint8_t)(mq >> (8 * e));
      float bb = 0.f;
      if (m & M_WET) {
        float div;
        if (div_in) {
          div = q_get(dq, e);
        } else if (G.cut && ap_byte(A.solid, e)) {
          div = 0.f;
        } else {
          float fl = u5[e], fr = u5[e + 1], ff = q_get(v0, e), fk = q_get(v1, e), fb = q_get(w0, e), ft = q_get(w1, e);
          if (G.cut) {
            fl = aperture(G, 0, fu + e, A.u[e]) * fl; fr = aperture(G, 0, fu + e + 1, A.u[e + 1]) * fr;
            ff = aperture(G, 1, fv0 + e, ap_byte(A.v0, e)) * ff; fk = aperture(G, 1, fv1 + e, ap_byte(A.v1, e)) * fk;
            fb = aperture(G, 2, fw0 + e, ap_byte(A.w0, e)) * fb; ft = aperture(G, 2, fw1 + e, ap_byte(A.w1, e)) * ft;
          }
          div = ((fr - fl + ft - fb) + (fk - ff)) / G.dx;
        }
        bb = scale * div;
        float diag = 0.f;
        if (!(m & M_GHOST)) {
          for (int f = 0; f < 6; ++f) if (m & (1u << f)) diag += 1.0f;
        } else {
          RowG R;
          full_row(G, i0 + e, j, k, R);
          for (int f = 0; f < 6; ++f) {
            if (R.type[f] == 1) diag += R.a[f];
            else if (R.type[f] == 2) {
              diag += R.a[f] * R.it[f];
              bb += R.a[f] * R.val[f] * R.it[f] * G.inv;
            }
          }
        }
        if (prec_out) q_set(pout, e, diag > 0.f ? 1.0f / (diag * G.inv) : 0.f);
      }
      q_set(out, e, bb);
      s2 += (double)bb * (double)bb;
    }
    st4(b, c0, out);
    if (prec_out) st4(prec_out, c0, pout);
  }
  reduce_finish<false>(s2, unused, OP_B2, ctl, partials);
}

// Velocity correction on quads (k_correct with four x-cells per thread: the
// x-, y- and z- faces of the quad and the closing faces on the +x, +y, +z
// walls; the solid flags, row codes and pressures of the quad and of the cells
// below it in x, y and z as 4-, 16- and 32-byte loads with indices clamped into
// the grid). Each face's arithmetic is k_correct's (corrected_side, ghost_up).
__global__ void __launch_bounds__(QBLOCK) k_correct_q(Fine G, QuadGrid Q, const float* us, const float* vs,
                                                      const float* ws, const double* p, float k1, float* u, float* v,
                                                      float* w, Ctl* ctl, double* partials) {
#pragma clang fp contract(off)
  double mx = 0.0, unused = 0.0;
  int qx, j, k;
  const bool valid = quad_coords(Q, qx, j, k);
  if (valid) {
    const int nx = G.nx, ny = G.ny, nz = G.nz, i0 = 4 * qx;
    const size_t plane = (size_t)nx * ny;
    const size_t c0 = cidx(G, i0, j, k);
    const size_t cy = j > 0 ? c0 - nx : c0, cz = k > 0 ? c0 - plane : c0, cx = i0 > 0 ? c0 - 1 : c0;
    auto u32 = [](const uint8_t* a, size_t c) { return *reinterpret_cast<const uint32_t*>(a + c); };
    auto byte = [](uint32_t wv, int e) { return (uint8_t)(wv >> (8 * e)); };
    auto pquad = [&](size_t c, double o[4]) {
      const double2 a = *reinterpret_cast<const double2*>(p + c), b = *reinterpret_cast<const double2*>(p + c + 2);
      o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    };
    const uint32_t sq = u32(G.solid, c0), mq = u32(G.mask, c0);
    const uint32_t sy = u32(G.solid, cy), my = u32(G.mask, cy), sz = u32(G.solid, cz), mz = u32(G.mask, cz);
    const uint8_t sx = G.solid[cx], mxl = G.mask[cx];
    double pq[4], py[4], pz[4];
    pquad(c0, pq);
    pquad(cy, py);
    pquad(cz, pz);
    const double px = p[cx];
    const size_t fu = fuidx(G, i0, j, k), fv = fvidx(G, i0, j, k), fw = fwidx(G, i0, j, k);
    float ux[5];
    uint8_t ax[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      ux[e] = e < 4 || i0 + 4 == nx ? us[fu + e] : 0.f;
      ax[e] = e < 4 || i0 + 4 == nx ? G.ou[fu + e] : 0;
    }
    const float4 vq = ld4(vs, fv), wq = ld4(ws, fw);
    const uint32_t avq = u32(G.ov, fv), awq = u32(G.ow, fw);
    const bool top_y = j == ny - 1, top_z = k == nz - 1;
    const size_t fvt = fvidx(G, i0, ny, k), fwt = fwidx(G, i0, j, nz);
    const float4 vt = top_y ? ld4(vs, fvt) : f4zero(), wt = top_z ? ld4(ws, fwt) : f4zero();
    const uint32_t awt = top_z ? u32(G.ow, fwt) : 0u;
    const double kk = (double)k1, dx = (double)G.dx;
    // x- faces of the four cells, and the closing +x face
    float uo[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = i0 + e;
      const uint8_t sl = e > 0 ? byte(sq, e - 1) : sx, ml = e > 0 ? byte(mq, e - 1) : mxl;
      const double pl = e > 0 ? pq[e - 1] : px;
      const bool inner = i > 0;  // i < nx holds inside the quad
      float val = ux[e];
      if (inner && ax[e] && !sl && !byte(sq, e))
        val = corrected_side(G, val, pl, pq[e], (ml & M_WET) != 0, (byte(mq, e) & M_WET) != 0, k1, i - 1, j, i, j, k);
      uo[e] = val;
      mx = fmax(mx, (double)fabsf(val));
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) u[fu + e] = uo[e];
    if (i0 + 4 == nx) {
      u[fu + 4] = ux[4];
      mx = fmax(mx, (double)fabsf(ux[4]));
    }
    // y- faces, and the closing +y faces
    float4 vo;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = i0 + e;
      float val = q_get(vq, e);
      if (j > 0 && byte(avq, e) && !byte(sy, e) && !byte(sq, e))
        val = corrected_side(G, val, py[e], pq[e], (byte(my, e) & M_WET) != 0, (byte(mq, e) & M_WET) != 0, k1, i,
                             j - 1, i, j, k);
      q_set(vo, e, val);
      mx = fmax(mx, (double)fabsf(val));
    }
    st4(v, fv, vo);
    if (top_y) {
      st4(v, fvt, vt);
      mx = fmax(mx, (double)fmaxf(fmaxf(fabsf(vt.x), fabsf(vt.y)), fmaxf(fabsf(vt.z), fabsf(vt.w))));
    }
    // z- faces, and the closing +z faces
    auto zface = [&](float val, uint8_t open, int kf, uint8_t mb, double pb, uint8_t ma, double pa, int i, bool sb,
                     bool sa) {
      if (G.phi) {
        // level set (apic3d.rs:1016-1048): a face next to a solid is not
        // corrected, and air may lie below water
        if (kf >= 1 && kf < nz && open && !sb && !sa) {
          const bool wb = (mb & M_WET) != 0, wa = (ma & M_WET) != 0;
          float a, b;
          if (wb && wa) val = (float)((double)val - kk * (pa - pb) / dx);
          else if (wb) { ghost_up(G, i, j, kf - 1, a, b); val = (float)((double)val - kk * ((double)b - pb) * (double)a / dx); }
          else if (wa) { ghost_phi(G, i, j, kf, i, j, kf - 1, a, b); val = (float)((double)val - kk * (pa - (double)b) * (double)a / dx); }
        }
        return val;
      }
      if (kf >= 1 && open && (mb & M_WET)) {
        if (kf < nz && (ma & M_WET)) {
          val = (float)((double)val - kk * (pa - pb) / dx);
        } else {
          float a, b;
          ghost_up(G, i, j, kf - 1, a, b);
          val = (float)((double)val - kk * ((double)b - pb) * (double)a / dx);
        }
      }
      return val;
    };
    float4 wo;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float val = zface(q_get(wq, e), byte(awq, e), k, byte(mz, e), pz[e], byte(mq, e), pq[e], i0 + e,
                              byte(sz, e) != 0, byte(sq, e) != 0);
      q_set(wo, e, val);
      mx = fmax(mx, (double)fabsf(val));
    }
    st4(w, fw, wo);
    if (top_z) {
      float4 wt2;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // face nz: below and (clamped) above are the top cell itself
        const float val = zface(q_get(wt, e), byte(awt, e), nz, byte(mq, e), pq[e], byte(mq, e), pq[e], i0 + e,
                                byte(sq, e) != 0, byte(sq, e) != 0);
        q_set(wt2, e, val);
        mx = fmax(mx, (double)fabsf(val));
      }
      st4(w, fwt, wt2);
    }
  }
  reduce_finish<true>(mx, unused, OP_UMAX, ctl, partials);
}

// max |div u| over wet rows and over plain rows on quads (k_divmetric with
// four x-cells per thread; divergence_at's arithmetic per cell).
__global__ void __launch_bounds__(QBLOCK) k_divmetric_q(Fine G, QuadGrid Q, const float* u, const float* v,
                                                        const float* w, Ctl* ctl, double* partials) {
#pragma clang fp contract(off)
  double all = 0.0, plain = 0.0;
  int qx, j, k;
  const bool valid = quad_coords(Q, qx, j, k);
  if (valid) {
    const int i0 = 4 * qx;
    const size_t c0 = cidx(G, i0, j, k);
    const uint32_t mq = *reinterpret_cast<const uint32_t*>(G.mask + c0);
    if (mq & 0x40404040u) {  // a wet cell in the quad
      const size_t fu = fuidx(G, i0, j, k), fv0 = fvidx(G, i0, j, k), fv1 = fvidx(G, i0, j + 1, k);
      const size_t fw0 = fwidx(G, i0, j, k), fw1 = fwidx(G, i0, j, k + 1);
      float u5[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) u5[e] = u[fu + e];
      const float4 v0 = ld4(v, fv0), v1 = ld4(v, fv1), w0 = ld4(w, fw0), w1 = ld4(w, fw1);
      const ApQuad A = G.cut ? ap_quad(G, c0, fu, fv0, fv1, fw0, fw1) : ApQuad{};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint8_t m = (uint8_t)(mq >> (8 * e));
        if (!(m & M_WET)) continue;
        float div = 0.f;
        if (!(G.cut && ap_byte(A.solid, e))) {
          float fl = u5[e], fr = u5[e + 1], ff = q_get(v0, e), fk = q_get(v1, e), fb = q_get(w0, e), ft = q_get(w1, e);
          if (G.cut) {
            fl = aperture(G, 0, fu + e, A.u[e]) * fl; fr = aperture(G, 0, fu + e + 1, A.u[e + 1]) * fr;
            ff = aperture(G, 1, fv0 + e, ap_byte(A.v0, e)) * ff; fk = aperture(G, 1, fv1 + e, ap_byte(A.v1, e)) * fk;
            fb = aperture(G, 2, fw0 + e, ap_byte(A.w0, e)) * fb; ft = aperture(G, 2, fw1 + e, ap_byte(A.w1, e)) * ft;
          }
          div = ((fr - fl + ft - fb) + (fk - ff)) / G.dx;
        }
        const double d = (double)fabsf(div);
        all = fmax(all, d);
        // plain: no ghost face (a row flagged for a fractional aperture alone is plain)
        bool plain_row = !(m & M_GHOST);
        if (!plain_row && G.frac) {
          RowG R;
          full_row(G, i0 + e, j, k, R);
          plain_row = true;
          for (int f = 0; f < 6; ++f) plain_row = plain_row && R.type[f] != 2;
        }
        if (plain_row) plain = fmax(plain, d);
      }
    }
  }
  reduce_finish<true>(all, plain, OP_DIVMAX, ctl, partials);
}

// Oettli-Prager backward error on quads (k_backward with four x-cells per
// thread, the pressures of the quad and of its y and z neighbours as 32-byte
// loads with indices clamped into the grid; k_backward's arithmetic per cell).
template <bool GEN = true>
__global__ void __launch_bounds__(QBLOCK) k_backward_q(Fine G, QuadGrid Q, const double* p, const float* b,
                                                       const float* r, Ctl* ctl, double* partials) {
#pragma clang fp contract(off)
  double worst = 0.0, unused = 0.0;
  int qx, y, k;
  const bool valid = quad_coords(Q, qx, y, k);
  if (valid) {
    const size_t plane = (size_t)G.nx * G.ny;
    const size_t c0 = ((size_t)k * G.ny + y) * G.nx + 4 * (size_t)qx;
    const uint32_t mq = *reinterpret_cast<const uint32_t*>(G.mask + c0);
    auto quad = [&](size_t c, double v[4]) {
      const double2 a = *reinterpret_cast<const double2*>(p + c), e = *reinterpret_cast<const double2*>(p + c + 2);
      v[0] = a.x; v[1] = a.y; v[2] = e.x; v[3] = e.y;
    };
    double pc[4], pym[4], pyp[4], pzm[4], pzp[4];
    quad(c0, pc);
    quad(y > 0 ? c0 - G.nx : c0, pym);
    quad(y + 1 < G.ny ? c0 + G.nx : c0, pyp);
    quad(k > 0 ? c0 - plane : c0, pzm);
    quad(k + 1 < G.nz ? c0 + plane : c0, pzp);
    const double pxm = p[qx > 0 ? c0 - 1 : c0], pxp = p[4 * qx + 4 < G.nx ? c0 + 4 : c0 + 3];
    const float4 bq = ld4(b, c0), rq = ld4(r, c0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint8_t m = (uint8_t)(mq >> (8 * e));
      if (!(m & M_WET)) continue;
      const double pe = fabs(pc[e]);
      const double pn[6] = {e > 0 ? pc[e - 1] : pxm, e < 3 ? pc[e + 1] : pxp, pym[e], pyp[e], pzm[e], pzp[e]};
      double acc = 0.0;
      if (!(m & M_GHOST)) {
        for (int f = 0; f < 6; ++f)
          if (m & (1u << f)) acc += pe + fabs(pn[f]);
      } else {
        uint8_t type[6];
        float it[6], a[6];
        ghost_row<GEN>(G, 4 * qx + e, y, k, m, type, it, a);
        for (int f = 0; f < 6; ++f) {
          if (type[f] == 1) acc += (double)a[f] * (pe + fabs(pn[f]));
          else if (type[f] == 2) acc += (double)a[f] * pe * (double)it[f];
        }
      }
      const double scale = fabs((double)q_get(bq, e)) + acc * (double)G.inv;
      if (scale > 0.0) worst = fmax(worst, fabs((double)q_get(rq, e)) / scale);
    }
  }
  reduce_finish<true>(worst, unused, OP_BACKWARD, ctl, partials);
}

// Cold start, p = 0: p64 = 0, r = b - A 0 = b exactly, |r|^2 into `op`
// (k_load_p and k_residual in one pass over b).
__global__ void __launch_bounds__(QBLOCK) k_start_cold(size_t nquads, const float* b, double* p, float* r, int op,
                                                       Ctl* ctl, double* partials) {
  double s = 0.0, unused = 0.0;
  const size_t t = (size_t)blockIdx.x * QBLOCK + threadIdx.x;
  if (t < nquads) {
    const size_t c = 4 * t;
    const float4 bq = ld4(b, c);
    double2* pp = reinterpret_cast<double2*>(p + c);
    pp[0] = make_double2(0.0, 0.0);
    pp[1] = make_double2(0.0, 0.0);
    st4(r, c, bq);
    s = (double)bq.x * bq.x + (double)bq.y * bq.y + (double)bq.z * bq.z + (double)bq.w * bq.w;
  }
  reduce_finish<false>(s, unused, op, ctl, partials);
}

// ---------------------------------------------------------------- residual + restriction
// The residual t = r - A z of the 16 fine cells of (fine quad column qx,
// coarse row J, coarse plane K), summed per coarse cell in the reference's
// child order and scaled by 1/8 (restrict3, delta3d_multigrid.rs:258-272),
// with the work split over two threads: the first 128 threads of a
// workgroup take the fine plane 2K, the last 128 the plane 2K + 1. The first half's
// partial sums (children of plane 2K, in the reference's order) reach the
// second half through LDS, which adds its own children in order: the sum,
// and so the coarse r, is the one-thread sum bit for bit, with twice the
// waves in flight to overlap the loads with the arithmetic (38 -> 28 us at
// 256x256x128). H: the restriction's quad grid for 128-thread halves.
template <int ST, bool GEN = true>
__global__ void __launch_bounds__(QBLOCK) k_q_restrict2(Fine G, QuadGrid H, const typename Store<ST>::T* z,
                                                        const float* r, Level L0, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  __shared__ float2 carry[QBLOCK / 2];
  if (!ctl->active) return;
  const int local = (int)(threadIdx.x & (QBLOCK / 2 - 1)), dz = (int)(threadIdx.x / (QBLOCK / 2));
  const unsigned Lb = xcd_block(blockIdx.x, gridDim.x);
  const unsigned bx = Lb % (unsigned)H.nbx, tb = Lb / (unsigned)H.nbx;
  const int qx = (int)bx * H.QX + (local & (H.QX - 1));
  const int J = (int)(tb % (unsigned)H.nby) * H.RY + local / H.QX;
  const int K = (int)(tb / (unsigned)H.nby);
  const bool valid = qx < H.nq && J < H.rows && K < H.planes;
  const size_t plane = (size_t)G.nx * G.ny;
  const int k = 2 * K + dz;
  float t[2][4];  // [dy][e]
#pragma unroll
  for (int dy = 0; dy < 2; ++dy) {
    const int y = 2 * J + dy;
    const size_t c0 = valid ? ((size_t)k * G.ny + y) * G.nx + 4 * (size_t)qx : 0;
    float4 zc = f4zero(), rq = f4zero(), ym = f4zero(), yp = f4zero(), zm = f4zero(), zp = f4zero();
    uint32_t mq = 0;
    if (valid) {
      mq = *reinterpret_cast<const uint32_t*>(G.mask + c0);
      zc = SV::ld4(z, c0);
      rq = ld4(r, c0);
      if (y > 0) ym = SV::ld4(z, c0 - G.nx);
      if (y + 1 < G.ny) yp = SV::ld4(z, c0 + G.nx);
      if (k > 0) zm = SV::ld4(z, c0 - plane);
      if (k + 1 < G.nz) zp = SV::ld4(z, c0 + plane);
    }
    float left, right;
    x_neighbours(H, valid, qx, zc, left, right, [&](int off) { return SV::ld(z, c0 + off); });
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint8_t m = (uint8_t)(mq >> (8 * e));
      const float pc = q_get(zc, e);
      const float xl = e == 0 ? left : q_get(zc, e - 1);
      const float xr = e == 3 ? right : q_get(zc, e + 1);
      const float nb[6] = {xl, xr, q_get(ym, e), q_get(yp, e), q_get(zm, e), q_get(zp, e)};
      float prec;
      const float az = valid ? row_local<GEN>(G, 4 * qx + e, y, k, m, pc, nb, prec) : 0.f;
      t[dy][e] = q_get(rq, e) - az;
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
  SV::st2(reinterpret_cast<typename SV::T*>(L0.r), lidx(L0, 2 * qx, J, K), out);
}

// A vector held in storage ST, widened to f32.
template <int ST>
__global__ void __launch_bounds__(BLOCK) k_unpack(size_t n, const typename Store<ST>::T* z, float* out) {
  FSP_FOR(c, n) out[c] = Store<ST>::ld(z, c);
}

// ---------------------------------------------------------------- CG steps on quads
// d = z (first) or z + beta d_old, q = A d, <d, q> (delta3d_mobile.rs:520-523, :460-465).
// z is held in storage ST.
template <int ST, bool GEN = true>
__global__ void __launch_bounds__(QBLOCK) k_q_spmv_dir(Fine G, QuadGrid Q, const typename Store<ST>::T* z, float* d0,
                                                       float* d1, float* q, Ctl* ctl, double* partials) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (!ctl->active) return;
  const int sel = ctl->dsel, first = ctl->first_dir;
  const float beta = (float)ctl->beta;
  const float* dold = sel ? d1 : d0;
  float* dnew = sel ? d0 : d1;
  int qx, y, k;
  const bool valid = quad_coords(Q, qx, y, k);
  const size_t plane = (size_t)G.nx * G.ny;
  const size_t c0 = valid ? ((size_t)k * G.ny + y) * G.nx + 4 * (size_t)qx : 0;
  auto dir4 = [&](size_t c) {
    float4 a = SV::ld4(z, c);
    if (!first) {
      const float4 b = ld4(dold, c);
      a.x = a.x + beta * b.x; a.y = a.y + beta * b.y; a.z = a.z + beta * b.z; a.w = a.w + beta * b.w;
    }
    return a;
  };
  float4 dc = f4zero(), ym = f4zero(), yp = f4zero(), zm = f4zero(), zp = f4zero();
  uint32_t mq = 0;
  if (valid) {
    mq = *reinterpret_cast<const uint32_t*>(G.mask + c0);
    dc = dir4(c0);
    if (y > 0) ym = dir4(c0 - G.nx);
    if (y + 1 < G.ny) yp = dir4(c0 + G.nx);
    if (k > 0) zm = dir4(c0 - plane);
    if (k + 1 < G.nz) zp = dir4(c0 + plane);
  }
  float left, right;
  x_neighbours(Q, valid, qx, dc, left, right, [&](int off) {
    const size_t c = c0 + off;
    return first ? SV::ld(z, c) : SV::ld(z, c) + beta * dold[c];
  });
  double s = 0.0, unused = 0.0;
  if (valid) {
    float4 qo;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint8_t m = (uint8_t)(mq >> (8 * e));
      const float pc = q_get(dc, e);
      const float xl = e == 0 ? left : q_get(dc, e - 1);
      const float xr = e == 3 ? right : q_get(dc, e + 1);
      const float nb[6] = {xl, xr, q_get(ym, e), q_get(yp, e), q_get(zm, e), q_get(zp, e)};
      float prec;
      q_set(qo, e, row_local<GEN>(G, 4 * qx + e, y, k, m, pc, nb, prec));
    }
    st4(dnew, c0, dc);
    st4(q, c0, qo);
    s = (double)dc.x * qo.x + (double)dc.y * qo.y + (double)dc.z * qo.z + (double)dc.w * qo.w;
  }
  reduce_finish<false>(s, unused, OP_DQ, ctl, partials);
}

// p += alpha d (f64), r -= alpha q, |r|^2; four cells per thread.
__global__ void __launch_bounds__(QBLOCK) k_q_update(size_t nquads, double* p, float* r, const float* d0,
                                                     const float* d1, const float* q, Ctl* ctl, double* partials) {
  if (!ctl->active) return;
  const double alpha64 = ctl->alpha;
  const float alpha = (float)alpha64;
  const float* d = ctl->dsel ? d1 : d0;
  double s = 0.0, unused = 0.0;
  const size_t t = (size_t)blockIdx.x * QBLOCK + threadIdx.x;
  if (t < nquads) {
    const size_t c = 4 * t;
    const float4 dv = ld4(d, c), qv = ld4(q, c);
    float4 rv = ld4(r, c);
    double2* pp = reinterpret_cast<double2*>(p + c);
    double2 p0 = pp[0], p1 = pp[1];
    p0.x += alpha64 * (double)dv.x; p0.y += alpha64 * (double)dv.y;
    p1.x += alpha64 * (double)dv.z; p1.y += alpha64 * (double)dv.w;
    pp[0] = p0;
    pp[1] = p1;
    rv.x = rv.x - alpha * qv.x; rv.y = rv.y - alpha * qv.y; rv.z = rv.z - alpha * qv.z; rv.w = rv.w - alpha * qv.w;
    st4(r, c, rv);
    s = (double)rv.x * rv.x + (double)rv.y * rv.y + (double)rv.z * rv.z + (double)rv.w * rv.w;
  }
  reduce_finish<false>(s, unused, OP_RN, ctl, partials);
}

// ---------------------------------------------------------------- small coarse levels in one workgroup
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_MAX_LEVELS = 8;
constexpr int TAIL_LDS_FLOATS = 64 * 1024 / 4;  // LDS of the tail workgroup

struct TailSet {
  Level lv[TAIL_MAX_LEVELS];  // global arrays of the tail levels (first .. last)
  const float* pack;          // Galerkin tail: every level's couplings in LDS order (k_gct_pack)
  int count;
  int smoother;               // FSP_SMOOTH_*
  int pre, post, coarse;
};

// A tail level inside LDS: sizes and float offsets of its arrays.
struct TailLvl {
  int nx, ny, nz, n;
  float inv;
  int o_kind, o_ou, o_ov, o_ow, o_diag, o_x, o_r, o_t, o_dd;
  int galerkin;
};

__device__ __forceinline__ void tail_decode(const TailLvl& L, int c, int& i, int& j, int& k) {
  const int t = c / L.nx;
  i = c - t * L.nx;
  k = t / L.ny;
  j = t - k * L.ny;
}

// row_sums (delta3d_multigrid.rs:216-255) on LDS arrays; x at offset xo.
__device__ __forceinline__ float tail_row(const float* lds, const TailLvl& L, int i, int j, int k, int c, int xo) {
#pragma clang fp contract(off)
  const int8_t* kind = reinterpret_cast<const int8_t*>(lds + L.o_kind);
  if (kind[c] <= 0) return 0.f;
  const float* x = lds + xo;
  const float* ou = lds + L.o_ou;
  const float* ov = lds + L.o_ov;
  const float* ow = lds + L.o_ow;
  const float pc = x[c];
  float acc = 0.f;
  const int nx = L.nx, ny = L.ny, nz = L.nz, plane = nx * ny;
  const bool gal = L.galerkin != 0;
  auto face = [&](float a, bool has, int n, bool top) {
    if (a == 0.f) return;
    if (has) {
      const int8_t kn = kind[n];
      if (kn > 0) acc += a * (pc - x[n]);
      else if (!gal && kn < 0) acc += 2.0f * a * pc;
    } else if (!gal && top) {
      acc += 2.0f * a * pc;
    }
  };
  face(ou[(k * ny + j) * (nx + 1) + i], i > 0, c - 1, false);
  face(ou[(k * ny + j) * (nx + 1) + i + 1], i + 1 < nx, c + 1, false);
  face(ov[(k * (ny + 1) + j) * nx + i], j > 0, c - nx, false);
  face(ov[(k * (ny + 1) + j + 1) * nx + i], j + 1 < ny, c + nx, false);
  face(ow[(k * ny + j) * nx + i], k > 0, c - plane, false);
  face(ow[((k + 1) * ny + j) * nx + i], k + 1 < nz, c + plane, true);
  if (gal) acc += lds[L.o_dd + c] * pc;
  return acc * L.inv;
}

// One smoothing pass on tail level l. RBGS: colour `col` in place (first: x
// starts at zero). Jacobi: t = x + w(r - Lx)/diag, then x = t. Stored values
// are rounded to storage ST.
template <int ST>
__device__ __forceinline__ void tail_pass(float* lds, const TailLvl& L, int smoother, int col, bool first) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  float* x = lds + L.o_x;
  const float* r = lds + L.o_r;
  const float* diag = lds + L.o_diag;
  if (smoother == 0) {
    float* t = lds + L.o_t;
    const float omega = 6.0f / 7.0f;
    for (int c = threadIdx.x; c < L.n; c += TAIL_THREADS) {
      int i, j, k;
      tail_decode(L, c, i, j, k);
      const float row = first ? 0.f : tail_row(lds, L, i, j, k, c, L.o_x);
      float xc = first ? 0.f : x[c];
      const float dg = diag[c];
      if (dg > 0.f) xc = SV::rnd(xc + omega * (r[c] - row) / dg);
      t[c] = xc;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < L.n; c += TAIL_THREADS) x[c] = t[c];
    __syncthreads();
    return;
  }
  for (int c = threadIdx.x; c < L.n; c += TAIL_THREADS) {
    int i, j, k;
    tail_decode(L, c, i, j, k);
    const float dg = diag[c];
    if (((i + j + k) & 1) != col) {
      if (first) x[c] = 0.f;
      continue;
    }
    if (first) { x[c] = dg > 0.f ? SV::rnd(r[c] / dg) : 0.f; continue; }
    if (dg > 0.f) x[c] = SV::rnd(x[c] + (r[c] - tail_row(lds, L, i, j, k, c, L.o_x)) / dg);
  }
  __syncthreads();
}

template <int ST>
__device__ __forceinline__ void tail_sweeps(float* lds, const TailLvl& L, int smoother, int count, bool forward,
                                            bool coarsest, bool from_zero) {
  if (count == 0) {
    if (from_zero) {
      for (int c = threadIdx.x; c < L.n; c += TAIL_THREADS) lds[L.o_x + c] = 0.f;
      __syncthreads();
    }
    return;
  }
  bool fz = from_zero;
  if (smoother == 0) {
    for (int s = 0; s < count; ++s) { tail_pass<ST>(lds, L, 0, 0, fz); fz = false; }
    return;
  }
  if (coarsest) {
    const int half = count / 2;
    for (int s = 0; s < half; ++s) { tail_pass<ST>(lds, L, 1, 0, fz); fz = false; tail_pass<ST>(lds, L, 1, 1, false); }
    if (count & 1) { tail_pass<ST>(lds, L, 1, 0, fz); fz = false; }
    for (int s = 0; s < half; ++s) { tail_pass<ST>(lds, L, 1, 1, fz); fz = false; tail_pass<ST>(lds, L, 1, 0, false); }
    return;
  }
  for (int s = 0; s < count; ++s) {
    const int c0 = forward ? 0 : 1;
    tail_pass<ST>(lds, L, 1, c0, fz);
    fz = false;
    tail_pass<ST>(lds, L, 1, 1 - c0, false);
  }
}

// Descends and ascends every tail level in LDS (coarse_cycle3,
// delta3d_multigrid.rs:303-332), starting from the first tail level whose r
// is already restricted; writes that level's x back to global.
// Level vectors x and r are held in storage ST, in global memory and (as
// rounded f32 values) in LDS.
template <int ST>
__global__ void __launch_bounds__(TAIL_THREADS) k_tail(TailSet T, Ctl* ctl) {
#pragma clang fp contract(off)
  using SV = Store<ST>;
  if (!ctl->active) return;
  extern __shared__ float lds[];
  __shared__ TailLvl tl[TAIL_MAX_LEVELS];
  if (threadIdx.x == 0) {
    int off = 0;
    for (int l = 0; l < T.count; ++l) {
      const Level& g = T.lv[l];
      TailLvl& L = tl[l];
      L.nx = g.nx; L.ny = g.ny; L.nz = g.nz; L.n = (int)g.n; L.inv = g.inv;
      L.o_kind = off; off += (L.n + 3) / 4;
      L.o_ou = off; off += (g.nx + 1) * g.ny * g.nz;
      L.o_ov = off; off += g.nx * (g.ny + 1) * g.nz;
      L.o_ow = off; off += g.nx * g.ny * (g.nz + 1);
      L.o_diag = off; off += L.n;
      L.o_x = off; off += L.n;
      L.o_r = off; off += L.n;
      if (T.smoother == 0) { L.o_t = off; off += L.n; } else { L.o_t = L.o_x; }
      L.galerkin = g.galerkin;
      if (g.galerkin) { L.o_dd = off; off += L.n; } else { L.o_dd = L.o_diag; }
    }
  }
  __syncthreads();
  for (int l = 0; l < T.count; ++l) {
    const Level& g = T.lv[l];
    const TailLvl L = tl[l];
    int8_t* kind = reinterpret_cast<int8_t*>(lds + L.o_kind);
    const int nu = (g.nx + 1) * g.ny * g.nz, nv = g.nx * (g.ny + 1) * g.nz, nw = g.nx * g.ny * (g.nz + 1);
    for (int q = threadIdx.x; q < L.n; q += TAIL_THREADS) {
      kind[q] = g.kind[q];
      lds[L.o_diag + q] = g.diag[q];
      if (L.galerkin) lds[L.o_dd + q] = g.dd[q];
    }
    for (int q = threadIdx.x; q < nu; q += TAIL_THREADS) lds[L.o_ou + q] = g.ou[q];
    for (int q = threadIdx.x; q < nv; q += TAIL_THREADS) lds[L.o_ov + q] = g.ov[q];
    for (int q = threadIdx.x; q < nw; q += TAIL_THREADS) lds[L.o_ow + q] = g.ow[q];
    if (l == 0)
      for (int q = threadIdx.x; q < L.n; q += TAIL_THREADS)
        lds[L.o_r + q] = SV::ld(reinterpret_cast<const typename SV::T*>(g.r), q);
  }
  __syncthreads();
  const int last = T.count - 1;
  for (int l = 0; l <= last; ++l) {
    const TailLvl L = tl[l];
    tail_sweeps<ST>(lds, L, T.smoother, l == last ? T.coarse : T.pre, true, l == last, true);
    if (l < last) {
      const TailLvl C = tl[l + 1];
      for (int c = threadIdx.x; c < C.n; c += TAIL_THREADS) {
        int i, j, k;
        tail_decode(C, c, i, j, k);
        float s = 0.f;
        for (int ch = 0; ch < 8; ++ch) {
          const int fi = 2 * i + (ch & 1), fj = 2 * j + ((ch >> 1) & 1), fk = 2 * k + ((ch >> 2) & 1);
          const int f = (fk * L.ny + fj) * L.nx + fi;
          s += lds[L.o_r + f] - tail_row(lds, L, fi, fj, fk, f, L.o_x);
        }
        lds[C.o_r + c] = SV::rnd(0.125f * s);
      }
      __syncthreads();
    }
  }
  for (int l = last - 1; l >= 0; --l) {
    const TailLvl L = tl[l];
    const TailLvl C = tl[l + 1];
    for (int c = threadIdx.x; c < L.n; c += TAIL_THREADS) {
      int i, j, k;
      tail_decode(L, c, i, j, k);
      lds[L.o_x + c] = SV::rnd(lds[L.o_x + c] + lds[C.o_x + ((k >> 1) * C.ny + (j >> 1)) * C.nx + (i >> 1)]);
    }
    __syncthreads();
    tail_sweeps<ST>(lds, L, T.smoother, T.post, false, false, false);
  }
  const TailLvl L0 = tl[0];
  for (int q = threadIdx.x; q < L0.n; q += TAIL_THREADS)
    SV::st(reinterpret_cast<typename SV::T*>(T.lv[0].x), q, lds[L0.o_x + q]);
}

// LDS bytes k_tail needs for the given levels.
inline size_t tail_lds_bytes(const Level* lv, int count, int smoother) {
  size_t floats = 0;
  for (int l = 0; l < count; ++l) {
    const Level& g = lv[l];
    const size_t nu = (size_t)(g.nx + 1) * g.ny * g.nz, nv = (size_t)g.nx * (g.ny + 1) * g.nz,
                 nw = (size_t)g.nx * g.ny * (g.nz + 1);
    floats += (g.n + 3) / 4 + nu + nv + nw + (smoother == 0 ? 4 : 3) * g.n + (g.galerkin ? g.n : 0);
  }
  return floats * sizeof(float);
}

}  // namespace fsp
