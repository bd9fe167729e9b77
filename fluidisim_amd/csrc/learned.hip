// Learned preconditioner: a 3-level 3D convolutional encoder-decoder
// (8/16/32 channels, 3x3x3 kernels, ReLU, additive skip connections) with
// its forward and backward passes as binary16 MFMA implicit GEMMs (f32
// accumulation), an Adam optimiser on f32 master weights, and the glue the
// projection uses to apply it inside flexible CG. No counterpart in the
// reference solver.
//
// Activations are channels-last binary16, cells (k*ny+j)*nx+i. Level l of
// the network is the grid coarsened 2^l times (average pooling down, nearest
// upsampling folded into the next convolution: in the forward pass a 2x2x2
// convolution of the coarse grid per output parity, k_conv_up; in the weight
// gradient the upsampled loads of k_wgrad). Layer weights are
// W[co][tap][ci], tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding at the
// domain boundary.
//
// Convolution as an implicit GEMM on v_mfma_f32_16x16x32_f16 (k_conv): the
// layer's weights (M = output channels, in 16-channel tiles) times a 16-cell
// row of its input (N), K over (tap, input channel) in chunks of 32, the
// input brick and the weights staged in LDS. The weight gradient (k_wgrad)
// uses v_mfma_f32_16x16x16_f16 with K over cells.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/fsp.h"
#include "../../include/fsp_learned.h"
#include "learned_internal.h"

namespace fsp {
namespace net {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int WAVES = 4;  // wavefronts per workgroup of the weight-gradient kernels
constexpr int THREADS = 64 * WAVES;
#ifndef FSP_CONV_WAVES
#define FSP_CONV_WAVES 8  // 8 measured faster than 4 at 256x256x128 (1.03 vs 1.31 ms per forward)
#endif
#ifndef FSP_CONV_MIN_WAVES
#define FSP_CONV_MIN_WAVES 4  // wavefronts per SIMD the register budget of k_conv leaves room for
#endif
constexpr int CWAVES = FSP_CONV_WAVES;  // wavefronts per workgroup of the convolution kernels
constexpr int CTHREADS = 64 * CWAVES;

struct LayerDesc {
  int level, cin, cout, relu, skip, up;  // skip: layer index added before the ReLU (-1 none); up: input from level+1
};
// e1a e1b | e2a e2b | e3a e3b | d2a d2b | d1a d1b | head
constexpr int NL = 11;
constexpr LayerDesc LAYERS[NL] = {
    {0, 4, 8, 1, -1, 0},  {0, 8, 8, 1, -1, 0},   {1, 8, 16, 1, -1, 0}, {1, 16, 16, 1, -1, 0},
    {2, 16, 32, 1, -1, 0}, {2, 32, 32, 1, -1, 0}, {1, 32, 16, 1, 3, 1}, {1, 16, 16, 1, -1, 0},
    {0, 16, 8, 1, 1, 1},  {0, 8, 8, 1, -1, 0},   {0, 8, 1, 0, -1, 0},
};

// K of a layer (27 taps x input channels) padded to the 32 of one product
__host__ __device__ constexpr int kpad(int cin) { return ((27 * cin + 31) / 32) * 32; }
__host__ __device__ constexpr int npad(int cout) { return ((cout + 15) / 16) * 16; }
// Forward layers with 4 or 8 input and at most 8 output channels take two
// output rows per product: the A operand's rows 0-7 are the weights for row y,
// rows 8-15 those for row y + 1, and K runs over (dz, ry, dx, ci) of the four
// input rows y - 1 .. y + 2 the pair reads: 36 x 8 = 288 = 9 products per row
// pair against 7 per row (8 channels), 144 -> 160 = 5 against 4 (4 channels).
__host__ __device__ constexpr bool pair_rows(int cin, int cout) { return (cin == 4 || cin == 8) && cout <= 8; }
// K of a layer's forward weights
__host__ __device__ constexpr int kfwd(int cin, int cout) {
  return pair_rows(cin, cout) ? ((36 * cin + 31) / 32) * 32 : kpad(cin);
}
// channels of the gradient flowing into a layer's backward-data convolution
__host__ __device__ constexpr int gch(int cout) { return cout < 4 ? 4 : cout; }

size_t param_count() {
  size_t n = 0;
  for (const LayerDesc& L : LAYERS) n += (size_t)L.cout * 27 * L.cin + L.cout;
  return n;
}

// ---------------------------------------------------------------- convolution
struct ConvArgs {
  int nx, ny, nz;        // output grid
  int cnx, cny, cnz;     // input grid: the coarse grid of k_conv_up (else = output grid)
  const _Float16* in;    // [cells_in][CIN]
  const _Float16* wt;    // [npad(COUT)][K]; K = kpad(CIN), kfwd for row pairs; k_conv_up: k_pack_up's layout
  const float* bias;     // [COUT] (forward only)
  const _Float16* skip;  // [cells][COUT] (k_conv_up)
  _Float16* out;         // [cells][COUT]
  float* yout;           // head: [cells]
  const int32_t* gate;   // return at once while *gate == 0 (null: always run)
};

// Output brick of one workgroup: 16 x-cells by BY rows by BZ planes; its
// input halo (18 x (BY+2) x (BZ+2) cells, CIN channels) is staged in LDS
// once, so each input value is read from memory once per brick instead of
// once per tap.
template <int CIN, int COUT>
struct ConvBrick {
  static constexpr int BX = 16, BY = 8, BZ = (CIN >= 32 || COUT >= 32) ? 4 : 8;
  static constexpr int HX = BX + 2, HY = BY + 2, HZ = BZ + 2;
  static constexpr int HCELLS = HX * HY * HZ;
  static constexpr int TILES = BY * BZ;  // 16-cell rows of the brick
  static constexpr int TPW = TILES / CWAVES;
  static_assert(TILES % CWAVES == 0, "rows per wavefront");
};

__host__ __device__ constexpr int tap_offset(int tap, int hx, int hy) {
  return (tap / 9) * hy * hx + ((tap / 3) % 3) * hx + tap % 3;
}

// EPI 0: bias + ReLU -> binary16; 1: bias -> f32 channel 0 (head);
// 2: plain product -> binary16 (backward data). The upsampling layers (input
// from the coarser level, skip added before the ReLU) run k_conv_up.
//
// D[co][cell] = W[co][k] X[k][cell] on v_mfma_f32_16x16x32_f16: the weights
// are the A operand (16 output channels), a 16-cell row of the brick the B
// operand, K = 32 values of (tap, input channel) per product. Lane (m, g)
// then holds output channels 4g..4g+3 of cell m, written as one 8-byte
// store. Wavefront w owns rows w*TPW .. w*TPW+TPW-1; for each K chunk the
// weights are read from LDS once and multiplied into every row's tile. In
// the row-pair form (pair_rows) a tile is two rows y, y + 1 (y even) and lane
// (m, g) holds channels 4(g & 1)..+3 of row y + (g >> 1).
template <int CIN, int COUT, int EPI>
__global__ void __launch_bounds__(CTHREADS, FSP_CONV_MIN_WAVES) k_conv(ConvArgs a) {
  using Bk = ConvBrick<CIN, COUT>;
  constexpr bool PAIR = EPI != 2 && pair_rows(CIN, COUT);
  constexpr int TPW = PAIR ? Bk::TPW / 2 : Bk::TPW;
  constexpr int NT = npad(COUT) / 16;
  constexpr int KP = PAIR ? kfwd(CIN, COUT) : kpad(CIN);
  constexpr int CHUNKS = KP / 32;
  static_assert(COUT % 4 == 0 || (COUT == 1 && EPI == 1), "output channels in groups of 4");
  constexpr int NP = npad(COUT);
  constexpr int KPL = KP + 8;  // LDS row stride of the weights (halves): 16-byte rows spread over the banks
  __shared__ __attribute__((aligned(16))) _Float16 halo[Bk::HCELLS * CIN];
  __shared__ __attribute__((aligned(16))) _Float16 wl[NP * KPL];
  if (a.gate && !*a.gate) return;
  const int x0 = blockIdx.x * Bk::BX, y0 = blockIdx.y * Bk::BY, z0 = blockIdx.z * Bk::BZ;

  // weights and halo (zero outside the domain) -> LDS, every load of the
  // thread in flight before the first LDS write
  {
    constexpr int WV = NP * KP / 8;  // 16-byte vectors of the weights
    constexpr int WPER = (WV + CTHREADS - 1) / CTHREADS;
    uint4 wbuf[WPER];
#pragma unroll
    for (int r = 0; r < WPER; ++r) {
      const int q = (int)threadIdx.x + r * CTHREADS;
      wbuf[r] = make_uint4(0u, 0u, 0u, 0u);
      if (q < WV) wbuf[r] = *reinterpret_cast<const uint4*>(a.wt + (size_t)q * 8);
    }
    constexpr int VH = CIN >= 8 ? 8 : 4;  // halves per vector
    constexpr int VPC = CIN / VH;
    constexpr int TOTAL = Bk::HCELLS * VPC;
    constexpr int PER = (TOTAL + CTHREADS - 1) / CTHREADS;
    typedef uint32_t hv __attribute__((ext_vector_type(VH / 2)));
    hv buf[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int q = (int)threadIdx.x + r * CTHREADS;
      const int cell = q / VPC, part = q - cell * VPC;
      const int lx = cell % Bk::HX, t = cell / Bk::HX, ly = t % Bk::HY, lz = t / Bk::HY;
      const int xx = x0 + lx - 1, yy = y0 + ly - 1, zz = z0 + lz - 1;
      const bool in = q < TOTAL && xx >= 0 && xx < a.nx && yy >= 0 && yy < a.ny && zz >= 0 && zz < a.nz;
      buf[r] = hv(0u);
      if (in) {
        const size_t v = ((size_t)zz * a.ny + yy) * a.nx + xx;
        buf[r] = *reinterpret_cast<const hv*>(a.in + v * CIN + part * VH);
      }
    }
#pragma unroll
    for (int r = 0; r < WPER; ++r) {
      const int q = (int)threadIdx.x + r * CTHREADS;
      if (q < WV) {
        const int row = q / (KP / 8), col = (q - row * (KP / 8)) * 8;
        *reinterpret_cast<uint4*>(wl + row * KPL + col) = wbuf[r];
      }
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int q = (int)threadIdx.x + r * CTHREADS;
      if (q < TOTAL) *reinterpret_cast<hv*>(halo + (size_t)q * VH) = buf[r];
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  f4 acc[TPW][NT];
#pragma unroll
  for (int i = 0; i < TPW; ++i)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[i][t] = f4{0.f, 0.f, 0.f, 0.f};
  // local row and plane of tile i of the wavefront (the first row of a pair)
  auto tile_y = [&](int i) {
    const int tile = wave * TPW + i;
    return PAIR ? 2 * (tile % (Bk::BY / 2)) : tile % Bk::BY;
  };
  auto tile_z = [&](int i) {
    const int tile = wave * TPW + i;
    return PAIR ? tile / (Bk::BY / 2) : tile / Bk::BY;
  };
  // halo index (halves) of this lane's cell (x = m) in tile i, tap (-1,-1,-1)
  auto row_base = [&](int i) { return ((tile_z(i) * Bk::HY + tile_y(i)) * Bk::HX + m) * CIN; };
#pragma unroll
  for (int ch = 0; ch < CHUNKS; ++ch) {
    // this lane group's 8 K values: k = 32 ch + 8 g .. +7
    const int kk = ch * 32 + 8 * g;
    h8 w[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) w[t] = *reinterpret_cast<const h8*>(wl + (t * 16 + m) * KPL + kk);
    // halo offset of entry tp = (dz*4 + ry)*3 + dx of a row pair's K (input row ry from y - 1)
    auto pair_off = [&](int tp) { return (((tp / 12) * Bk::HY + (tp / 3) % 4) * Bk::HX + tp % 3) * CIN; };
    if (PAIR && CIN == 8) {
      // 8 channels of one entry
      const int off = pair_off(kk / CIN);
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        const h8 xv = *reinterpret_cast<const h8*>(halo + row_base(i) + off);
        acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w[0], xv, acc[i][0], 0, 0, 0);
      }
    } else if (PAIR) {
      // CIN == 4: two entries of 4 channels; entries from 36 on are K padding
      const int tp0 = kk / 4, tp1 = tp0 + 1;
      const bool live0 = tp0 < 36, live1 = tp1 < 36;
      const int off0 = live0 ? pair_off(tp0) : 0, off1 = live1 ? pair_off(tp1) : 0;
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        h4 lo = *reinterpret_cast<const h4*>(halo + row_base(i) + off0);
        h4 hi = *reinterpret_cast<const h4*>(halo + row_base(i) + off1);
        if (!live0) lo = h4{0, 0, 0, 0};
        if (!live1) hi = h4{0, 0, 0, 0};
        const h8 xv = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
        acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w[0], xv, acc[i][0], 0, 0, 0);
      }
    } else if (CIN >= 8) {
      // one tap, 8 consecutive channels
      const int tap = kk / CIN, ci = kk - tap * CIN;
      const bool live = tap < 27;  // K padding: zero operand
      const int off = (live ? tap_offset(tap, Bk::HX, Bk::HY) : 0) * CIN + ci;
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        h8 xv = *reinterpret_cast<const h8*>(halo + row_base(i) + off);
        if (!live) xv = h8(0);
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w[t], xv, acc[i][t], 0, 0, 0);
      }
    } else {
      // CIN == 4: two taps of 4 channels
      const int tap0 = kk / 4, tap1 = tap0 + 1;
      const bool live0 = tap0 < 27, live1 = tap1 < 27;
      const int off0 = (live0 ? tap_offset(tap0, Bk::HX, Bk::HY) : 0) * CIN;
      const int off1 = (live1 ? tap_offset(tap1, Bk::HX, Bk::HY) : 0) * CIN;
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        h4 lo = *reinterpret_cast<const h4*>(halo + row_base(i) + off0);
        h4 hi = *reinterpret_cast<const h4*>(halo + row_base(i) + off1);
        if (!live0) lo = h4{0, 0, 0, 0};
        if (!live1) hi = h4{0, 0, 0, 0};
        const h8 xv = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w[t], xv, acc[i][t], 0, 0, 0);
      }
    }
  }

  // epilogue: lane (m, g) holds D[channel t*16 + 4g + v][cell m]; in the
  // row-pair form channels 4(g & 1)..+3 of row y + (g >> 1)
  const int rsel = PAIR ? g >> 1 : 0, cg = PAIR ? (g & 1) : g;
  f4 bias[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    bias[t] = f4{0.f, 0.f, 0.f, 0.f};
    if (EPI == 1 && t == 0) bias[t][0] = a.bias[0];
    if (EPI == 0 && t * 16 + 4 * cg < COUT)
      bias[t] = *reinterpret_cast<const f4*>(a.bias + t * 16 + 4 * cg);
  }
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    const int y = y0 + tile_y(i), z = z0 + tile_z(i);
    if (y >= a.ny || z >= a.nz) continue;  // uniform per wavefront (ny is even: a pair is whole)
    const size_t cell = ((size_t)z * a.ny + y + rsel) * a.nx + x0 + m;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const f4 val = acc[i][t];
      if (EPI == 1) {
        if (cg == 0) a.yout[cell] = val[0] + bias[0][0];
        continue;
      }
      const int co = t * 16 + 4 * cg;
      if (co >= COUT) continue;
      h4 o;
      if (EPI == 2) {
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = (_Float16)val[e];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float u = val[e] + bias[t][e];
          o[e] = (_Float16)(u > 0.f ? u : 0.f);
        }
      }
      *reinterpret_cast<h4*>(a.out + cell * COUT + co) = o;
    }
  }
}

// ---------------------------------------------------------------- upsampling convolution
// The 3x3x3 convolution of the nearest-neighbour upsampled coarse input as a
// 2x2x2 convolution of the coarse input itself. Per axis, the three taps of
// an output cell of parity p (x = 2c + p) fall on two coarse cells,
// c - 1 + p + o for o = 0, 1: taps {-1} and {0, +1} when p = 0, {-1, 0} and
// {+1} when p = 1. One set of summed weights per parity (k_pack_up) gives
// K = 8 CIN instead of 27 CIN, and the LDS halo holds coarse cells.
//
// Output brick 32 x BY x 4 cells (BY = 8, or 4 for 32 input channels, whose
// weights take twice the registers); wavefront w takes the cells of parity
// (px, py, pz) = (w & 1, (w >> 1) & 1, w >> 2): a 16-lane row holds the
// cells x = x0 + 2m + px of one row (y, z), so every lane of a product uses
// the same weights, and the wavefront keeps its parity's weights in
// registers. Its rows: y = y0 + 2a + py (a < BY / 2), z = z0 + 2b + pz (b < 2).
// Epilogue: bias + skip + ReLU -> binary16.
template <int CIN>
struct UpBrick {
  static constexpr int BX = 32, BY = CIN >= 32 ? 4 : 8, BZ = 4;
  static constexpr int HX = BX / 2 + 2, HY = BY / 2 + 2, HZ = BZ / 2 + 2;
  static constexpr int HCELLS = HX * HY * HZ;
  static constexpr int AY = BY / 2;  // rows of one parity per plane
  static constexpr int TPW = AY * (BZ / 2);
};
static_assert(CWAVES == 8, "one wavefront per parity of the upsampling convolution");
// With at most 8 output channels a product covers both y parities: the A
// operand's rows 0-7 are parity py = 0, rows 8-15 py = 1, and K runs over
// (oz, ry, ox, ci) of the three coarse rows cy - 1 + ry the two rows read
// (py = 0: ry = oy, py = 1: ry = oy + 1): 12 CIN, 6 products per row pair
// against 4 per row. Wavefront w then takes parities (px, pz) = (w & 1,
// (w >> 1) & 1) and half w >> 2 of their row pairs.
__host__ __device__ constexpr bool up_pairs(int cout) { return cout <= 8; }
// binary16 values of an upsampling layer's packed weights
__host__ __device__ constexpr int up_weights(int cin, int cout) {
  return up_pairs(cout) ? 4 * 16 * 12 * cin : 8 * npad(cout) * 8 * cin;
}

template <int CIN, int COUT>
__global__ void __launch_bounds__(CTHREADS, FSP_CONV_MIN_WAVES) k_conv_up(ConvArgs a) {
  using Bk = UpBrick<CIN>;
  constexpr bool PY = up_pairs(COUT);
  constexpr int NT = npad(COUT) / 16;
  constexpr int NP = npad(COUT);
  constexpr int K = PY ? 12 * CIN : 8 * CIN;
  constexpr int CHUNKS = K / 32;
  constexpr int TPW = PY ? Bk::TPW / 2 : Bk::TPW;
  static_assert(CIN % 8 == 0 && K % 32 == 0 && COUT % 4 == 0 && (!PY || NT == 1), "channel groups");
  __shared__ __attribute__((aligned(16))) _Float16 halo[Bk::HCELLS * CIN];
  if (a.gate && !*a.gate) return;
  const int x0 = blockIdx.x * Bk::BX, y0 = blockIdx.y * Bk::BY, z0 = blockIdx.z * Bk::BZ;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  // parities of the wavefront (py from the lane group in the pair form) and its first tile
  const int px = wave & 1, py = PY ? g >> 1 : (wave >> 1) & 1, pz = PY ? (wave >> 1) & 1 : wave >> 2;
  const int pq = PY ? (wave & 3) : wave, j0 = PY ? (wave >> 2) * TPW : 0;

  // coarse halo (zero outside the coarse grid) -> LDS; origin (x0/2 - 1, y0/2 - 1, z0/2 - 1)
  {
    constexpr int VPC = CIN / 8;
    constexpr int TOTAL = Bk::HCELLS * VPC;
    constexpr int PER = (TOTAL + CTHREADS - 1) / CTHREADS;
    const int cx0 = x0 / 2 - 1, cy0 = y0 / 2 - 1, cz0 = z0 / 2 - 1;
    uint4 buf[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int q = (int)threadIdx.x + r * CTHREADS;
      const int cell = q / VPC, part = q - cell * VPC;
      const int lx = cell % Bk::HX, t = cell / Bk::HX, ly = t % Bk::HY, lz = t / Bk::HY;
      const int cx = cx0 + lx, cy = cy0 + ly, cz = cz0 + lz;
      const bool in = q < TOTAL && cx >= 0 && cx < a.cnx && cy >= 0 && cy < a.cny && cz >= 0 && cz < a.cnz;
      buf[r] = make_uint4(0u, 0u, 0u, 0u);
      if (in) buf[r] = *reinterpret_cast<const uint4*>(a.in + (((size_t)cz * a.cny + cy) * a.cnx + cx) * CIN + part * 8);
    }
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      const int q = (int)threadIdx.x + r * CTHREADS;
      if (q < TOTAL) *reinterpret_cast<uint4*>(halo + (size_t)q * 8) = buf[r];
    }
  }
  __syncthreads();

  // this parity's weights: lane (m, g) holds row t*16 + m, k = 32 ch + 8 g .. +7
  h8 w[CHUNKS][NT];
#pragma unroll
  for (int ch = 0; ch < CHUNKS; ++ch)
#pragma unroll
    for (int t = 0; t < NT; ++t)
      w[ch][t] = *reinterpret_cast<const h8*>(a.wt + (size_t)(pq * NP + t * 16 + m) * K + 32 * ch + 8 * g);
  // skip values of the epilogue (8 channels of a cell per item, see below),
  // loaded before the products
  constexpr int CELLS = Bk::BX * Bk::BY * Bk::BZ;
  constexpr int GROUPS = COUT / 8;  // 8-channel groups of a cell
  constexpr int IPT = CELLS * GROUPS / CTHREADS;
  static_assert(COUT % 8 == 0 && (CELLS * GROUPS) % CTHREADS == 0, "epilogue items");
  h8 skv[IPT];
  size_t ocell[IPT];
  bool olive[IPT];
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
    const int q = (int)threadIdx.x + r * CTHREADS, cl = q / GROUPS, part = q - cl * GROUPS;
    const int lx = cl % Bk::BX, t = cl / Bk::BX, ly = t % Bk::BY, lz = t / Bk::BY;
    const int y = y0 + ly, z = z0 + lz;
    olive[r] = y < a.ny && z < a.nz;
    ocell[r] = (((size_t)z * a.ny + y) * a.nx + x0 + lx) * COUT + part * 8;
    skv[r] = h8(0);
    if (olive[r]) skv[r] = *reinterpret_cast<const h8*>(a.skip + ocell[r]);
  }
  f4 acc[TPW][NT];
#pragma unroll
  for (int i = 0; i < TPW; ++i)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[i][t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ch = 0; ch < CHUNKS; ++ch) {
    // 8 consecutive channels of one coarse tap (ox, oy, oz); in the pair form
    // of entry (oz, ry, ox), coarse row cy - 1 + ry
    const int kk = 32 * ch + 8 * g;
    const int tap = kk / CIN, ci = kk - tap * CIN;
    const int ox = tap & 1, oz = PY ? tap / 6 : tap >> 2;
    const int ry = PY ? (tap >> 1) % 3 : py + ((tap >> 1) & 1);
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int j = j0 + i;
      const int hx = m + px + ox, hy = j % Bk::AY + ry, hz = j / Bk::AY + pz + oz;
      const h8 xv = *reinterpret_cast<const h8*>(halo + ((hz * Bk::HY + hy) * Bk::HX + hx) * CIN + ci);
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w[ch][t], xv, acc[i][t], 0, 0, 0);
    }
  }

  // epilogue through LDS: each wavefront stages acc + bias (f32) of its cells,
  // then a thread finishes 8 channels of a cell per item (+ skip, ReLU), so the
  // skip loads and the output stores are contiguous 16-byte accesses. Lane
  // (m, g) holds channels t*16 + 4g..+3, in the pair form 4(g & 1)..+3 of py.
  __shared__ __attribute__((aligned(16))) float stage[CELLS * COUT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = PY ? 4 * (g & 1) : t * 16 + 4 * g;
    if (co >= COUT) continue;
    const f4 bias = *reinterpret_cast<const f4*>(a.bias + co);
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int j = j0 + i;
      const int lx = 2 * m + px, ly = 2 * (j % Bk::AY) + py, lz = 2 * (j / Bk::AY) + pz;
      *reinterpret_cast<f4*>(stage + ((lz * Bk::BY + ly) * Bk::BX + lx) * COUT + co) = acc[i][t] + bias;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
    if (!olive[r]) continue;
    const int q = (int)threadIdx.x + r * CTHREADS;
    const f4 lo = *reinterpret_cast<const f4*>(stage + q * 8);
    const f4 hi = *reinterpret_cast<const f4*>(stage + q * 8 + 4);
    h8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float u = e < 4 ? lo[e] : hi[e - 4];
      u += (float)skv[r][e];
      o[e] = (_Float16)(u > 0.f ? u : 0.f);
    }
    *reinterpret_cast<h8*>(a.out + ocell[r]) = o;
  }
}

// ---------------------------------------------------------------- weight gradient
// Partial dW[co][k] = sum over the cells of one partition of g[cell][co] *
// in[cell + tap][ci] (k = tap*CIN + ci), one wavefront per (partition, k
// chunk, co tile); D = A B with A[co][cell] = g, B[cell][k] = shifted input.
struct WgradArgs {
  int nx, ny, nz, cnx, cny, cnz;
  const _Float16* g;    // [cells][gs]: gradient at the layer's pre-activation (channels 0..COUT-1)
  int gs;
  const _Float16* in;   // layer input
  float* partial;       // [parts][npad(COUT)][kpad(CIN)]
  int parts;
};

template <int CIN, int COUT, int UP>
__global__ void __launch_bounds__(THREADS) k_wgrad(WgradArgs a) {
  constexpr int MT = npad(COUT) / 16;
  constexpr int KP = kpad(CIN);
  constexpr int CHUNKS = KP / 16;
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * WAVES + (threadIdx.x >> 6);
  const int jobs = CHUNKS * MT;
  if (wave >= jobs * a.parts) return;
  const int job = wave % jobs, part = wave / jobs;
  const int ch = job % CHUNKS, mt = job / CHUNKS;
  const int m = lane & 15, g4 = lane >> 4;
  const int kk = ch * 16 + m;  // this lane's B column
  const int tap = kk / CIN, ci = kk - tap * CIN;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int co = mt * 16 + m;  // this lane's A row
  const int segs = a.nx / 16;
  const long total = (long)segs * a.ny * a.nz;
  const long per = (total + a.parts - 1) / a.parts;
  const long s0 = per * part, s1 = s0 + per < total ? s0 + per : total;
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  for (long s = s0; s < s1; ++s) {
    const int seg = (int)(s % segs);
    const long row = s / segs;
    const int y = (int)(row % a.ny), z = (int)(row / a.ny);
    const size_t rowbase = ((size_t)z * a.ny + y) * a.nx + seg * 16;
    h4 av, bv;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int xc = seg * 16 + 4 * g4 + v;
      av[v] = co < COUT ? a.g[(rowbase + 4 * g4 + v) * a.gs + co] : (_Float16)0;
      _Float16 b = 0;
      const int xx = xc + dx, yy = y + dy, zz = z + dz;
      if (tap < 27 && xx >= 0 && xx < a.nx && yy >= 0 && yy < a.ny && zz >= 0 && zz < a.nz) {
        size_t q;
        if (UP) q = ((size_t)(zz >> 1) * a.cny + (yy >> 1)) * a.cnx + (xx >> 1);
        else q = ((size_t)zz * a.ny + yy) * a.nx + xx;
        b = a.in[q * CIN + ci];
      }
      bv[v] = b;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv, acc, 0, 0, 0);
  }
  // lane holds D[co 4*g4+v][k m] of this (co tile, k chunk)
  float* P = a.partial + (size_t)part * npad(COUT) * KP;
#pragma unroll
  for (int v = 0; v < 4; ++v) P[(size_t)(mt * 16 + 4 * g4 + v) * KP + ch * 16 + m] = acc[v];
}

// Sum of the partitions in order; dW in the flat parameter layout, scaled
// by 1/scale; flags a non-finite value.
__global__ void k_wgrad_reduce(const float* partial, int parts, int cin, int cout, float inv_scale, float* grad,
                               int* bad) {
  const int KP = kpad(cin), NP = npad(cout);
  const int n = cout * 27 * cin;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int co = q / (27 * cin), k = q - co * 27 * cin;
    float s = 0.f;
    for (int p = 0; p < parts; ++p) s += partial[((size_t)p * NP + co) * KP + k];
    s *= inv_scale;
    if (!isfinite(s)) *bad = 1;
    grad[q] = s;
  }
}

// Bias gradient: sum over cells of g[cell][co], one workgroup per channel.
__global__ void __launch_bounds__(256) k_bgrad(const _Float16* g, size_t cells, int gs, float inv_scale, float* grad,
                                               int* bad) {
  __shared__ float red[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (size_t c = threadIdx.x; c < cells; c += 256) s += (float)g[c * gs + co];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float v = red[0] * inv_scale;
    if (!isfinite(v)) *bad = 1;
    grad[co] = v;
  }
}

// ---------------------------------------------------------------- elementwise
// Mean over each 2x2x2 block of a channels-last activation with C = 8 G channels.
// A thread averages one group of eight channels (16 bytes) of one coarse cell, in
// the order dz, dy, dx of the children. Block (64, 4): x runs over (coarse x,
// group), y and blockIdx.y over coarse rows (y, z).
template <int G>
__global__ void k_pool(const _Float16* in, _Float16* out, int cnx, int cny, int cnz, const int32_t* gate) {
  if (gate && !*gate) return;
  constexpr int C = 8 * G;
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= cnx * G) return;
  const int i = t / G, g = t % G;
  const int fnx = 2 * cnx, fny = 2 * cny, rows = cny * cnz;
  for (int row = blockIdx.y * 4 + threadIdx.y; row < rows; row += gridDim.y * 4) {
    const int j = row % cny, k = row / cny;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy) {
        const size_t f = ((size_t)(2 * k + dz) * fny + 2 * j + dy) * fnx + 2 * i;
        const h8 a = *reinterpret_cast<const h8*>(in + f * C + 8 * g);
        const h8 b = *reinterpret_cast<const h8*>(in + (f + 1) * C + 8 * g);
        for (int c = 0; c < 8; ++c) {
          s[c] += (float)a[c];
          s[c] += (float)b[c];
        }
      }
    h8 o;
    for (int c = 0; c < 8; ++c) o[c] = (_Float16)(s[c] * 0.125f);
    *reinterpret_cast<h8*>(out + ((size_t)row * cnx + i) * C + 8 * g) = o;
  }
}
// g_fine = g_coarse[parent] / 8 (+ add[fine])
__global__ void k_pool_back(const _Float16* gc, const _Float16* add, _Float16* gf, int fnx, int fny, int fnz, int C) {
  const size_t total = (size_t)fnx * fny * fnz * C;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    const size_t cell = q / C;
    const int i = (int)(cell % fnx), j = (int)((cell / fnx) % fny), k = (int)(cell / ((size_t)fnx * fny));
    const size_t p = ((size_t)(k >> 1) * (fny >> 1) + (j >> 1)) * (fnx >> 1) + (i >> 1);
    float v = (float)gc[p * C + c] * 0.125f;
    if (add) v += (float)add[q];
    gf[q] = (_Float16)v;
  }
}
// g_coarse = sum of the eight children's g_fine
__global__ void k_up_back(const _Float16* gf, _Float16* gc, int cnx, int cny, int cnz, int C) {
  const size_t total = (size_t)cnx * cny * cnz * C;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    const size_t cell = q / C;
    const int i = (int)(cell % cnx), j = (int)((cell / cnx) % cny), k = (int)(cell / ((size_t)cnx * cny));
    const int fnx = 2 * cnx, fny = 2 * cny;
    float s = 0.f;
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx)
          s += (float)gf[(((size_t)(2 * k + dz) * fny + 2 * j + dy) * fnx + 2 * i + dx) * C + c];
    gc[q] = (_Float16)s;
  }
}
// g = g * (h > 0)
__global__ void k_relu_back(const _Float16* h, _Float16* g, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    if (!((float)h[q] > 0.f)) g[q] = (_Float16)0.f;
}
// Loss: mean of (y - t)^2 over wet cells; per-block f64 partials; the
// output gradient scale * 2 (y - t) wet / n_wet as a 4-channel binary16
// tensor (channels 1-3 zero) for the head's backward-data convolution.
__global__ void __launch_bounds__(256) k_loss(const float* y, const float* t, const uint8_t* wet, size_t cells,
                                              double inv_n, float gscale, _Float16* gy, double* partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < cells; c += (size_t)gridDim.x * 256) {
    const float d = wet[c] ? y[c] - t[c] : 0.f;
    s += (double)d * d;
    const float gv = (float)(2.0 * inv_n) * d * gscale;
    gy[4 * c] = (_Float16)gv;
    gy[4 * c + 1] = (_Float16)0.f;
    gy[4 * c + 2] = (_Float16)0.f;
    gy[4 * c + 3] = (_Float16)0.f;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] * inv_n;
}

// Forward weights (W[co][tap][ci] -> [npad(cout)][kpad(cin)], k = tap*cin+ci;
// for row pairs [16][kfwd], row h*8 + co, k = ((dz*4 + ry)*3 + dx)*cin + ci
// holding W[co][dz][ry - h][dx][ci] where 0 <= ry - h <= 2, else 0; zero
// K padding)
// and backward-data weights (W'[ci][tap'][c] = W[c][26-tap'][ci], c over
// gch(cout) channels) in binary16, from the f32 master parameters.
__global__ void k_pack(const float* theta, int cin, int cout, _Float16* wt, _Float16* wb) {
  const int KP = kfwd(cin, cout), NP = npad(cout);
  const int gc = gch(cout), KPb = kpad(gc), NPb = npad(cin);
  const int nf = NP * KP, nb = NPb * KPb;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nf + nb; q += gridDim.x * blockDim.x) {
    if (q < nf && pair_rows(cin, cout)) {
      const int row = q / KP, k = q - row * KP, h = row >> 3, co = row & 7;
      const int tp = k / cin, ci = k - tp * cin, dz = tp / 12, dy = (tp / 3) % 4 - h, dx = tp % 3;
      float v = 0.f;
      if (co < cout && tp < 36 && dy >= 0 && dy <= 2) v = theta[((size_t)co * 27 + dz * 9 + dy * 3 + dx) * cin + ci];
      wt[q] = (_Float16)v;
    } else if (q < nf) {
      const int co = q / KP, k = q - co * KP;
      float v = 0.f;
      if (co < cout && k < 27 * cin) v = theta[(size_t)co * 27 * cin + k];
      wt[q] = (_Float16)v;
    } else if (wb) {
      const int r = q - nf;
      const int ci = r / KPb, k = r - ci * KPb;
      float v = 0.f;
      if (ci < cin && k < 27 * gc) {
        const int tp = k / gc, c = k - tp * gc;
        if (c < cout) v = theta[((size_t)c * 27 + (26 - tp)) * cin + ci];
      }
      wb[r] = (_Float16)v;
    }
  }
}

// Summed weights of an upsampling layer for k_conv_up: [8 parities][npad(cout)]
// [8 coarse taps x cin], coarse tap (ox, oy, oz) of parity (px, py, pz) the
// f32 sum of the layer's taps d (index d + 1 per axis) that fall on it:
// {0} and {1, 2} for p = 0, {0, 1} and {2} for p = 1; rounded once to binary16.
// With up_pairs: [4 parities (px, pz)][16 rows py*8 + co][12 entries x cin],
// entry (oz, ry, ox) holding coarse tap oy = ry - py (zero outside 0..1).
__global__ void k_pack_up(const float* theta, int cin, int cout, _Float16* wu) {
  const bool pairs = up_pairs(cout);
  const int NP = pairs ? 16 : npad(cout), K = (pairs ? 12 : 8) * cin, total = up_weights(cin, cout);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    const int P = q / (NP * K), r = q - P * NP * K, row = r / K, k = r - row * K;
    const int tap = k / cin, ci = k - tap * cin;
    int p[3], o[3], co = row;
    if (pairs) {
      co = row & 7;
      p[0] = P & 1; p[1] = row >> 3; p[2] = P >> 1;
      o[0] = tap & 1; o[1] = (tap >> 1) % 3 - p[1]; o[2] = tap / 6;
    } else {
      p[0] = P & 1; p[1] = (P >> 1) & 1; p[2] = P >> 2;
      o[0] = tap & 1; o[1] = (tap >> 1) & 1; o[2] = tap >> 2;
    }
    float v = 0.f;
    if (co < cout && o[1] >= 0 && o[1] <= 1) {
      int lo[3], hi[3];
      for (int d = 0; d < 3; ++d) {
        lo[d] = p[d] == 0 ? (o[d] == 0 ? 0 : 1) : (o[d] == 0 ? 0 : 2);
        hi[d] = p[d] == 0 ? (o[d] == 0 ? 0 : 2) : (o[d] == 0 ? 1 : 2);
      }
      for (int dz = lo[2]; dz <= hi[2]; ++dz)
        for (int dy = lo[1]; dy <= hi[1]; ++dy)
          for (int dx = lo[0]; dx <= hi[0]; ++dx) v += theta[((size_t)co * 27 + dz * 9 + dy * 3 + dx) * cin + ci];
    }
    wu[q] = (_Float16)v;
  }
}

// Adam (Kingma & Ba 2015) on the f32 master parameters; skipped entirely
// when a gradient was non-finite.
__global__ void k_adam(float* theta, float* m1, float* m2, const float* grad, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, const int* bad) {
  if (*bad) return;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const float gq = grad[q];
    const float a = b1 * m1[q] + (1.f - b1) * gq;
    const float b = b2 * m2[q] + (1.f - b2) * gq * gq;
    m1[q] = a;
    m2[q] = b;
    theta[q] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
  }
}

// Network input of a residual: (r * sigma, wet, ghost row, solid) in binary16,
// sigma = sqrt(n_wet / |r|^2) taken from the device (f64) or given. With VEC a
// thread converts four consecutive cells through 16-byte loads and stores (the
// cell count is a multiple of four); without it, one cell per step.
template <bool VEC>
__global__ void k_features(const float* r, const uint8_t* mask, const uint8_t* solid, size_t cells,
                           const double* rr, const uint32_t* nwet, float sigma_given, _Float16* x, float* sigma_out,
                           const int32_t* gate) {
  if (gate && !*gate) return;
  float sigma = sigma_given;
  if (rr) {
    const double s = *rr > 0.0 ? sqrt((double)(*nwet > 0 ? *nwet : 1u) / *rr) : 1.0;
    sigma = (float)s;
  }
  if (sigma_out && blockIdx.x == 0 && threadIdx.x == 0) *sigma_out = sigma;
  auto cell = [&](float rc, uint32_t m, uint32_t sd) {
    const bool wet = (m & 64u) != 0;
    h4 v;
    v[0] = (_Float16)(wet ? rc * sigma : 0.f);
    v[1] = (_Float16)(wet ? 1.f : 0.f);
    v[2] = (_Float16)((wet && (m & 128u)) ? 1.f : 0.f);
    v[3] = (_Float16)(sd ? 1.f : 0.f);
    return v;
  };
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  if (VEC) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < cells / 4; q += stride) {
      const f4 rv = reinterpret_cast<const f4*>(r)[q];
      const uint32_t m4 = reinterpret_cast<const uint32_t*>(mask)[q];
      const uint32_t s4 = solid ? reinterpret_cast<const uint32_t*>(solid)[q] : 0u;
      h4 v[4];
      for (int e = 0; e < 4; ++e) v[e] = cell(rv[e], (m4 >> (8 * e)) & 255u, (s4 >> (8 * e)) & 255u);
      h8* o = reinterpret_cast<h8*>(x + 16 * q);
      o[0] = h8{v[0][0], v[0][1], v[0][2], v[0][3], v[1][0], v[1][1], v[1][2], v[1][3]};
      o[1] = h8{v[2][0], v[2][1], v[2][2], v[2][3], v[3][0], v[3][1], v[3][2], v[3][3]};
    }
  } else {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += stride)
      *reinterpret_cast<h4*>(x + 4 * c) = cell(r[c], mask[c], solid ? solid[c] : 0u);
  }
}

// z = y / sigma on wet cells, 0 elsewhere; four cells per thread with VEC.
template <bool VEC>
__global__ void k_unscale(const float* y, const uint8_t* mask, const float* sigma, size_t cells, float* z,
                          const int32_t* gate) {
  if (gate && !*gate) return;
  const float s = *sigma;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  if (VEC) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < cells / 4; q += stride) {
      const f4 yv = reinterpret_cast<const f4*>(y)[q];
      const uint32_t m4 = reinterpret_cast<const uint32_t*>(mask)[q];
      f4 zv;
      for (int e = 0; e < 4; ++e) zv[e] = ((m4 >> (8 * e)) & 64u) ? yv[e] / s : 0.f;
      reinterpret_cast<f4*>(z)[q] = zv;
    }
  } else {
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += stride)
      z[c] = (mask[c] & 64u) ? y[c] / s : 0.f;
  }
}

// One 16x16x16 product, for checking the MFMA operand layout.
__global__ void k_mfma_probe(const _Float16* A, const _Float16* B, float* D) {
  const int l = threadIdx.x, m = l & 15, g = l >> 4;
  h4 a, b;
  for (int v = 0; v < 4; ++v) {
    a[v] = A[m * 16 + 4 * g + v];
    b[v] = B[(4 * g + v) * 16 + m];
  }
  f4 d = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  for (int v = 0; v < 4; ++v) D[(4 * g + v) * 16 + m] = d[v];
}

}  // namespace net
}  // namespace fsp

// ================================================================ host side
namespace fsp {
namespace net {

#define NET_HIP(x)                          \
  do {                                      \
    if ((x) != hipSuccess) return FSP_DEVICE; \
  } while (0)

inline unsigned blocks_for(size_t n, unsigned threads = 256, unsigned cap = 4096) {
  size_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  return (unsigned)(b < cap ? b : cap);
}

// One thread per item (up to 2^24 blocks): for elementwise kernels, whose time a
// capped grid-stride loop lengthens by serialising each thread's steps.
inline unsigned blocks_per_item(size_t n, unsigned threads = 256) { return blocks_for(n, threads, 1u << 24); }

// Whether four cells per thread fit the pointers: a16/b16 (four floats or four f16
// quadruples) on 16 bytes, a4/b4 (four byte masks) on 4 bytes, a cell count that
// is a multiple of four.
inline bool quad_cells(size_t cells, const void* a16, const void* b16, const void* a4, const void* b4) {
  auto al = [](const void* p, uintptr_t a) { return p == nullptr || ((uintptr_t)p & (a - 1)) == 0; };
  return cells % 4 == 0 && al(a16, 16) && al(b16, 16) && al(a4, 4) && al(b4, 4);
}

}  // namespace net
}  // namespace fsp

using namespace fsp::net;

struct fsp_net {
  int device = 0;
  hipStream_t stream = nullptr;  // the network's own stream
  hipStream_t cur = nullptr;     // stream of the call in progress
  const int32_t* gate = nullptr; // gate of the forward pass in progress
  std::vector<void*> allocs;
  size_t nparam_allocs = 0;  // leading allocations that belong to the parameters (kept across shapes)
  // parameters
  size_t P = 0;
  float *theta = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
  _Float16* wt[NL] = {};
  _Float16* wb[NL] = {};
  _Float16* wu[NL] = {};  // summed per-parity weights of the upsampling layers (k_conv_up)
  size_t poff[NL] = {};  // offset of layer l's W in theta (b follows)
  int* bad = nullptr;
  int* hbad = nullptr;
  double* hloss = nullptr;
  uint64_t step = 0;
  float gscale = 0.f;
  // activations for one shape
  int nx = 0, ny = 0, nz = 0;
  uint64_t generation = 0;  // bumped whenever the shape buffers are reallocated
  size_t V[3] = {};
  _Float16* x0 = nullptr;
  _Float16* act[NL] = {};
  float* y = nullptr;
  _Float16 *p1 = nullptr, *p2 = nullptr;
  // backward
  _Float16* gpre[NL] = {};  // gradient at each layer's pre-activation
  _Float16 *gy = nullptr, *gf16 = nullptr, *gf32 = nullptr, *gp2 = nullptr, *gp1 = nullptr;
  float* partial = nullptr;
  size_t partial_floats = 0;
  double* lpart = nullptr;
  float* sigma = nullptr;
  uint8_t* wet = nullptr;
};

namespace {

template <class T>
fsp_status nalloc(fsp_net* N, T** out, size_t count) {
  void* p = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 256;
  if (hipMalloc(&p, bytes) != hipSuccess) return FSP_DEVICE;
  if (hipMemsetAsync(p, 0, bytes, N->stream) != hipSuccess) return FSP_DEVICE;
  N->allocs.push_back(p);
  *out = static_cast<T*>(p);
  return FSP_OK;
}

void free_shape(fsp_net* N, size_t keep) {
  while (N->allocs.size() > keep) {
    (void)hipFree(N->allocs.back());
    N->allocs.pop_back();
  }
}

template <int CIN, int COUT, int EPI>
void conv(hipStream_t st, int nx, int ny, int nz, const _Float16* in, const _Float16* wt, const float* bias,
          _Float16* out, float* yout, const int32_t* gate = nullptr) {
  ConvArgs a{};
  a.gate = gate;
  a.nx = nx; a.ny = ny; a.nz = nz;
  a.cnx = nx; a.cny = ny; a.cnz = nz;
  a.in = in; a.wt = wt; a.bias = bias; a.out = out; a.yout = yout;
  using Bk = ConvBrick<CIN, COUT>;
  const dim3 grid((unsigned)(nx / Bk::BX), (unsigned)((ny + Bk::BY - 1) / Bk::BY),
                  (unsigned)((nz + Bk::BZ - 1) / Bk::BZ));
  k_conv<CIN, COUT, EPI><<<grid, CTHREADS, 0, st>>>(a);
}

// Upsampling layer (bias + skip + ReLU) on an output grid nx x ny x nz (nx % 32 == 0,
// ny and nz even) from the coarse grid of half its size; wu from k_pack_up.
template <int CIN, int COUT>
void conv_up(hipStream_t st, int nx, int ny, int nz, const _Float16* in, const _Float16* wu, const float* bias,
             const _Float16* skip, _Float16* out, const int32_t* gate) {
  ConvArgs a{};
  a.nx = nx; a.ny = ny; a.nz = nz;
  a.cnx = nx / 2; a.cny = ny / 2; a.cnz = nz / 2;
  a.in = in; a.wt = wu; a.bias = bias; a.skip = skip; a.out = out; a.gate = gate;
  using Bk = UpBrick<CIN>;
  const dim3 grid((unsigned)(nx / Bk::BX), (unsigned)((ny + Bk::BY - 1) / Bk::BY), (unsigned)((nz + Bk::BZ - 1) / Bk::BZ));
  k_conv_up<CIN, COUT><<<grid, CTHREADS, 0, st>>>(a);
}

// dW of layer l (g: gradient at its pre-activation with channel stride gs)
template <int CIN, int COUT, int UP>
void wgrad(fsp_net* N, int l, int nx, int ny, int nz, const _Float16* g, const _Float16* in, int gs = COUT) {
  constexpr int MT = npad(COUT) / 16, CH = kpad(CIN) / 16;
  const size_t segs = (size_t)(nx / 16) * ny * nz;
  int parts = (int)((8192 + CH * MT - 1) / (CH * MT));
  if ((size_t)parts > segs) parts = (int)segs;
  if ((size_t)parts * npad(COUT) * kpad(CIN) > N->partial_floats) parts = (int)(N->partial_floats / (npad(COUT) * kpad(CIN)));
  if (parts < 1) parts = 1;
  WgradArgs a{};
  a.nx = nx; a.ny = ny; a.nz = nz;
  a.cnx = UP ? nx / 2 : nx; a.cny = UP ? ny / 2 : ny; a.cnz = UP ? nz / 2 : nz;
  a.g = g; a.gs = gs; a.in = in; a.partial = N->partial; a.parts = parts;
  const size_t waves = (size_t)CH * MT * parts;
  k_wgrad<CIN, COUT, UP><<<(unsigned)((waves + WAVES - 1) / WAVES), THREADS, 0, N->cur>>>(a);
  const float inv = 1.0f / N->gscale;
  k_wgrad_reduce<<<blocks_for((size_t)COUT * 27 * CIN), 256, 0, N->cur>>>(N->partial, parts, CIN, COUT, inv,
                                                                             N->grad + N->poff[l], N->bad);
  k_bgrad<<<COUT, 256, 0, N->cur>>>(g, (size_t)nx * ny * nz, gs, inv, N->grad + N->poff[l] + (size_t)COUT * 27 * CIN,
                                       N->bad);
}

int lvl_n(const fsp_net* N, int level, int axis) {
  const int d = axis == 0 ? N->nx : axis == 1 ? N->ny : N->nz;
  return d >> level;
}

// The buffers of one forward pass: input features, activations, pooled
// levels and the output.
struct Acts {
  int nx, ny, nz;
  const _Float16* x0;   // [cells][4]
  _Float16* act[NL];    // layer outputs (the head writes y)
  _Float16 *p1, *p2;    // pooled level-1 and level-2 inputs
  float* y;             // [cells]
};

Acts net_acts(fsp_net* N) {
  Acts A{};
  A.nx = N->nx; A.ny = N->ny; A.nz = N->nz;
  A.x0 = N->x0;
  for (int l = 0; l < NL; ++l) A.act[l] = N->act[l];
  A.p1 = N->p1; A.p2 = N->p2; A.y = N->y;
  return A;
}

// k_pool's grid over coarse rows of cnx cells and g channel groups.
dim3 pool_grid(int cnx, int cny, int cnz, int g) {
  const unsigned rows = (unsigned)cny * (unsigned)cnz;
  return dim3(((unsigned)cnx * g + 63) / 64, std::min<unsigned>((rows + 3) / 4, 65535u));
}

void forward(fsp_net* N, const Acts& A) {
  hipStream_t st = N->cur;
  const int32_t* gt = N->gate;
  const int X0 = A.nx, Y0 = A.ny, Z0 = A.nz;
  const int X1 = X0 / 2, Y1 = Y0 / 2, Z1 = Z0 / 2, X2 = X0 / 4, Y2 = Y0 / 4, Z2 = Z0 / 4;
  const size_t V1 = (size_t)X1 * Y1 * Z1, V2 = (size_t)X2 * Y2 * Z2;
  auto B = [&](int l) { return N->theta + N->poff[l] + (size_t)LAYERS[l].cout * 27 * LAYERS[l].cin; };
  conv<4, 8, 0>(st, X0, Y0, Z0, A.x0, N->wt[0], B(0), A.act[0], nullptr, gt);
  conv<8, 8, 0>(st, X0, Y0, Z0, A.act[0], N->wt[1], B(1), A.act[1], nullptr, gt);
  k_pool<1><<<pool_grid(X1, Y1, Z1, 1), dim3(64, 4), 0, st>>>(A.act[1], A.p1, X1, Y1, Z1, gt);
  conv<8, 16, 0>(st, X1, Y1, Z1, A.p1, N->wt[2], B(2), A.act[2], nullptr, gt);
  conv<16, 16, 0>(st, X1, Y1, Z1, A.act[2], N->wt[3], B(3), A.act[3], nullptr, gt);
  k_pool<2><<<pool_grid(X2, Y2, Z2, 2), dim3(64, 4), 0, st>>>(A.act[3], A.p2, X2, Y2, Z2, gt);
  conv<16, 32, 0>(st, X2, Y2, Z2, A.p2, N->wt[4], B(4), A.act[4], nullptr, gt);
  conv<32, 32, 0>(st, X2, Y2, Z2, A.act[4], N->wt[5], B(5), A.act[5], nullptr, gt);
  conv_up<32, 16>(st, X1, Y1, Z1, A.act[5], N->wu[6], B(6), A.act[3], A.act[6], gt);
  conv<16, 16, 0>(st, X1, Y1, Z1, A.act[6], N->wt[7], B(7), A.act[7], nullptr, gt);
  conv_up<16, 8>(st, X0, Y0, Z0, A.act[7], N->wu[8], B(8), A.act[1], A.act[8], gt);
  conv<8, 8, 0>(st, X0, Y0, Z0, A.act[8], N->wt[9], B(9), A.act[9], nullptr, gt);
  conv<8, 1, 1>(st, X0, Y0, Z0, A.act[9], N->wt[10], B(10), nullptr, A.y, gt);
}

void forward(fsp_net* N) { forward(N, net_acts(N)); }

void relu_back(fsp_net* N, const _Float16* h, _Float16* g, size_t n) {
  k_relu_back<<<blocks_for(n), 256, 0, N->cur>>>(h, g, n);
}

// Gradients of the loss (its output gradient already in N->gy) into N->grad.
void backward(fsp_net* N) {
  hipStream_t st = N->cur;
  const int X0 = N->nx, Y0 = N->ny, Z0 = N->nz;
  const int X1 = X0 / 2, Y1 = Y0 / 2, Z1 = Z0 / 2, X2 = X0 / 4, Y2 = Y0 / 4, Z2 = Z0 / 4;
  const size_t V0 = N->V[0], V1 = N->V[1], V2 = N->V[2];
  _Float16** g = N->gpre;
  // head (gradient gy: 4 channels, channel 0 used)
  wgrad<8, 1, 0>(N, 10, X0, Y0, Z0, N->gy, N->act[9], 4);
  conv<4, 8, 2>(st, X0, Y0, Z0, N->gy, N->wb[10], nullptr, g[9], nullptr);
  relu_back(N, N->act[9], g[9], V0 * 8);
  wgrad<8, 8, 0>(N, 9, X0, Y0, Z0, g[9], N->act[8]);
  conv<8, 8, 2>(st, X0, Y0, Z0, g[9], N->wb[9], nullptr, g[8], nullptr);
  relu_back(N, N->act[8], g[8], V0 * 8);
  wgrad<16, 8, 1>(N, 8, X0, Y0, Z0, g[8], N->act[7]);
  conv<8, 16, 2>(st, X0, Y0, Z0, g[8], N->wb[8], nullptr, N->gf16, nullptr);
  k_up_back<<<blocks_for(V1 * 16), 256, 0, st>>>(N->gf16, g[7], X1, Y1, Z1, 16);
  relu_back(N, N->act[7], g[7], V1 * 16);
  wgrad<16, 16, 0>(N, 7, X1, Y1, Z1, g[7], N->act[6]);
  conv<16, 16, 2>(st, X1, Y1, Z1, g[7], N->wb[7], nullptr, g[6], nullptr);
  relu_back(N, N->act[6], g[6], V1 * 16);
  wgrad<32, 16, 1>(N, 6, X1, Y1, Z1, g[6], N->act[5]);
  conv<16, 32, 2>(st, X1, Y1, Z1, g[6], N->wb[6], nullptr, N->gf32, nullptr);
  k_up_back<<<blocks_for(V2 * 32), 256, 0, st>>>(N->gf32, g[5], X2, Y2, Z2, 32);
  relu_back(N, N->act[5], g[5], V2 * 32);
  wgrad<32, 32, 0>(N, 5, X2, Y2, Z2, g[5], N->act[4]);
  conv<32, 32, 2>(st, X2, Y2, Z2, g[5], N->wb[5], nullptr, g[4], nullptr);
  relu_back(N, N->act[4], g[4], V2 * 32);
  wgrad<16, 32, 0>(N, 4, X2, Y2, Z2, g[4], N->p2);
  conv<32, 16, 2>(st, X2, Y2, Z2, g[4], N->wb[4], nullptr, N->gp2, nullptr);
  k_pool_back<<<blocks_for(V1 * 16), 256, 0, st>>>(N->gp2, g[6], g[3], X1, Y1, Z1, 16);
  relu_back(N, N->act[3], g[3], V1 * 16);
  wgrad<16, 16, 0>(N, 3, X1, Y1, Z1, g[3], N->act[2]);
  conv<16, 16, 2>(st, X1, Y1, Z1, g[3], N->wb[3], nullptr, g[2], nullptr);
  relu_back(N, N->act[2], g[2], V1 * 16);
  wgrad<8, 16, 0>(N, 2, X1, Y1, Z1, g[2], N->p1);
  conv<16, 8, 2>(st, X1, Y1, Z1, g[2], N->wb[2], nullptr, N->gp1, nullptr);
  k_pool_back<<<blocks_for(V0 * 8), 256, 0, st>>>(N->gp1, g[8], g[1], X0, Y0, Z0, 8);
  relu_back(N, N->act[1], g[1], V0 * 8);
  wgrad<8, 8, 0>(N, 1, X0, Y0, Z0, g[1], N->act[0]);
  conv<8, 8, 2>(st, X0, Y0, Z0, g[1], N->wb[1], nullptr, g[0], nullptr);
  relu_back(N, N->act[0], g[0], V0 * 8);
  wgrad<4, 8, 0>(N, 0, X0, Y0, Z0, g[0], N->x0);
}

}  // namespace

namespace fsp {

bool net_shape_ok(int nx, int ny, int nz) { return nx > 0 && ny > 0 && nz > 0 && nx % 64 == 0 && ny % 4 == 0 && nz % 4 == 0; }

fsp_status net_reserve(fsp_net* N, int nx, int ny, int nz) {
  if (!net_shape_ok(nx, ny, nz)) return FSP_SHAPE;
  if (N->nx == nx && N->ny == ny && N->nz == nz) return FSP_OK;
  NET_HIP(hipSetDevice(N->device));
  NET_HIP(hipStreamSynchronize(N->stream));
  free_shape(N, N->nparam_allocs);
  N->nx = N->ny = N->nz = 0;
  const size_t V0 = (size_t)nx * ny * nz, V1 = V0 / 8, V2 = V0 / 64;
  N->V[0] = V0; N->V[1] = V1; N->V[2] = V2;
  fsp_status st;
  if ((st = nalloc(N, &N->x0, V0 * 4)) || (st = nalloc(N, &N->y, V0)) || (st = nalloc(N, &N->p1, V1 * 8)) ||
      (st = nalloc(N, &N->p2, V2 * 16)) || (st = nalloc(N, &N->gy, V0 * 4)) || (st = nalloc(N, &N->gf16, V0 * 16)) ||
      (st = nalloc(N, &N->gf32, V1 * 32)) || (st = nalloc(N, &N->gp2, V2 * 16)) || (st = nalloc(N, &N->gp1, V1 * 8)) ||
      (st = nalloc(N, &N->sigma, 1)) || (st = nalloc(N, &N->wet, V0)) || (st = nalloc(N, &N->lpart, 4096)))
    return st;
  for (int l = 0; l < NL; ++l) {
    const size_t Vl = N->V[LAYERS[l].level];
    if (l < NL - 1 && (st = nalloc(N, &N->act[l], Vl * LAYERS[l].cout))) return st;
    if ((st = nalloc(N, &N->gpre[l], Vl * gch(LAYERS[l].cout)))) return st;
  }
  // weight-gradient partials: up to 8192 wavefronts' worth, at least one partition per layer
  size_t most = 0;
  for (const LayerDesc& L : LAYERS) most = std::max(most, (size_t)npad(L.cout) * kpad(L.cin));
  N->partial_floats = std::max(most * 64, (size_t)8192 * 256);
  if ((st = nalloc(N, &N->partial, N->partial_floats))) return st;
  N->nx = nx; N->ny = ny; N->nz = nz;
  ++N->generation;
  return FSP_OK;
}

uint64_t net_generation(const fsp_net* N) { return N ? N->generation : 0; }

fsp_status net_features(fsp_net* N, hipStream_t stream, const float* r, const uint8_t* mask, const uint8_t* solid,
                        const double* rr, const uint32_t* nwet, float sigma, uint16_t* x_out) {
  const size_t V = N->V[0];
  _Float16* x = reinterpret_cast<_Float16*>(x_out);
  if (quad_cells(V, r, x, mask, solid))
    k_features<true><<<blocks_per_item(V / 4), 256, 0, stream>>>(r, mask, solid, V, rr, nwet, sigma, x, N->sigma, nullptr);
  else
    k_features<false><<<blocks_for(V), 256, 0, stream>>>(r, mask, solid, V, rr, nwet, sigma, x, N->sigma, nullptr);
  return hipGetLastError() == hipSuccess ? FSP_OK : FSP_DEVICE;
}

fsp_status net_apply(fsp_net* N, hipStream_t stream, const float* r, const uint8_t* mask, const uint8_t* solid,
                     const double* rr, const uint32_t* nwet, const int32_t* gate, float* z) {
  const size_t V = N->V[0];
  if (quad_cells(V, r, N->x0, mask, solid))
    k_features<true><<<blocks_per_item(V / 4), 256, 0, stream>>>(r, mask, solid, V, rr, nwet, 0.f, N->x0, N->sigma, gate);
  else
    k_features<false><<<blocks_for(V), 256, 0, stream>>>(r, mask, solid, V, rr, nwet, 0.f, N->x0, N->sigma, gate);
  N->cur = stream;
  N->gate = gate;
  forward(N);
  N->gate = nullptr;
  if (quad_cells(V, N->y, z, mask, nullptr))
    k_unscale<true><<<blocks_per_item(V / 4), 256, 0, stream>>>(N->y, mask, N->sigma, V, z, gate);
  else
    k_unscale<false><<<blocks_for(V), 256, 0, stream>>>(N->y, mask, N->sigma, V, z, gate);
  return hipGetLastError() == hipSuccess ? FSP_OK : FSP_DEVICE;
}

}  // namespace fsp

namespace {

fsp_status pack_all(fsp_net* N) {
  for (int l = 0; l < NL; ++l) {
    const LayerDesc& L = LAYERS[l];
    const size_t total = (size_t)npad(L.cout) * kfwd(L.cin, L.cout) + (size_t)npad(L.cin) * kpad(gch(L.cout));
    k_pack<<<blocks_for(total), 256, 0, N->cur>>>(N->theta + N->poff[l], L.cin, L.cout, N->wt[l], l > 0 ? N->wb[l] : nullptr);
    if (L.up)
      k_pack_up<<<blocks_for((size_t)up_weights(L.cin, L.cout)), 256, 0, N->cur>>>(N->theta + N->poff[l], L.cin,
                                                                                    L.cout, N->wu[l]);
  }
  return hipGetLastError() == hipSuccess ? FSP_OK : FSP_DEVICE;
}

// loss (host) and gradients into N->grad for the current x0 / target / wet
fsp_status loss_and_grad(fsp_net* N, const float* target, const uint8_t* wet, double* loss, bool* bad) {
  const size_t V = N->V[0];
  NET_HIP(hipMemsetAsync(N->bad, 0, sizeof(int), N->cur));
  forward(N);
  // wet count for the mean
  std::vector<uint8_t> hw;
  size_t nwet = 0;
  {
    hw.resize(V);
    NET_HIP(hipMemcpyAsync(hw.data(), wet, V, hipMemcpyDeviceToHost, N->cur));
    NET_HIP(hipStreamSynchronize(N->cur));
    for (uint8_t b : hw) nwet += b ? 1 : 0;
  }
  if (nwet == 0) nwet = 1;
  if (N->gscale <= 0.f) N->gscale = std::ldexp(1.0f, (int)std::lround(std::log2((double)nwet)));
  const unsigned blocks = std::min<unsigned>(4096, blocks_for(V));
  k_loss<<<blocks, 256, 0, N->cur>>>(N->y, target, wet, V, 1.0 / (double)nwet, N->gscale, N->gy, N->lpart);
  backward(N);
  std::vector<double> lp(blocks);
  NET_HIP(hipMemcpyAsync(lp.data(), N->lpart, blocks * sizeof(double), hipMemcpyDeviceToHost, N->cur));
  NET_HIP(hipMemcpyAsync(N->hbad, N->bad, sizeof(int), hipMemcpyDeviceToHost, N->cur));
  NET_HIP(hipStreamSynchronize(N->cur));
  double s = 0.0;
  for (double v : lp) s += v;
  if (loss) *loss = s;
  *bad = *N->hbad != 0;
  return FSP_OK;
}

// ---------------------------------------------------------------- forward pass that reserves nothing
// Activations of fsp_net_forward_static_device: a file-scope device array
// of STATIC_BYTES, one copy per GPU, held until the process ends and shared
// by every such call on that GPU (calls that overlap on one GPU race). The
// largest grid it holds is STATIC_MAX_CELLS cells (84.5 bytes of
// activations per cell).
constexpr size_t STATIC_BYTES = (size_t)24 << 20;
constexpr size_t STATIC_MAX_CELLS = 262144;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[STATIC_BYTES];

// The device address of `scratch` on each GPU (host code never passes the
// array's name to a kernel: on the host it is a stand-in, not a device
// address), taken once per device with that device current.
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};

fsp_status scratch_address(int device) {
  if (device < 0 || device >= MAX_DEVICES) return FSP_DEVICE;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[device]) return FSP_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    std::fprintf(stderr, "fsp: hipGetSymbolAddress(scratch) on device %d failed: %s\n", device, hipGetErrorString(e));
    return FSP_DEVICE;
  }
  scratch_ptr[device] = p;
  return FSP_OK;
}

// Activation buffers of an nx x ny x nz forward pass carved from `base`
// (256-byte aligned, in layer order); returns the bytes used.
size_t carve_acts(unsigned char* base, int nx, int ny, int nz, Acts* A) {
  const size_t V[3] = {(size_t)nx * ny * nz, (size_t)nx * ny * nz / 8, (size_t)nx * ny * nz / 64};
  size_t off = 0;
  auto take = [&](size_t halves) {
    _Float16* q = base ? reinterpret_cast<_Float16*>(base + off) : nullptr;
    off += (halves * sizeof(_Float16) + 255) & ~(size_t)255;
    return q;
  };
  A->nx = nx; A->ny = ny; A->nz = nz;
  for (int l = 0; l < NL; ++l)
    A->act[l] = l < NL - 1 ? take(V[LAYERS[l].level] * LAYERS[l].cout) : nullptr;
  A->p1 = take(V[1] * 8);
  A->p2 = take(V[2] * 16);
  return off;
}

}  // namespace

extern "C" {

uint32_t fsp_net_param_count(void) { return (uint32_t)param_count(); }

fsp_status fsp_net_create(int32_t device, fsp_net** out) {
  if (!out) return FSP_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return FSP_DEVICE;
  fsp_net* N = new (std::nothrow) fsp_net();
  if (!N) return FSP_DEVICE;
  N->device = device;
  auto fail = [&](fsp_status s) { fsp_net_destroy(N); return s; };
  if (hipSetDevice(device) != hipSuccess) return fail(FSP_DEVICE);
  // the static activation array's address on this device, for the forward
  // pass that reserves nothing
  if (const fsp_status ss = scratch_address(device)) return fail(ss);
  if (hipStreamCreateWithFlags(&N->stream, hipStreamNonBlocking) != hipSuccess) return fail(FSP_DEVICE);
  N->cur = N->stream;
  N->P = param_count();
  size_t off = 0;
  for (int l = 0; l < NL; ++l) {
    N->poff[l] = off;
    off += (size_t)LAYERS[l].cout * 27 * LAYERS[l].cin + LAYERS[l].cout;
  }
  fsp_status st;
  if ((st = nalloc(N, &N->theta, N->P)) || (st = nalloc(N, &N->m1, N->P)) || (st = nalloc(N, &N->m2, N->P)) ||
      (st = nalloc(N, &N->grad, N->P)) || (st = nalloc(N, &N->bad, 1)))
    return fail(st);
  for (int l = 0; l < NL; ++l) {
    const LayerDesc& L = LAYERS[l];
    if ((st = nalloc(N, &N->wt[l], (size_t)npad(L.cout) * kfwd(L.cin, L.cout))) ||
        (st = nalloc(N, &N->wb[l], (size_t)npad(L.cin) * kpad(gch(L.cout)))) ||
        (L.up && (st = nalloc(N, &N->wu[l], (size_t)up_weights(L.cin, L.cout)))))
      return fail(st);
  }
  N->nparam_allocs = N->allocs.size();
  if (hipHostMalloc((void**)&N->hbad, sizeof(int), hipHostMallocDefault) != hipSuccess) return fail(FSP_DEVICE);
  if (hipStreamSynchronize(N->stream) != hipSuccess) return fail(FSP_DEVICE);
  *out = N;
  return FSP_OK;
}

void fsp_net_destroy(fsp_net* N) {
  if (!N) return;
  (void)hipSetDevice(N->device);
  if (N->stream) (void)hipStreamSynchronize(N->stream);
  free_shape(N, 0);
  if (N->hbad) (void)hipHostFree(N->hbad);
  if (N->stream) (void)hipStreamDestroy(N->stream);
  delete N;
}

fsp_status fsp_net_set_params(fsp_net* N, const float* theta, size_t n) {
  if (!N || !theta) return FSP_INVALID;
  if (n != N->P) return FSP_SHAPE;
  for (size_t q = 0; q < n; ++q)
    if (!std::isfinite(theta[q])) return FSP_NOT_FINITE;
  NET_HIP(hipSetDevice(N->device));
  N->cur = N->stream;
  NET_HIP(hipMemcpyAsync(N->theta, theta, n * sizeof(float), hipMemcpyHostToDevice, N->stream));
  NET_HIP(hipMemsetAsync(N->m1, 0, n * sizeof(float), N->stream));
  NET_HIP(hipMemsetAsync(N->m2, 0, n * sizeof(float), N->stream));
  N->step = 0;
  N->gscale = 0.f;
  fsp_status st = pack_all(N);
  if (st) return st;
  NET_HIP(hipStreamSynchronize(N->stream));
  return FSP_OK;
}

fsp_status fsp_net_get_params(fsp_net* N, float* theta, size_t n) {
  if (!N || !theta) return FSP_INVALID;
  if (n != N->P) return FSP_SHAPE;
  NET_HIP(hipSetDevice(N->device));
  NET_HIP(hipMemcpyAsync(theta, N->theta, n * sizeof(float), hipMemcpyDeviceToHost, N->stream));
  NET_HIP(hipStreamSynchronize(N->stream));
  return FSP_OK;
}

fsp_status fsp_net_reserve(fsp_net* N, int32_t nx, int32_t ny, int32_t nz) {
  if (!N) return FSP_INVALID;
  return fsp::net_reserve(N, nx, ny, nz);
}

fsp_status fsp_net_forward_device(fsp_net* N, int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_x, float* d_y) {
  if (!N || !d_x || !d_y) return FSP_INVALID;
  NET_HIP(hipSetDevice(N->device));
  fsp_status st = fsp::net_reserve(N, nx, ny, nz);
  if (st) return st;
  N->cur = N->stream;
  NET_HIP(hipMemcpyAsync(N->x0, d_x, N->V[0] * 4 * sizeof(uint16_t), hipMemcpyDeviceToDevice, N->stream));
  forward(N);
  NET_HIP(hipMemcpyAsync(d_y, N->y, N->V[0] * sizeof(float), hipMemcpyDeviceToDevice, N->stream));
  NET_HIP(hipStreamSynchronize(N->stream));
  return FSP_OK;
}

fsp_status fsp_net_forward_static_device(fsp_net* N, int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_x,
                                         float* d_y) {
  if (!N || !d_x || !d_y) return FSP_INVALID;
  if (!fsp::net_shape_ok(nx, ny, nz)) return FSP_SHAPE;
  Acts A{};
  const size_t cells = (size_t)nx * ny * nz;
  if (cells > STATIC_MAX_CELLS || carve_acts(nullptr, nx, ny, nz, &A) > STATIC_BYTES) {
    std::fprintf(stderr, "fsp: fsp_net_forward_static_device: %d x %d x %d = %zu cells exceed the static "
                 "activation array (%zu cells at most)\n", nx, ny, nz, cells, STATIC_MAX_CELLS);
    return FSP_SHAPE;
  }
  NET_HIP(hipSetDevice(N->device));
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    base = static_cast<unsigned char*>(scratch_ptr[N->device]);
  }
  if (!base) return FSP_DEVICE;
  carve_acts(base, nx, ny, nz, &A);
  A.x0 = reinterpret_cast<const _Float16*>(d_x);
  A.y = d_y;
  N->cur = N->stream;
  N->gate = nullptr;
  forward(N, A);
  NET_HIP(hipGetLastError());
  NET_HIP(hipStreamSynchronize(N->stream));
  return FSP_OK;
}

fsp_status fsp_net_loss_grad_device(fsp_net* N, int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_x,
                                    const float* d_target, const uint8_t* d_wet, double* loss, float* grad) {
  if (!N || !d_x || !d_target || !d_wet) return FSP_INVALID;
  NET_HIP(hipSetDevice(N->device));
  fsp_status st = fsp::net_reserve(N, nx, ny, nz);
  if (st) return st;
  N->cur = N->stream;
  NET_HIP(hipMemcpyAsync(N->x0, d_x, N->V[0] * 4 * sizeof(uint16_t), hipMemcpyDeviceToDevice, N->stream));
  bool bad = false;
  if ((st = loss_and_grad(N, d_target, d_wet, loss, &bad))) return st;
  if (bad) return FSP_NOT_FINITE;
  if (grad) {
    NET_HIP(hipMemcpyAsync(grad, N->grad, N->P * sizeof(float), hipMemcpyDeviceToHost, N->stream));
    NET_HIP(hipStreamSynchronize(N->stream));
  }
  return FSP_OK;
}

fsp_status fsp_net_train_step_device(fsp_net* N, int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_x,
                                     const float* d_target, const uint8_t* d_wet, float lr, double* loss,
                                     int32_t* skipped) {
  if (!N || !d_x || !d_target || !d_wet || !(lr > 0.f)) return FSP_INVALID;
  NET_HIP(hipSetDevice(N->device));
  fsp_status st = fsp::net_reserve(N, nx, ny, nz);
  if (st) return st;
  N->cur = N->stream;
  NET_HIP(hipMemcpyAsync(N->x0, d_x, N->V[0] * 4 * sizeof(uint16_t), hipMemcpyDeviceToDevice, N->stream));
  bool bad = false;
  if ((st = loss_and_grad(N, d_target, d_wet, loss, &bad))) return st;
  if (skipped) *skipped = bad ? 1 : 0;
  if (bad) {
    N->gscale *= 0.5f;
    return FSP_OK;
  }
  N->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)N->step), c2 = 1.f - std::pow(b2, (float)N->step);
  k_adam<<<blocks_for(N->P), 256, 0, N->stream>>>(N->theta, N->m1, N->m2, N->grad, N->P, lr, b1, b2, 1e-8f, c1, c2,
                                                  N->bad);
  if ((st = pack_all(N))) return st;
  NET_HIP(hipStreamSynchronize(N->stream));
  return FSP_OK;
}

fsp_status fsp_net_mfma_probe(const uint16_t* d_a, const uint16_t* d_b, float* d_d) {
  if (!d_a || !d_b || !d_d) return FSP_INVALID;
  k_mfma_probe<<<1, 64>>>(reinterpret_cast<const _Float16*>(d_a), reinterpret_cast<const _Float16*>(d_b), d_d);
  NET_HIP(hipDeviceSynchronize());
  return FSP_OK;
}

}  // extern "C"
