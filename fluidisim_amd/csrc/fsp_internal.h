// Internal device structures of the MI355X projection library.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fsp {

// Per-cell row code of the fine operator, rebuilt once per projection.
// Bits 0-5: the face couples to a wet neighbour, in the reference's row order
// x-, x+, y-, y+, z-, z+ (delta3d_mobile.rs:149-188). Bit 6: the cell is wet.
// Bit 7: the row carries at least one ghost (free-surface) face; such rows
// are evaluated from the geometry itself.
enum : uint8_t {
  M_XM = 1, M_XP = 2, M_YM = 4, M_YP = 8, M_ZM = 16, M_ZP = 32, M_WET = 64, M_GHOST = 128
};

constexpr int BLOCK = 256;        // threads per workgroup (4 wavefronts of 64)
constexpr int MAX_BLOCKS = 2048;  // grid cap of grid-stride kernels (8 per CU)
constexpr int NUM_SLOTS = 16;     // reduction slots
constexpr int MAX_PARTIALS = 65536;  // blocks a reducing kernel may have
constexpr int G_SLOTS = 8;        // ghost rows tabulated per column

// Reduction slots / finish operations.
enum Op : int {
  OP_B2 = 0,        // |b|^2 (and finiteness)
  OP_RR = 1,        // |r|^2 of a recomputed true residual
  OP_DQ = 2,        // <d, A d>  -> alpha, or stop
  OP_RN = 3,        // |r|^2 after the update
  OP_RZ = 4,        // <r, z> at the end of an iteration -> beta, it++, loop test
  OP_RZ_PRIME = 5,  // <r, z> of a prime (beta = 0)
  OP_MEAN_B = 6,    // sum of b and count of active rows (gauge)
  OP_MEAN_P = 7,    // sum of p and count of active rows (gauge)
  OP_DIVMAX = 8,    // max |div| over wet rows, over plain rows
  OP_UMAX = 9,      // max |u| over all faces
  OP_BACKWARD = 10, // backward error max
  OP_COUNT = 11,    // wet cells, ghost faces
  OP_GENERIC = 12,  // plain sum (tests)
};

// Device-resident control block of the conjugate gradient.
struct Ctl {
  double b2, rr, rn, rz, dq, alpha, beta, tol_b2, physical_target;
  double actual;         // last true residual |b - Ap|^2
  double red[NUM_SLOTS][2];  // raw totals of every reduction slot
  double mean_b, mean_p;
  uint32_t it, max_iters, pass_limit;
  int32_t active;        // 1 while the CG loop runs; every iteration kernel returns early at 0
  int32_t first_dir;     // 1: the next direction is d = z
  int32_t dsel;          // which direction buffer holds the current d
  int32_t nonfinite;     // a reduction met a non-finite value
  int32_t stop_dq;       // the loop stopped on <d, A d> <= 0
  int32_t pure_neumann;  // no Dirichlet face on the fine level (and gauge on)
  int32_t gauge;
  uint32_t wet_cells, ghost_faces;
  uint32_t counters[NUM_SLOTS];
};

// Fine-level geometry, passed by value to kernels.
struct Fine {
  int nx, ny, nz;
  int cut;               // aperture-weighted divergence (reference cut geometry)
  size_t n;              // cells
  float dx, inv;         // dx, 1/dx^2
  float rho, g, rest;
  int homogeneous;
  const uint8_t* solid;  // cells
  const uint8_t* ou;     // u-face apertures (0/1)
  const uint8_t* ov;
  const uint8_t* ow;
  const float* eta;      // columns
  const float* eta_rt;   // columns
  const uint8_t* mask;   // cells (row codes)
  const uint16_t* gk0;   // columns: lowest ghost row (0xFFFF: none)
  const float* gtab;     // columns x G_SLOTS x 8: 1/theta of the ghost faces
  float prec_tab[7];     // 1/(n * inv) for a plain row with n faces
};

// One coarse level (delta3d_multigrid.rs:37-52).
struct Level {
  int nx, ny, nz;
  size_t n;
  float inv;
  int8_t* kind;          // 1 active, -1 air, 0 solid
  float* ou;
  float* ov;
  float* ow;
  float* diag;
  float* x;
  float* r;
  float* t;
  float* dd;             // Galerkin levels: Dirichlet part of the diagonal (units of inv)
  int galerkin;          // 1: P^T A P / 16 coarse operator, 0: re-discretised
};

}  // namespace fsp
