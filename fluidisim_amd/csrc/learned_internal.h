// Interface between the projection (fsp.hip) and the learned preconditioner
// (learned.hip), both linked into libfsp.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fsp.h"

struct fsp_net;

namespace fsp {

// Network shape rule: nx % 64 == 0, ny % 4 == 0, nz % 4 == 0.
bool net_shape_ok(int nx, int ny, int nz);

// Reserves the network's buffers for a grid (outside any graph capture).
fsp_status net_reserve(fsp_net* N, int nx, int ny, int nz);

// Changes whenever net_reserve reallocates the shape buffers: a graph that
// captured a forward pass is stale once it differs from the value it saw.
uint64_t net_generation(const fsp_net* N);

// Network input from a residual r on `stream`; sigma from the device f64
// |r|^2 (rr) and wet-cell count (nwet) when rr is given, else `sigma`.
fsp_status net_features(fsp_net* N, hipStream_t stream, const float* r, const uint8_t* mask, const uint8_t* solid,
                        const double* rr, const uint32_t* nwet, float sigma, uint16_t* x_out);

// z = network(features(r)) / sigma on wet cells, 0 elsewhere, on `stream`;
// every kernel returns at once while *gate == 0 (gate may be null).
fsp_status net_apply(fsp_net* N, hipStream_t stream, const float* r, const uint8_t* mask, const uint8_t* solid,
                     const double* rr, const uint32_t* nwet, const int32_t* gate, float* z);

}  // namespace fsp
