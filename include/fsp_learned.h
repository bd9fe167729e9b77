/* Learned preconditioner of the MI355X projection library (libfsp.so).
 *
 * None of these entry points replaces a function of the reference solver:
 * they add the network preconditioner the projection applies inside
 * flexible CG (preconditioner FSP_PRECOND_LEARNED of fsp.h), with plain
 * MGPCG as the fallback. The network is a 3-level 3D convolutional
 * encoder-decoder (8/16/32 channels, 3x3x3 kernels, ReLU, additive skip
 * connections), 80 729 parameters, binary16 weights and activations with f32
 * accumulation on MFMA, trained by its own forward and backward kernels.
 *
 * Parameters travel as one f32 vector in layer order e1a e1b e2a e2b e3a
 * e3b d2a d2b d1a d1b head, each layer W[co][tap][ci] then b[co], tap =
 * (dz+1)*9 + (dy+1)*3 + (dx+1).
 *
 * Shapes: nx % 64 == 0, ny % 4 == 0, nz % 4 == 0 (two poolings, 16-cell rows
 * on the coarsest network level). Device pointers are HIP device memory on
 * the network's device; calls on one network are not thread-safe.
 */
#ifndef FSP_LEARNED_H
#define FSP_LEARNED_H

#include <stddef.h>
#include <stdint.h>

#include "fsp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fsp_net fsp_net;

/* Number of f32 parameters (80 729). */
uint32_t fsp_net_param_count(void);

/* A network on `device` with zero parameters. */
fsp_status fsp_net_create(int32_t device, fsp_net** out);
void fsp_net_destroy(fsp_net* net);

/* Host f32 parameter vector in, or out. Setting resets the optimiser. */
fsp_status fsp_net_set_params(fsp_net* net, const float* theta, size_t n);
fsp_status fsp_net_get_params(fsp_net* net, float* theta, size_t n);

/* Reserves activation buffers for a grid (done implicitly by the calls below
 * when the shape changes). FSP_SHAPE for an unsupported shape. */
fsp_status fsp_net_reserve(fsp_net* net, int32_t nx, int32_t ny, int32_t nz);

/* Network input from a residual: x[c] = (r*sigma, wet, ghost row, solid) in
 * binary16, 4 channels per cell, from the row codes of `solver`'s current
 * geometry; sigma = sqrt(wet cells / |r|^2 over wet cells) when sigma <= 0.
 * Returns the sigma used through sigma_out (host). */
fsp_status fsp_net_features_device(fsp_net* net, fsp_solver* solver, const float* d_r, uint16_t* d_x,
                                   float sigma, float* sigma_out);

/* y = network(x), x binary16 [cells][4], y f32 [cells]. */
fsp_status fsp_net_forward_device(fsp_net* net, int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_x, float* d_y);

/* No counterpart in the reference: fsp_net_forward_device for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops
 * over patches of changing shapes). It reserves nothing: the activations live
 * in a static device array of 24 MiB per GPU, which the library holds until
 * the process ends and every call on that GPU shares (calls that overlap on
 * one GPU race; each call completes before it returns). Grids of at most
 * 262 144 cells (e.g. 64 x 64 x 64); larger ones return FSP_SHAPE. y is
 * fsp_net_forward_device's y bit for bit. */
fsp_status fsp_net_forward_static_device(fsp_net* net, int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_x,
                                         float* d_y);

/* Loss = mean over wet cells of (network(x) - target)^2 and its gradient
 * w.r.t. the parameters (host f32, fsp_net_param_count values). */
fsp_status fsp_net_loss_grad_device(fsp_net* net, int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_x,
                                    const float* d_target, const uint8_t* d_wet, double* loss, float* grad);

/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on the same loss.
 * *skipped = 1 when a gradient overflowed binary16 (the step is dropped and
 * the gradient scale halved). */
fsp_status fsp_net_train_step_device(fsp_net* net, int32_t nx, int32_t ny, int32_t nz, const uint16_t* d_x,
                                     const float* d_target, const uint8_t* d_wet, float lr, double* loss,
                                     int32_t* skipped);

/* The network used by `solver` when its preconditioner is FSP_PRECOND_LEARNED
 * (the solver does not own it; null detaches). */
fsp_status fsp_attach_net(fsp_solver* solver, fsp_net* net);

/* Checks the MFMA operand layout: D = A B for 16x16 binary16 A, B (row
 * major, device), D f32. */
fsp_status fsp_net_mfma_probe(const uint16_t* d_a, const uint16_t* d_b, float* d_d);

#ifdef __cplusplus
}
#endif
#endif
