/* The learned preconditioner's network for wavecore's SGN solve, on gfx950 (MI355X).
 *
 * No counterpart in the reference: wavecore preconditions its PCG with the Jacobi diagonal
 * (crates/wavecore/src/dispersion.rs:568-589). This network is an additional preconditioner
 * for a flexible PCG with the same tolerance, the reference's own preconditioner staying the
 * fallback (wavecore_hip.h: wc_solver_set_preconditioner).
 *
 * The network: a 4-level U-Net on the solver's grid, 3x3 convolutions with 16/32/64/128
 * channels, BF16 weights and activations, FP32 accumulation (MFMA). Inputs per cell: depth h,
 * the bed gradient (db/dx, db/dy) and the residual (rx, ry) scaled by its RMS; output: the
 * correction (zx, zy). Layers, in order (L0 is the grid, each level halves both sizes):
 *   E0 3x3 5->16 L0, E1 3x3 16->32 L1, E2 3x3 32->64 L2, E3 3x3 64->128 L3 (2x2 average pooling
 *   between levels), B 3x3 128->128 L3, T3 1x1 128->64 L3, D2 3x3 64->64 L2 (input: T3's output
 *   upsampled by 2, nearest, plus E2's), T2 1x1 64->32, D1 3x3 32->32 L1, T1 1x1 32->16,
 *   D0 3x3 16->16 L0, O 1x1 16->2 L0. ReLU after every 3x3 layer. Parameters: per layer the
 *   weights W[tap][in][out] then the biases [out], FP32, layers in the order above.
 *
 * Training runs on the GPU through this API: a network instance owns its buffers for one grid
 * size; the caller supplies samples (grid-layout device arrays) and drives the steps. All
 * pointers are device pointers on the current HIP device unless named *_host; `stream` is a
 * hipStream_t (0 for the default stream). Returns WC_OK (0) or an error code of
 * wavecore_hip.h, with the message in wc_last_error(). */
#ifndef WAVECORE_UNET_H
#define WAVECORE_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wcu_net wcu_net;

/* Total FP32 parameters, and the offset and size of layer l (0..11, order above). */
uint64_t wcu_param_count(void);
int wcu_layer(int32_t l, uint64_t* offset, uint64_t* weights, uint64_t* biases, int32_t* taps,
              int32_t* cin, int32_t* cout);

/* An instance for grids of up to nx x ny interior cells. */
int wcu_net_new(uint64_t nx, uint64_t ny, wcu_net** out);
void wcu_net_free(wcu_net* n);

/* Loads FP32 parameters (device pointer, wcu_param_count() values) and packs the BF16 MFMA
 * fragments of every layer, forward and transposed. */
int wcu_net_set_params(wcu_net* n, const float* params, void* stream);

/* The network's input from grid-layout arrays (row pitch `pitch` floats, interior at
 * (ghost, ghost), nx x ny cells): channels h * h_scale, bgx, bgy, rx * r_scale, ry * r_scale. */
int wcu_net_input(wcu_net* n, uint64_t nx, uint64_t ny, uint64_t pitch, uint64_t ghost,
                  const float* h, const float* bgx, const float* bgy, const float* rx,
                  const float* ry, float h_scale, float r_scale, void* stream);

/* Forward pass of the loaded input; the output (2 FP32 channels per cell, nx x ny, row pitch
 * nx) in out_xy (device, 2 * nx * ny floats, (x, y) interleaved). */
int wcu_net_forward(wcu_net* n, float* out_xy, void* stream);

/* Mean squared error of the last forward output against target (2 * nx * ny floats,
 * interleaved), over the cells where weight != 0 (nx * ny floats; the sum of the weights is
 * the divisor); writes d loss / d output into the instance and the loss to *loss_host after a
 * synchronisation of `stream`. */
int wcu_net_loss(wcu_net* n, const float* target, const float* weight, double* loss_host,
                 void* stream);

/* Backward pass from the stored d loss / d output: FP32 gradients of all parameters into grads
 * (wcu_param_count() floats, overwritten). */
int wcu_net_backward(wcu_net* n, float* grads, void* stream);

/* Adam (Kingma and Ba, 2015) on FP32 parameters with FP32 moments m, v; step t >= 1. */
int wcu_adam(uint64_t count, float* params, const float* grads, float* m, float* v, float lr,
             float beta1, float beta2, float eps, int32_t t, void* stream);

/* Copies one of the instance's tensors to host memory for checks (BF16 values widened to
 * FP32, the interior only, NHWC): ids 0..17 activations x0, e0, p1, e1, p2, e2, p3, e3, y3, t3,
 * u2, y2, t2, u1, y1, t1, u0, y0; 100 + the same numbering for their gradients where they exist
 * (d pre-activation for ReLU layers). *h, *w, *c receive the shape; host may be null to query. */
int wcu_net_tensor(wcu_net* n, int32_t id, float* host, uint64_t* h, uint64_t* w, uint64_t* c);

#ifdef __cplusplus
}
#endif

#endif
