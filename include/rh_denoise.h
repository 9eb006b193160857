/* rh_denoise.h -- C ABI of the learned moment denoiser on MI355X (gfx950).
 *
 * The denoiser has no counterpart in the reference: it is the optional
 * post-deposit correction of the noisy ion moments (n, J) of a low-ppc run
 * towards a high-ppc run, applied before the field update.  Network: four
 * 3x3x3 convolutions, 7 -> 32 -> 32 -> 32 -> 4 channels, no biases (64,800
 * weights), ReLU after the first three, residual on (n, Jx, Jy, Jz):
 *
 *   x = (n, Jx, Jy, Jz, Bx, By, Bz) / scale          (7 channels, per voxel)
 *   y = x[0:4] + conv4(relu(conv3(relu(conv2(relu(conv1(x)))))))
 *   output = y * scale[0:4]
 *
 * Convolutions are zero-padded ("same") over each volume of X*Y*Z voxels.
 * Activations and weights are BF16 on the device, products accumulate in
 * FP32 on the matrix cores; the residual and the loss are FP32.
 *
 * Layouts (all device pointers are hipMalloc'd memory of the caller; output
 * volumes 16-byte aligned):
 *   volumes   float [P][Z][Y][X][C], C = 7 for inputs, 4 for outputs/targets
 *   weights   float, the four layers concatenated, each [cout][cin][kz][ky][kx]
 *             (64,800 values: 6,048 + 27,648 + 27,648 + 3,456)
 * X must be a multiple of 16, Y and Z multiples of 4.
 *
 * Every call returns RH_DN_OK (0) or a negative code; rh_dn_last_error()
 * gives the message.  All work of one denoiser is ordered on one HIP stream.
 */
#ifndef RH_DENOISE_H
#define RH_DENOISE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { RH_DN_OK = 0, RH_DN_ERR_ARG = -1, RH_DN_ERR_HIP = -2, RH_DN_ERR_CAPACITY = -4 };
enum { RH_DN_N_WEIGHTS = 64800, RH_DN_IN_CH = 7, RH_DN_OUT_CH = 4 };

typedef struct rh_denoiser rh_denoiser;

/* max_voxels bounds P*X*Y*Z of any later call (activation buffers). */
int rh_dn_create(int64_t max_voxels, rh_denoiser** out);
void rh_dn_destroy(rh_denoiser* d);
const char* rh_dn_last_error(const rh_denoiser* d);

/* host FP32 weights in the layout above; resets the Adam state on set */
int rh_dn_set_weights(rh_denoiser* d, const float* w, int64_t n);
int rh_dn_get_weights(rh_denoiser* d, float* w, int64_t n);
/* per-channel normalisation of the 7 input channels (n, J, B) */
int rh_dn_set_scales(rh_denoiser* d, const float scale[7]);

/* y = denoise(x) on device volumes */
int rh_dn_forward(rh_denoiser* d, const float* x, float* y, int P, int X, int Y, int Z);

/* rh_dn_forward without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same kernels
 * and the same result as rh_dn_forward, but the packed input and the
 * activations live in one static device array of the library (36 MiB on
 * each GPU it runs on, held until the process ends) instead of the
 * denoiser's buffers, so P*X*Y*Z may exceed the capacity given at creation
 * up to RH_DN_STATIC_MAX_VOXELS (RH_DN_ERR_CAPACITY above it).  The call
 * makes no hipMalloc, hipFree or event: it records no kernel times.  Every
 * call on one GPU shares that GPU's array, so calls that overlap on one GPU
 * (two denoisers on different streams) race; calls ordered on one stream do
 * not.  Nothing in the array needs clearing between calls. */
enum { RH_DN_STATIC_MAX_VOXELS = 262144 };
int rh_dn_forward_static(rh_denoiser* d, const float* x, float* y, int P, int X, int Y, int Z);

/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean squared error
 * between the normalised output and target/scale over the 4 output channels;
 * loss (before the step) is written to *loss_host. */
int rh_dn_train_step(rh_denoiser* d, const float* x, const float* target, int P, int X, int Y, int Z,
                     float lr, float* loss_host);

/* gradient of that loss w.r.t. the 64,800 weights (host FP32), without a step */
int rh_dn_gradient(rh_denoiser* d, const float* x, const float* target, int P, int X, int Y, int Z,
                   float* grad_host, float* loss_host);

int rh_dn_synchronize(rh_denoiser* d);

/* Order the denoiser's later work on `hip_stream` (a hipStream_t; NULL goes
 * back to the denoiser's own stream), e.g. the stream of a simulation
 * context that applies it every step. */
int rh_dn_set_stream(rh_denoiser* d, void* hip_stream);

/* per-kernel device time of the denoiser's launches since the last reset */
int rh_dn_set_timing(rh_denoiser* d, int on);
int rh_dn_kernel_time(rh_denoiser* d, const char* name, double* ms, int64_t* launches);
int rh_dn_reset_timing(rh_denoiser* d);

#ifdef __cplusplus
}
#endif

#endif /* RH_DENOISE_H */
