/*
 * unet_hip.h -- learned warm start of the MBIR solver (libmbir4d.so).
 *
 * No counterpart in the reference: the reference starts every
 * multi-resolution stage from zero or from the up-sampled coarser stage
 * (XT_ICD_update.c:1127-1180).  This network maps images computed directly
 * from a scan (the filtered back-projection of each frame and its mean over
 * the frames, mbir_hip_fbp; or the normalised back-projection,
 * mbir_hip_bp_normalized) to an estimate of the converged MBIR image; the
 * solver then starts from it and still runs to convergence, so the estimate
 * changes the work, not the result.
 *
 * Network: 3-level 3D U-Net, channels c / 2c / 4c, two 3x3x3 convolutions
 * (zero padding, ReLU) per level, 2x2x2 average pooling, nearest 2x
 * up-sampling, skips concatenated ([up-sampled | skip] channels), a 3x3x3
 * output convolution and a residual connection: y = x_0 + net(x), x_0 the
 * first input channel.  BF16 weights and activations, FP32 accumulation
 * (MFMA), FP32 master weights, Adam.  Inputs are [N][D][H][W][in_channels]
 * FP32 in host memory (channels last), outputs and targets [N][D][H][W];
 * D and H divisible by 4, W by 64.
 *
 * Parameter order (unet_get_params / unet_set_params / unet_gradients): per
 * layer (enc1a, enc1b, enc2a, enc2b, bott_a, bott_b, dec2a, dec2b, dec1a,
 * dec1b, out) the weights [cout][27][cin] with tap = (dz+1)*9 + (dy+1)*3 +
 * (dx+1) (cross-correlation: out(v) = sum_tap w[tap] in(v + d_tap)), then
 * the bias [cout].
 *
 * Functions return 0 or -1 (unet_last_error() says why).
 */
#ifndef UNET_HIP_H_
#define UNET_HIP_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct unet_ctx unet_ctx;

const char* unet_last_error(void);

/* channels c in {8, 16, 24, 32}, in_channels in [1, 8]; He-normal initial
 * weights from seed. */
int unet_create(unet_ctx** out, int32_t channels, int32_t in_channels, uint32_t seed);
void unet_destroy(unet_ctx* ctx);
int64_t unet_param_count(const unet_ctx* ctx);
int unet_get_params(unet_ctx* ctx, float* params);
int unet_set_params(unet_ctx* ctx, const float* params);

/* y = x_0 + net(x) for a batch of N volumes D x H x W. */
int unet_forward(unet_ctx* ctx, int32_t N, int32_t D, int32_t H, int32_t W, const float* x,
                 float* y);

/* Loss mean((x_0 + net(x) - t)^2) and its gradient with respect to every
 * parameter (in the parameter order above), without an update. */
int unet_gradients(unet_ctx* ctx, int32_t N, int32_t D, int32_t H, int32_t W, const float* x,
                   const float* t, float* grads, double* loss);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8, learning rate lr) on the loss
 * above; *loss is the loss before the step. */
int unet_train_step(unet_ctx* ctx, int32_t N, int32_t D, int32_t H, int32_t W, const float* x,
                    const float* t, float lr, double* loss);

/* One convolution layer through the network's kernels (BF16 operands, FP32
 * results), for parity tests.  Activations [N][D][H][W][channels] FP32.
 *   mode 0: out [v][cout] = conv(in) + bias, ReLU when relu;
 *   mode 1: out [v][cin]  = input gradient of the convolution for gout;
 *   mode 2: out [cout][27][cin] = weight gradient for gout and in. */
int unet_conv3d(int32_t mode, int32_t N, int32_t D, int32_t H, int32_t W, int32_t cin,
                int32_t cout, const float* in, const float* w, const float* bias, int32_t relu,
                const float* gout, float* out);

/* One MFMA (v_mfma_f32_16x16x32_bf16) through the lane layout the kernels
 * use: D[16][16] = A[16][32] B[32][16], A and B rounded to BF16. */
int unet_mfma_probe(const float* A, const float* B, float* D);

#ifdef __cplusplus
}
#endif
#endif
