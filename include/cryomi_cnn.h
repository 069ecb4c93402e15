/*
 * cryomi CNN primitives: BF16 MFMA (FP32 accumulate) building blocks of the
 * pose-proposal network, its input pipeline and its optimiser.
 *
 * No counterpart in the reference: cryoJAX has no learned pose proposals.
 * The network proposes orientations that cryomi_ncc_search_candidates
 * (cryomi_search.h) then scores exactly on the reference grid-search's own
 * grid points (cryojax/jax_util/_grid_search/search_loop.py:24-158), with the
 * exhaustive search as the fallback. Conventions as in cryomi.h; bf16 tensors
 * are uint16 bit patterns (round-to-nearest-even from float32); activations
 * are NHWC.
 */
#ifndef CRYOMI_CNN_H
#define CRYOMI_CNN_H

#include "cryomi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CRYOMI_GEMM_RELU 0x1u       /* output = max(output, 0) */
#define CRYOMI_GEMM_ACCUMULATE 0x2u /* c32 += alpha * A B with atomics (no bias / relu / mask / c16) */

/*
 * C[m, n] = epilogue(alpha * sum_k A(m, k) B(k, n)) with bf16 A, B and FP32
 * accumulation on v_mfma_f32_16x16x32_bf16 (128 x 128, 128 x 64, 64 x 128 or
 * 64 x 64 tiles, the largest that still fills the GPU).
 *   A(m, k) = a[m * lda_m + k * lda_k], B(k, n) = b[k * ldb_k + n * ldb_n]
 *   (element strides; for each operand exactly one of its two strides is 1).
 * Epilogue (plain mode): + bias[n] (if bias), ReLU (CRYOMI_GEMM_RELU), zero
 * where the bf16 mask(m, n) = mask[m * ld_mask + n] is not positive (if
 * mask), then store to c32 (float) and/or c16 (bf16).
 * CRYOMI_GEMM_ACCUMULATE: atomically add alpha * A B into c32; split_k > 1
 * (only allowed here) splits the k range over that many workgroup layers.
 * Requirements: k % 8 == 0; 16-byte aligned a and b; the non-unit strides
 * multiples of 8; m % 8 == 0 when A is contiguous along m, n % 8 == 0 when B
 * is contiguous along n.
 */
typedef struct cryomi_gemm_bf16_args {
  const void* a;
  long long lda_m, lda_k;
  const void* b;
  long long ldb_k, ldb_n;
  int m, n, k;
  float alpha;
  const float* bias;
  const void* mask;
  long long ld_mask;
  float* c32;
  long long ldc32;
  void* c16;
  long long ldc16;
  unsigned flags;
  int split_k;
} cryomi_gemm_bf16_args;

int cryomi_gemm_bf16(const cryomi_gemm_bf16_args* args, struct ihipStream_t* stream);

/*
 * 3x3 / stride-2 / padding-1 convolution layers as implicit GEMMs on the same
 * MFMA kernel, the patches gathered from the NHWC activations while the
 * operands are staged (no im2col or col2im buffers). x is the layer input
 * (n, h, w, c) bf16, h, w (>= 2), c and co (>= 8) powers of two; weights (co, 9 c)
 * bf16 with column (ty * 3 + tx) * c + ci; dy = dL/dy (n, h/2, w/2, co) bf16.
 *   CRYOMI_CONV_FORWARD     : y = patches(x) . weights^T + bias, ReLU with
 *                             CRYOMI_GEMM_RELU; c16 (bf16) and / or c32
 *                             (float), (n h/2 w/2, co);
 *   CRYOMI_CONV_WEIGHT_GRAD : c32 (co, 9 c) = dy^T . patches(x) and, if
 *                             bias_grad, bias_grad (co) = column sums of dy
 *                             (one more, all-ones patch column); added to the
 *                             outputs instead with CRYOMI_GEMM_ACCUMULATE. The
 *                             pixels split into split_k ranges whose partial
 *                             products go to workspace slices, summed in a
 *                             fixed order (deterministic, no atomics);
 *   CRYOMI_CONV_DATA_GRAD   : c16 (n h w, c) = [mask > 0] (the adjoint of the
 *                             patch gather applied to dy . weights); mask is the
 *                             layer input (its ReLU mask) or NULL.
 */
#define CRYOMI_CONV_FORWARD 0
#define CRYOMI_CONV_WEIGHT_GRAD 1
#define CRYOMI_CONV_DATA_GRAD 2
typedef struct cryomi_conv_bf16_args {
  int op;
  int n, h, w, c, co;
  const void* x;
  const void* weights;
  const void* dy;
  const float* bias;
  const void* mask;
  float* c32;
  void* c16;
  unsigned flags;
  int split_k;
  float* bias_grad;       /* weight gradient: (co) = dL/dbias, or NULL */
  void* workspace;        /* weight gradient: split_k * co * (9 c + (bias_grad != NULL)) floats */
  size_t workspace_bytes;
} cryomi_conv_bf16_args;

int cryomi_conv3x3_s2_bf16(const cryomi_conv_bf16_args* args, struct ihipStream_t* stream);

/* 3x3, stride 2, padding 1 patches of an NHWC bf16 tensor (n, h, w, c), h and
 * w even: out (n * h/2 * w/2, kp) bf16, column (ty * 3 + tx) * c + ci, zero
 * for taps outside the image and for columns >= 9 c (kp >= 9 c, kp % 8 == 0). */
int cryomi_im2col_s2(const void* x, int n, int h, int w, int c, int kp, void* out, struct ihipStream_t* stream);

/* Adjoint of cryomi_im2col_s2 followed by the ReLU mask of the layer input:
 * dx (n, h, w, c) bf16 = [x > 0] * sum over the patches that read each
 * element of dcol (n * h/2 * w/2, kp) float. x may be NULL (no mask). */
int cryomi_col2im_s2(const float* dcol, const void* x, int n, int h, int w, int c, int kp, void* dx,
                     struct ihipStream_t* stream);

/* grad[j] += sum_i dy[i * ld + j] for a bf16 (m, n) matrix (bias gradients). */
int cryomi_column_sums_bf16(const void* dy, int m, int n, long long ld, float* grad, struct ihipStream_t* stream);

/* float -> bf16 (n elements), and zero-fill of n floats (graph-safe, no memset). */
int cryomi_to_bf16(const float* x, size_t n, void* out, struct ihipStream_t* stream);
int cryomi_fill_zero(float* x, size_t n, struct ihipStream_t* stream);

/*
 * Pose head loss over a batch. out (n, 96) float per image:
 * [0, 72) eight 3x3 rotation hypotheses (row-major, unconstrained),
 * [72, 88) eight (x, y) offsets, [88, 96) eight confidence logits.
 * Targets: rot (n, 9) row-major rotation matrices, off (n, 2). With
 * d_h = ||R_h - R||_F^2 and the winner w = argmin d_h:
 *   L = sum_h c_h (d_h + off_weight ||t_h - t||^2) + CE(logits, w),
 * c_w = 1 - eps, c_h = eps / 7 otherwise; the batch loss is the mean over
 * images. Writes dout = dL/dout (n, 96) as bf16 (and float if dout32) and
 * per-image terms (n, 3): total, min_h d_h, winner index.
 */
int cryomi_pose_loss(const float* out, const float* rot, const float* off, int n, float eps, float off_weight,
                     void* dout, float* dout32, float* terms, struct ihipStream_t* stream);

/* Adam on flat float32 parameters; also writes their bf16 copy p16 (may be
 * NULL). bias_c1 = 1 - beta1^t, bias_c2 = 1 - beta2^t. */
int cryomi_adam(float* p, float* m, float* v, const float* g, void* p16, size_t n, float lr, float beta1,
                float beta2, float eps, float bias_c1, float bias_c2, struct ihipStream_t* stream);

/*
 * Network input, step 1: rfft half-planes (n, box, box/2+1) (complex128 if
 * `complex128`, else complex64; zero frequency in the corner, as
 * rfftn(ifftshift(image))) cropped to (n, out_box, out_box/2+1) complex64.
 * Flags: CRYOMI_CROP_PHASE_FLIP multiplies by the sign of each image's CTF
 * (params as in cryomi_simulate, evaluated on the box grid);
 * CRYOMI_CROP_NORMALIZE makes the box image zero-mean and unit-variance
 * first (Parseval over the half-plane) and scales it by sqrt(snr[i]).
 * CRYOMI_CROP_HARTLEY: the crop of fourier_crop_downsample's real branch
 * (cryojax/ndimage/_downsample.py:202-222, a centred Hartley crop to
 * out_box x out_box): out = rfftn of its real-space output. Interior bins are
 * the input's; a bin q on the unpaired -out_box/2 edge (row -out_box/2 or
 * column out_box/2) is ((1 - i) H(q) + (1 + i) H(q')) / 2, H = Re - Im of the
 * input spectrum, q' = -q folded into the crop window.
 * CRYOMI_CROP_BAND_INPUT: spec is cryomi_simulate_band's layout (n,
 * out_box + 1, out_box/2 + 1) of a box image (out_box < box) instead of the
 * full half-plane; CRYOMI_CROP_NORMALIZE then normalises the band-limited
 * image (its variance from the band's Parseval sum).
 * cryomi_irfft2(out_box) then gives the fourier-cropped image.
 */
#define CRYOMI_CROP_PHASE_FLIP 0x1u
#define CRYOMI_CROP_NORMALIZE 0x2u
#define CRYOMI_CROP_HARTLEY 0x4u
#define CRYOMI_CROP_BAND_INPUT 0x8u
int cryomi_crop_spectrum(const void* spec, int complex128, int n, int box, int out_box,
                         const cryomi_image_params* params, const float* snr, unsigned flags, void* out,
                         struct ihipStream_t* stream);

/* Network input, step 2: x (n, s, s) float + noise_std * N(0, 1) (Philox4x32-10
 * counter stream keyed by seed), then each image made zero-mean and
 * unit-variance, stored as bf16 (n, s, s). */
int cryomi_noise_normalize(const float* x, int n, int s, float noise_std, unsigned long long seed, void* out,
                           struct ihipStream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_CNN_H */
