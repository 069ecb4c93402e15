/*
 * cryomi image operations around the image-formation path.
 *
 * Fourier-crop downsampling of real images.
 * Replaces: cryojax.ndimage.fourier_crop_downsample_to_shape and
 *   fourier_crop_downsample for real 2D input
 *   (cryojax/ndimage/_downsample.py:91-155 argument handling, :158-199
 *   preserve_mean, :202-222 the Hartley-transform branch), with the crop of
 *   crop_to_shape (cryojax/ndimage/_edges.py:38-46) and cryoJAX's centred
 *   transforms fftn(x) = fft(ifftshift(x)), rfftn likewise
 *   (cryojax/ndimage/_fft.py:34-99).
 *
 * The map image (h, w) -> (oh, ow) is linear. With centred indices
 * k~ = k - K//2 on an axis of length K, cas t = cos t + sin t, and
 * a = 2 pi q~ m~ / h etc., the Hartley array is
 *   H = Cr x (Cc + Sc)^T + Sr x (Cc - Sc)^T
 * (cas(a + b) = cos a cas b + sin a cas(-b)); its centre oh x ow block is
 * inverted the same way and divided by oh ow. The library applies it as
 * three batched FP32 MFMA contractions with operators built on device in
 * FP64 (angles reduced exactly in integers):
 *   Zt   (2 ow, h)  = [Cc + Sc ; Cc - Sc] . x^T
 *   L    (2 oh, ow) = sum_b [P0 U_b ; P1 U_b] . Zt_b^T   (U = Cr, Sr; P = cos, sin on the coarse rows)
 *   y    (oh, ow)   = sum_a L_a . Q_a^T                   (Q = cas, cas(-) on the coarse columns)
 * and, with CRYOMI_DS_OUTPUT_FOURIER, two more that take rfftn(y).
 * Conventions as in cryomi.h.
 */
#ifndef CRYOMI_NDIMAGE_H
#define CRYOMI_NDIMAGE_H

#include "cryomi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CRYOMI_DS_PRESERVE_MEAN 0x1u  /* preserve_mean=True: scale by oh ow / (h w) */
#define CRYOMI_DS_OUTPUT_FOURIER 0x2u /* outputs_real_space=False: rfftn(y) as (oh, ow/2+1) complex64 */

/* Size in floats of the operator buffer for one (h, w) -> (oh, ow) map and flags; 0 if unsupported.
 * Sizes: 1 <= oh <= h <= 4096, 1 <= ow <= w <= 4096. */
size_t cryomi_fourier_crop_operator_floats(int h, int w, int oh, int ow, unsigned flags);

/* Build the operators (FP64 on device, stored FP32). The buffer depends only on the arguments: build
 * once and reuse it for every batch of that shape. */
int cryomi_fourier_crop_operators(int h, int w, int oh, int ow, unsigned flags, float* ops,
                                  struct ihipStream_t* stream);

size_t cryomi_fourier_crop_downsample_workspace_bytes(int n, int h, int w, int oh, int ow, unsigned flags);

/* images float32 (n, h, w) -> out float32 (n, oh, ow), or complex64 (n, oh, ow/2+1) with
 * CRYOMI_DS_OUTPUT_FOURIER. `ops` from cryomi_fourier_crop_operators with the same sizes and flags.
 * Images are processed in chunks sized by the workspace function. */
int cryomi_fourier_crop_downsample(const float* images, int n, int h, int w, int oh, int ow, unsigned flags,
                                   const float* ops, void* workspace, size_t workspace_bytes, void* out,
                                   struct ihipStream_t* stream);

/* Batched FP32 MFMA product used by the downsampling stages, exposed for tests:
 * for i < batch: C_i = sum_{s < nseg} A_{i,s} . B_{i,s}^T with A_{i,s} = a + i a_batch + s a_seg
 * (m x k, row stride lda) and B_{i,s} = b + i b_batch + s b_seg (n x k, row stride ldb); C_i = c + i c_batch
 * (m x n, row stride ldc). Strides are in floats and may be 0. */
typedef struct cryomi_bgemm_args {
  const float* a;
  long long a_batch, a_seg;
  int lda;
  const float* b;
  long long b_batch, b_seg;
  int ldb;
  float* c;
  long long c_batch;
  int ldc;
  int m, n, k, nseg, batch;
} cryomi_bgemm_args;

int cryomi_bgemm_nt(const cryomi_bgemm_args* args, struct ihipStream_t* stream);

/*
 * Fourier-crop downsampling of volumes (3D real input) and of complex images and volumes is
 * composed from the FP64 centred transform T(x) = fftshift(fftn(ifftshift(x))) of real arrays
 * (cryomi_fourier_voxel_grid, cryomi_fft.h; an image is a volume of depth 1) and the three calls
 * below. Replaces the 3D branch of fourier_crop_downsample (cryojax/ndimage/_downsample.py:135-147)
 * and the complex branch _fft_ds_complex_signal_to_shape (:190-191, 225-238):
 *   real:    y = (Re - Im) T(crop(Re T(x) - Im T(x))) / (od oh ow)       (Hartley crop, both ways)
 *   complex: c = crop(T(x)) with T(a + i b) = T(a) + i T(b);
 *            y = fftshift(ifftn(ifftshift(c))) = (conj T(Re c) + i conj T(Im c)) / (od oh ow)
 */

/* Centred crop of a complex128 (d, h, w) array to (od, oh, ow): out(o) = scale * in(o + n/2 - on/2) per
 * axis (crop_to_shape, cryojax/ndimage/_edges.py:15-62); with CRYOMI_SPEC_UNCENTRE instead
 * out(o) = scale * in((o + n/2) mod n) (ifftshift; a last axis of ow = w/2 + 1 keeps rfftn's half).
 * CRYOMI_SPEC_HARTLEY writes the real scale * (Re - Im); CRYOMI_SPEC_OUT_F32 writes float32 /
 * complex64 instead of float64 / complex128. od <= d, oh <= h, ow <= w. */
#define CRYOMI_SPEC_HARTLEY 0x1u
#define CRYOMI_SPEC_UNCENTRE 0x2u
#define CRYOMI_SPEC_OUT_F32 0x4u
int cryomi_spectrum_crop(const void* in, int d, int h, int w, int od, int oh, int ow, unsigned flags, double scale,
                         void* out, struct ihipStream_t* stream);

/* re[i], im[i] = the float64 parts of count complex values (complex128 when complex128, else complex64). */
int cryomi_complex_split(const void* in, int complex128, size_t count, double* re, double* im,
                         struct ihipStream_t* stream);

/* out = scale * (a + i b) (forward) or scale * (conj a + i conj b) (inverse) for complex128 a, b; out is
 * complex64 when out_complex64, else complex128. */
int cryomi_complex_combine(const void* a, const void* b, size_t count, int inverse, double scale, int out_complex64,
                           void* out, struct ihipStream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_NDIMAGE_H */
