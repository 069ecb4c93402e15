/*
 * cryomi noise model and image normalisation on device.
 *
 * Replaces, for batches of images in HBM:
 *   - AbstractImageModel._normalize_image / normalizes_signal=True
 *     (cryojax/simulator/_image_model/base_image_model.py:152-155,181-189):
 *     each image minus its mean, divided by its (population) standard
 *     deviation, both taken over `signal_region` (all pixels when NULL);
 *   - GaussianWhiteNoiseModel.compute_noise
 *     (cryojax/simulator/_noise_model/gaussian_noise_model.py:171-205): a
 *     real-valued Gaussian rfft half-plane of standard deviation
 *     sqrt(N^2 variance) with the zero frequency set to 0, which
 *     cryomi_irfft2 (the image model's postprocess) turns into the noise
 *     image. The random stream is Philox4x32-10 keyed by `seed`, not JAX's
 *     threefry: draws differ from the reference, their distribution does not;
 *   - GaussianWhiteNoiseModel.sample / compute_signal (:46-111): out =
 *     scale * signal + offset + noise;
 *   - GaussianWhiteNoiseModel.log_likelihood (:207-251):
 *     -sum((sim - obs)^2 / (2 var) + log(2 pi var) / 2) per image, in FP64.
 * Conventions as in cryomi.h.
 */
#ifndef CRYOMI_NOISE_H
#define CRYOMI_NOISE_H

#include "cryomi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* images (n, npix) float32, normalised in place; region (npix) uint8 or NULL */
int cryomi_normalize_images(float* images, int n, int npix, const unsigned char* region,
                            struct ihipStream_t* stream);

/*
 * Centred crop or zero pad of a batch of real images, (n, h, w) float32 -> (n, oh, ow) float32,
 * each axis on its own: out(y, x) = in(y + h/2 - oh/2, x + w/2 - ow/2) and 0 outside the input.
 * Replaces crop_to_shape / resize_with_crop_or_pad with constant zero padding
 * (cryojax/ndimage/_edges.py:15-62,110-172), as BasicImageConfig.crop_to_shape and
 * crop_or_pad_to_padded_shape (cryojax/simulator/_image_config.py:235-246) use them. With
 * CRYOMI_RESIZE_NORMALIZE each output image is then normalised as cryomi_normalize_images does
 * (over region (oh * ow) uint8 when not NULL): the crop-then-normalise of
 * AbstractImageModel.postprocess (base_image_model.py:146-155). out must not alias in.
 */
#define CRYOMI_RESIZE_NORMALIZE 0x1u
int cryomi_resize_images(const float* in, int n, int h, int w, int oh, int ow, unsigned flags,
                         const unsigned char* region, float* out, struct ihipStream_t* stream);

/* out[i] = (complex64) in[i] for count complex128 values (a spectrum handed to the float32 kernels) */
int cryomi_complex128_to_complex64(const void* in, size_t count, void* out, struct ihipStream_t* stream);

/* out (n, box, box/2+1) complex64 noise spectra */
int cryomi_white_noise_spectrum(int n, int box, double variance, unsigned long long seed, void* out,
                                struct ihipStream_t* stream);

/* out[i] = scale * signal[i] + offset + noise[i] for i < count (noise may be NULL) */
int cryomi_noisy_images(const float* signal, const float* noise, size_t count, float scale, float offset,
                        float* out, struct ihipStream_t* stream);

/* element types of cryomi_multiply_broadcast */
#define CRYOMI_F32 0
#define CRYOMI_C64 1
#define CRYOMI_F64 2
#define CRYOMI_C128 3

/*
 * data (n, npix) *= factor (npix), in place. data_type / factor_type: CRYOMI_F32, CRYOMI_C64,
 * CRYOMI_F64 or CRYOMI_C128, both of one precision (a complex factor needs complex data).
 * Replaces the multiplicative transforms around the image model, applied to a batch:
 *   - AbstractFilter.__call__ (cryojax/ndimage/transforms/_filters.py:30-33): image * get(), on
 *     the rfft image before the inverse FFT (base_image_model.py:127-128,137-150), and on the 3D
 *     transform in FourierVoxelGridVolume.from_real_voxel_grid (voxel_volume.py:105-109);
 *   - AbstractMask.__call__ (cryojax/ndimage/transforms/_masks.py:25-28): image * get(), on the
 *     real image after crop and normalisation (base_image_model.py:156-157);
 *   - ContrastTransferTheory's envelope (cryojax/simulator/_transfer_theory/
 *     transfer_theory.py:117-118): contrast_spectrum * envelope(frequency_grid).
 */
int cryomi_multiply_broadcast(void* data, long long n, long long npix, int data_type, const void* factor,
                              int factor_type, struct ihipStream_t* stream);

/* out (n) float64 log-likelihoods of observed (n, npix) under simulated (n, npix) */
int cryomi_gaussian_log_likelihood(const float* simulated, const float* observed, int n, int npix,
                                   double variance, double* out, struct ihipStream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_NOISE_H */
