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

/* out (n, box, box/2+1) complex64 noise spectra */
int cryomi_white_noise_spectrum(int n, int box, double variance, unsigned long long seed, void* out,
                                struct ihipStream_t* stream);

/* out[i] = scale * signal[i] + offset + noise[i] for i < count (noise may be NULL) */
int cryomi_noisy_images(const float* signal, const float* noise, size_t count, float scale, float offset,
                        float* out, struct ihipStream_t* stream);

/* out (n) float64 log-likelihoods of observed (n, npix) under simulated (n, npix) */
int cryomi_gaussian_log_likelihood(const float* simulated, const float* observed, int n, int npix,
                                   double variance, double* out, struct ihipStream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_NOISE_H */
