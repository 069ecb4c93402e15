/*
 * cryomi FP64 Fourier transforms of real grids and images on device.
 *
 * Both transforms are exact centred DFTs applied one axis at a time as
 * batched FP64 MFMA contractions against cos / sin operators built on device
 * (angles reduced exactly in integers): the first axis takes the real input
 * to real and imaginary planes, every further axis is a complex contraction
 * with two operator segments, and the last one stores interleaved complex.
 * Each contraction leaves its output axis first, so the next axis to
 * contract is always the contiguous one. FP64 throughout (DFT error grows
 * like the transform length times the FP64 unit roundoff).
 * Conventions as in cryomi.h.
 */
#ifndef CRYOMI_FFT_H
#define CRYOMI_FFT_H

#include "cryomi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Input element types. */
#define CRYOMI_REAL_F32 0
#define CRYOMI_REAL_F64 1

/*
 * fftshift(fftn(ifftshift(pad_to_shape(x, (pd, ph, pw))))) of a real volume x (d, h, w), zero padding.
 * Replaces: FourierVoxelGridVolume.from_real_voxel_grid with pad_mode="constant" and no filter
 *   (cryojax/simulator/_volume/representations/voxel_volume.py:68-119), with pad_to_shape's split
 *   pad_left = pd/2 - d/2 per axis (cryojax/ndimage/_edges.py:110-147, 175-185) and cryoJAX's
 *   fftn = fft(ifftshift(x)) (cryojax/ndimage/_fft.py:34-54).
 * Outputs (either may be NULL, not both): complex128 and / or complex64 (pd, ph, pw), zero
 * frequency at index p/2 per axis. Sizes 1 <= d <= pd <= 1024 (likewise h, w).
 * Unpadded power-of-two volumes (16 <= d, h <= 512, 32 <= w <= 512) run in-place FFT passes
 * over the complex128 output (over the workspace when out_c128 is NULL) and need 16-byte aligned
 * buffers (CRYOMI_ERR_INVALID_ARGUMENT otherwise); the others run DFT contractions.
 */
size_t cryomi_fourier_voxel_grid_workspace_bytes(int d, int h, int w, int pd, int ph, int pw);
int cryomi_fourier_voxel_grid(const void* real_grid, int input_type, int d, int h, int w, int pd, int ph, int pw,
                              void* out_c128, void* out_c64, void* workspace, size_t workspace_bytes,
                              struct ihipStream_t* stream);

/*
 * rfftn(ifftshift(y)) over the last two axes of real images y (n, h, w) -> complex128 (n, h, w/2+1).
 * Replaces: cryojax.ndimage.rfftn on images (cryojax/ndimage/_fft.py:80-99), the observed-image
 *   spectra a Fourier-space likelihood or the pose search scores against. 1 <= h, w <= 4096.
 * Power-of-two sides (16 <= h <= 512, 32 <= w <= 512) run FFTs in place in out_c128, need only a
 * token workspace and 16-byte aligned images and out_c128 (CRYOMI_ERR_INVALID_ARGUMENT
 * otherwise); other sizes run DFT contractions in image chunks of the workspace.
 */
size_t cryomi_rfft2_f64_workspace_bytes(int n, int h, int w);
int cryomi_rfft2_f64(const void* images, int input_type, int n, int h, int w, void* out_c128, void* workspace,
                     size_t workspace_bytes, struct ihipStream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_FFT_H */
