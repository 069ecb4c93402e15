/*
 * cryomi: C ABI of the MI355X (gfx950) image-formation kernels.
 *
 * This is the drop-in boundary for cryoJAX's Fourier-slice image-formation
 * path. cryoJAX has no native code and no FFI (every op is an XLA primitive),
 * so each entry point below names the Python interface it replaces; a JAX
 * build binds them as XLA FFI custom calls, other hosts call them directly.
 *
 * Conventions:
 *   - every pointer argument except `params` tables documented as host is a
 *     DEVICE pointer; buffers are caller-owned, row-major, unaliased;
 *   - nothing here allocates device memory or synchronises the host; work is
 *     enqueued on `stream` (a hipStream_t; NULL = the default stream);
 *   - calls are reentrant; every call returns CRYOMI_OK (0) or an error code,
 *     and cryomi_last_error() returns the calling thread's last message;
 *   - complex numbers are interleaved float32 pairs (complex64).
 */
#ifndef CRYOMI_H
#define CRYOMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct ihipStream_t; /* hipStream_t == struct ihipStream_t* */

enum cryomi_status {
  CRYOMI_OK = 0,
  CRYOMI_ERR_INVALID_ARGUMENT = 1, /* bad shape, NULL pointer, non-finite parameter */
  CRYOMI_ERR_UNSUPPORTED = 2,      /* valid request the kernels do not implement */
  CRYOMI_ERR_WORKSPACE = 3,        /* workspace smaller than *_workspace_bytes() */
  CRYOMI_ERR_LAUNCH = 4            /* HIP runtime error at launch */
};

/* flags for cryomi_simulate */
#define CRYOMI_APPLY_CTF 0x1u         /* ContrastTransferTheory present (LinearImageModel) */
#define CRYOMI_APPLY_TRANSLATION 0x2u /* applies_translation=True, translate_mode="fft" */
#define CRYOMI_OUTPUT_FOURIER 0x4u    /* outputs_real_space=False: write the rfft half-plane */
#define CRYOMI_GRID_BRICKED 0x8u      /* grid is in the layout written by cryomi_grid_to_bricks */
/*
 * Out-of-range corner rule of the trilinear gather. The reference reads each
 * corner with input.at[indices].get(mode="fill", fill_value=0)
 * (cryojax/ndimage/_map_coordinates.py:117), and JAX normalises negative
 * integer indices (i -> i + box) before its fill-or-drop gather: by default a
 * corner index in [-box, -1] wraps and any other index outside [0, box) reads
 * 0. With CRYOMI_GATHER_FILL every corner index outside [0, box) reads 0. Only
 * slices reaching |R k| > 1/2 cycle per pixel (the plane's corners) differ.
 */
#define CRYOMI_GATHER_FILL 0x10u

/*
 * Per-image parameters (one struct per image, in DEVICE memory). Mirrors the
 * leaves of the reference's pose / CTF / transfer theory / image config:
 *   rotation        : SO3 of the pose as a row-major 3x3 matrix; the sample at
 *                     frequency k is read at R k (AbstractPose.rotate_coordinates,
 *                     cryojax/simulator/_pose.py:29-66)
 *   offset_x/_y     : AbstractPose.offset_in_angstroms[0:2] (_pose.py:97-114)
 *   defocus         : AstigmaticCTF.defocus_in_angstroms (+ offset_z when the
 *                     pose carries one, base_image_model.py:318-323)
 *   astigmatism     : AstigmaticCTF.astigmatism_in_angstroms
 *   astigmatism_angle : AstigmaticCTF.astigmatism_angle, degrees
 *   spherical_aberration_mm : AstigmaticCTF.spherical_aberration_in_mm
 *   wavelength      : Angstrom, wavelength_from_kilovolts(voltage)
 *   pixel_size      : BasicImageConfig.pixel_size, Angstrom
 *   phase_shift     : radians; deg2rad(ContrastTransferTheory.phase_shift) +
 *                     atan(ac / sqrt(1 - ac^2)) (transfer_function.py:66-77)
 *   amplitude       : multiplies the extracted slice (pixel_size when
 *                     FourierSliceExtraction.outputs_integral, else 1)
 */
typedef struct cryomi_image_params {
  double rotation[9];
  double offset_x;
  double offset_y;
  double defocus;
  double astigmatism;
  double astigmatism_angle;
  double spherical_aberration_mm;
  double wavelength;
  double pixel_size;
  double phase_shift;
  double amplitude;
  double reserved;
} cryomi_image_params; /* 20 doubles, 160 bytes */

const char* cryomi_last_error(void);
int cryomi_abi_version(void);

/*
 * Launch tuning (process-wide). No counterpart in the reference (XLA picks
 * its own launch shapes); set by the build's autotuner. Keys:
 *   "persistent_wg_per_cu"       : resident workgroups per CU of the
 *                                  persistent image kernels, every box;
 *                                  0 (default) = occupancy limit
 *   "persistent_wg_per_cu.<box>" : the same for one box (overrides the
 *                                  above; -1 = unset)
 *   "groups.threads"             : workgroup size of the grouped kernel
 *                                  (cryomi_batch.h), 256 (default) or 512;
 *                                  sets its column / row tile shapes
 *   "groups.wg_per_cu"           : resident workgroups per CU of the
 *                                  grouped kernel; 0 (default) = occupancy
 * Changing a key changes *_workspace_bytes(); query it again afterwards.
 */
int cryomi_set_tuning(const char* key, int value);
int cryomi_get_tuning(const char* key); /* -1 for an unknown key */

/* Workspace (device bytes) cryomi_simulate needs for this request. */
size_t cryomi_simulate_workspace_bytes(int box, int n_images, unsigned flags);

/*
 * Batched fused forward: Fourier-slice extraction (trilinear, fill mode) of a
 * centred complex64 (box, box, box) grid on each image's rotated central plane,
 * self-conjugate fix-ups, AstigmaticCTF, in-plane shift phase, then
 * fftshift(irfft2) to a (box, box) float32 image per image.
 *
 * Replaces: LinearImageModel.simulate() under vmap over pose / CTF leaves
 *   (cryojax/simulator/_image_model/base_image_model.py:68-105, 296-328) with
 *   FourierSliceExtraction.integrate
 *   (cryojax/simulator/_volume_integrator/fourier_voxel_extract.py:77-146),
 *   ContrastTransferTheory.propagate_object
 *   (cryojax/simulator/_transfer_theory/transfer_theory.py:53-120) and
 *   irfftn (cryojax/ndimage/_fft.py:57-77). Without CRYOMI_APPLY_CTF this is
 *   ProjectionImageModel.simulate(); with CRYOMI_OUTPUT_FOURIER it is
 *   compute_fourier_image() (rfft half-plane, zero frequency in the corner).
 *
 * fourier_voxel_grid : complex64 [box][box][box], zero frequency at box/2, or
 *                      the cryomi_grid_to_bricks layout with CRYOMI_GRID_BRICKED
 * params             : cryomi_image_params[n_images] (device)
 * out                : float32 [n_images][box][box], or complex64
 *                      [n_images][box][box/2+1] with CRYOMI_OUTPUT_FOURIER
 * box                : even, 8 <= box <= 512, box = 2^a 3^b 5^c
 */
int cryomi_simulate(const void* fourier_voxel_grid, int box, int n_images,
                    const cryomi_image_params* params, unsigned flags, void* workspace,
                    size_t workspace_bytes, void* out, struct ihipStream_t* stream);

/*
 * Allocation-free variant of cryomi_simulate. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, flags, checks and images as cryomi_simulate, without the
 * workspace: the persistent image kernels keep their per-workgroup scratch
 * slots (box * (box/2 + 1) complex64 each) in a static device array of
 * CRYOMI_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, and the
 * persistent grid is capped to the slots it holds (64 MiB: 254 resident
 * workgroups at box 256, 1008 at box 128; any number of images); a box whose
 * slot would not fit is rejected with CRYOMI_ERR_WORKSPACE. The images are
 * formed in their natural order (no schedule table). The call allocates and
 * frees nothing. Calls on one GPU share its array: calls that overlap in time
 * on one GPU (other streams, other threads) race and must be serialised by
 * the caller.
 */
#define CRYOMI_NOALLOC_SCRATCH_BYTES (64u << 20)
int cryomi_simulate_noalloc(const void* fourier_voxel_grid, int box, int n_images,
                            const cryomi_image_params* params, unsigned flags, void* out,
                            struct ihipStream_t* stream);

/*
 * Vector-Jacobian product of cryomi_simulate (real-space output): given the
 * cotangent dL/dimage, accumulates
 *   grid_grad   += dL/dRe G + i dL/dIm G   (complex64 [box]^3, summed in
 *                  FP32 per brick of voxels on chip, then added once; the
 *                  summation order varies between runs; the caller zeroes it;
 *                  JAX's convention is the conjugate)
 *   param_grads += per image [dL/dR00, dL/dR01, dL/dR10, dL/dR11, dL/dR20,
 *                  dL/dR21, dL/doffset_x, dL/doffset_y, dL/ddefocus]
 *                  (double [n_images][9]; R is cryomi_image_params.rotation,
 *                  the caller chains dL/dR to its pose parametrisation)
 * Replaces: reverse-mode AD of LinearImageModel.simulate() (jax.grad /
 * eqx.filter_grad, README.md:99-123), whose transpose of the trilinear gather
 * (cryojax/ndimage/_map_coordinates.py:117) is an XLA scatter-add.
 * flags: CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_GATHER_FILL as in
 * the forward; with CRYOMI_GRID_BRICKED, fourier_voxel_grid is the
 * cryomi_grid_to_bricks layout (box % 4 == 0) and the gather reads one aligned
 * x pair per (z, y) row, as the forward does; grid_grad keeps the standard
 * layout either way.
 * cotangent: float32 [n_images][box][box].
 * workspace: cryomi_simulate_vjp_workspace_bytes(box, n_images) bytes: 16 B
 * per rfft sample plus, up to box 256, for one pass of at most 2048 images the
 * per-brick sample lists (at most 32 B per sample) and the list places of the
 * pass's sample blocks (4 B per block and brick, at most 256 MiB).
 */
size_t cryomi_simulate_vjp_workspace_bytes(int box, int n_images);
int cryomi_simulate_vjp(const void* fourier_voxel_grid, int box, int n_images,
                        const cryomi_image_params* params, unsigned flags, const float* cotangent,
                        void* workspace, size_t workspace_bytes, void* grid_grad, double* param_grads,
                        struct ihipStream_t* stream);

/*
 * One-time relayout of a centred complex64 (box, box, box) grid for the
 * gather: 2x2x4 (z, y, x) bricks of one 128-byte cache line each, consecutive
 * bricks of a row overlapping by one voxel in x (brick b holds x = 3b - 1 ..
 * 3b + 2, x = -1 the wrapped x = box - 1, x >= box zero), so that every
 * trilinear x pair is one 16-byte load from one line. Pass the result to
 * cryomi_simulate with CRYOMI_GRID_BRICKED. No counterpart in the reference
 * (its grid is an XLA array); it is the storage
 * FourierVoxelGridVolume.from_real_voxel_grid prepares for the kernels
 * (cryojax/simulator/_volume/representations/voxel_volume.py:68-119).
 * box % 4 == 0; out holds cryomi_bricked_grid_bytes(box) bytes (about
 * 4/3 box^3 complex64) and must not alias the input.
 */
size_t cryomi_bricked_grid_bytes(int box);
int cryomi_grid_to_bricks(const void* fourier_voxel_grid, int box, void* out, struct ihipStream_t* stream);

/*
 * fftshift(irfft2(x, s=(box, box))) for a batch of rfft half-planes.
 * Replaces: cryojax.ndimage.irfftn (cryojax/ndimage/_fft.py:57-77) for 2D
 * input with zero frequency in the corner. The imaginary parts of the
 * self-conjugate x bins are ignored (C2R semantics).
 * in  : complex64 [n][box][box/2+1];  out : float32 [n][box][box]
 */
size_t cryomi_irfft2_workspace_bytes(int box, int n);
int cryomi_irfft2(const void* in, int box, int n, void* workspace, size_t workspace_bytes,
                  void* out, struct ihipStream_t* stream);

/*
 * The image model's steps after integration, applied to a given complex64 rfft half-plane per
 * image (box, box/2+1): the self-conjugate fix-ups, amplitude, AstigmaticCTF and shift phase of
 * cryomi_simulate (flags CRYOMI_APPLY_CTF / CRYOMI_APPLY_TRANSLATION, params as there, rotation
 * unused), then fftshift(irfft2) to float32 (box, box) images, or, with CRYOMI_OUTPUT_FOURIER, the
 * complex64 half-plane itself. Replaces ContrastTransferTheory.propagate_object
 * (cryojax/simulator/_transfer_theory/transfer_theory.py:53-120), _phase_shift_translate
 * (base_image_model.py:160-171) and irfftn at padded_shape for a slice the integrator resized to
 * padded_shape (fourier_voxel_extract.py:132-146: irfftn, crop_or_pad, rfftn). Workspace:
 * cryomi_irfft2_workspace_bytes(box, n_images) (none with CRYOMI_OUTPUT_FOURIER).
 */
int cryomi_simulate_from_spectrum(const void* spectrum, int box, int n_images,
                                  const cryomi_image_params* params, unsigned flags, void* workspace,
                                  size_t workspace_bytes, void* out, struct ihipStream_t* stream);

/*
 * The low-frequency band of cryomi_simulate's CRYOMI_OUTPUT_FOURIER half-plane: out (n_images,
 * band + 1, band/2 + 1) complex64, row r = the box image's rfft row ky = r - band/2 (ky from
 * -band/2 to band/2), columns kx = 0 .. band/2; every sample equals the one cryomi_simulate
 * writes at (ky mod box, kx). These are the frequencies a Fourier crop of the box image to
 * band x band reads (fourier_crop_downsample, cryojax/ndimage/_downsample.py:91-222), so a crop
 * of a simulated image costs (band / box)^2 of the slice. Flags as cryomi_simulate (the output is
 * always Fourier); 2 <= band < box, band even. No workspace.
 */
int cryomi_simulate_band(const void* fourier_voxel_grid, int box, int n_images, const cryomi_image_params* params,
                         unsigned flags, int band, void* out, struct ihipStream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_H */
