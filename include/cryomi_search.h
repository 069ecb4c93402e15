/*
 * cryomi pose search: exhaustive normalised-cross-correlation grid search of
 * observed images over (rotation, in-plane offset) grid points, on gfx950.
 *
 * Replaces the reference's pose-search pattern run_grid_search(fn,
 * MinimumSearchMethod(batch_size), tree_grid=(rotations, offsets), args)
 * (cryojax/jax_util/_grid_search/search_loop.py:24-158,
 * search_method.py:149-310) with fn = -NCC of the observed image against
 * LinearImageModel.compute_fourier_image at the grid point
 * (cryojax/simulator/_image_model/base_image_model.py:296-328). The minimum is
 * per image; the raveled index is rot * n_offsets + off (C order over the two
 * leaves) and, as in MinimumSearchMethod, the lowest raveled index wins exact
 * ties.
 *
 * NCC over the rfft half-plane without DC and the three self-conjugate bins,
 * weights w = 1 on columns 0 and box/2, 2 elsewhere:
 *   NCC = sum w Re(Y conj T) / sqrt(sum w |Y|^2 * sum w |T|^2),
 *   T = CTF_image * slice(R) * exp(-2 pi i q . t).
 * Scores are computed in FP32 on MFMA (v_mfma_f32_16x16x4f32) as one GEMM of
 * CTF- and shift-folded observed spectra against the templates; every grid
 * point whose FP32 score lies within `tie_window` of an image's best is
 * re-scored in FP64 from the complex128 grid, and the FP64 minimum decides.
 *
 * Same conventions as cryomi.h (device pointers, no allocation, no host sync,
 * status codes, cryomi_last_error()).
 */
#ifndef CRYOMI_SEARCH_H
#define CRYOMI_SEARCH_H

#include "cryomi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cryomi_ncc_search_args {
  const void* grid_c64;   /* complex64 [box][box][box], zero frequency at box/2 (templates) */
  const void* grid_c128;  /* complex128, the same grid in float64 (FP64 re-scoring) */
  int box;                /* even, 8 <= box <= 512, 2^a 3^b */
  int n_images;
  const void* observed_rfft; /* complex128 [n_images][box][box/2+1], rfftn(ifftshift(y)) */
  /* per image: defocus, astigmatism, astigmatism_angle, spherical_aberration_mm,
     wavelength, pixel_size, phase_shift; rotation / offset / amplitude unused */
  const cryomi_image_params* image_params;
  int n_rotations;
  const double* rotations; /* [n_rotations][9] row-major rotation matrices */
  int n_offsets;
  const double* offsets; /* [n_offsets][2] in-plane offsets (x, y), Angstrom */
  double tie_window;     /* FP32 scores within this of the best are re-scored; <= 0 selects 2e-4 */
  long long* best_index; /* out [n_images]: raveled rot * n_offsets + off */
  double* best_score;    /* out [n_images]: FP64 -NCC at best_index */
  int* overflow;         /* out [n_images]: 1 if the candidate list overflowed (result still FP32-exact) */
} cryomi_ncc_search_args;

size_t cryomi_ncc_search_workspace_bytes(int box, int n_images, int n_rotations, int n_offsets);
int cryomi_ncc_search(const cryomi_ncc_search_args* args, void* workspace, size_t workspace_bytes,
                      struct ihipStream_t* stream);

/*
 * C[M][N] = A[M][K] . B[N][K]^T in FP32 on MFMA (the contraction used by the
 * search, exposed for testing and reuse). Row-major, K % 16 == 0, any M, N.
 * No counterpart in the reference (XLA emits its own GEMMs).
 */
int cryomi_sgemm_nt(const float* A, const float* B, float* C, int M, int N, int K, struct ihipStream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_SEARCH_H */
