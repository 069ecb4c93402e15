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
  int box;                /* even, 8 <= box <= 512, 2^a 3^b 5^c */
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
  const float* bank;     /* optional: templates of every rotation from cryomi_ncc_template_bank (same
                            grid_c64 and rotations); NULL builds them per chunk. Ignored by
                            cryomi_ncc_search_candidates, which takes its bank separately. */
  unsigned flags;        /* 0 or CRYOMI_GATHER_FILL: the out-of-range corner rule of the templates and of the
                            FP64 re-scoring (cryomi.h); a bank must have been built with the same rule */
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

/*
 * Learned-proposal search support. The grid, the scores, the raveled index
 * and the tie rule are those of cryomi_ncc_search; only the set of rotations
 * scored per image is restricted, so an image's result equals the exhaustive
 * result whenever the exhaustive winner's rotation is among its candidates.
 * No counterpart in the reference (it has no proposal mechanism); the
 * exhaustive cryomi_ncc_search stays the fallback and the judge of the error.
 */

/* floats per template row of the bank: 2 box (box/2+1) rounded up to 16,
 * plus 32 floats of padding that spread rows over the L2 channels */
size_t cryomi_ncc_bank_row_floats(int box);

/* FP32 templates (bare trilinear slices, no CTF / shift) of every rotation:
 * bank [n_rotations][cryomi_ncc_bank_row_floats(box)], interleaved (re, im).
 * flags: 0 or CRYOMI_GATHER_FILL (the gather's out-of-range corner rule). */
int cryomi_ncc_template_bank(const void* grid_c64, int box, const double* rotations, int n_rotations,
                             unsigned flags, float* bank, struct ihipStream_t* stream);

/*
 * Per image, the grid rotations within max_angle_deg (geodesic) of any of its
 * hypotheses, in ascending rotation index: cand [n_images][cap],
 * cand_count [n_images] (capped at cap), overflow [n_images] (1 if more than
 * cap matched). Hypotheses: either n_hyp unconstrained 3x3 row-major float
 * matrices per image at hyp + image * hyp_stride + 9 h (orthonormalised by
 * Gram-Schmidt on their first two columns: the network's output), or, when
 * center_index is not NULL, the single grid rotation center_index[image] /
 * n_offsets (a raveled search result).
 */
int cryomi_pose_candidates(const float* hyp, long long hyp_stride, int n_hyp, const long long* center_index,
                           int n_offsets, int n_images, const double* rotations, int n_rotations,
                           double max_angle_deg, int cap, int* cand, int* cand_count, int* overflow,
                           struct ihipStream_t* stream);

#define CRYOMI_NCC_REUSE_PREPARED 0x1u /* the same workspace, from a call with the same box, images, offsets,
                                          grid and args->flags, already holds the folded observations and
                                          bricked grid */
#define CRYOMI_NCC_MERGE 0x2u          /* keep the better of the new and the existing best_index / best_score */

typedef struct cryomi_ncc_candidate_args {
  const float* bank;      /* from cryomi_ncc_template_bank over all args->rotations */
  const int* cand;        /* [n_images][cap] rotation indices, ascending */
  const int* cand_count;  /* [n_images] */
  int cap;                /* candidates per image, 1..8192 */
  unsigned flags;         /* CRYOMI_NCC_REUSE_PREPARED | CRYOMI_NCC_MERGE */
  int* rescored;          /* optional out [n_images]: grid points re-scored in FP64 by this call */
} cryomi_ncc_candidate_args;

size_t cryomi_ncc_search_candidates_workspace_bytes(int box, int n_images, int n_offsets, int cap);
/* args->n_offsets <= 32. args->overflow reports tie-list overflow as in
 * cryomi_ncc_search (OR-ed with the existing value under CRYOMI_NCC_MERGE). */
int cryomi_ncc_search_candidates(const cryomi_ncc_search_args* args, const cryomi_ncc_candidate_args* cand,
                                 void* workspace, size_t workspace_bytes, struct ihipStream_t* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_SEARCH_H */
