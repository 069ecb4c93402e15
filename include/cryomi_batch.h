/*
 * cryomi batching: uneven (volume, box, pose-count) groups in one persistent
 * launch, and its capture into a replayable HIP graph.
 *
 * The reference batches with jax.vmap over the poses of one image model
 * (cryojax/simulator/_image_model/base_image_model.py:68-105 under vmap) and
 * compiles one XLA program per (volume, box) shape; several volumes mean
 * several programs run one after another. cryomi_simulate_groups forms the
 * images of up to CRYOMI_MAX_GROUPS such vmapped image models (each its own
 * FourierVoxelGridVolume, box and pose table) in ONE kernel launch whose
 * workgroups pull images from a shared queue, largest box first.
 *
 * Each group is equivalent to one cryomi_simulate(grid, box, n_images,
 * params, flags | layout, ...) call and produces bit-identical images.
 * Conventions as in cryomi.h.
 */
#ifndef CRYOMI_BATCH_H
#define CRYOMI_BATCH_H

#include "cryomi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CRYOMI_MAX_GROUPS 8

/* One vmapped image model. Host struct; the pointers are DEVICE pointers. */
typedef struct cryomi_volume_group {
  const void* fourier_voxel_grid;    /* complex64 grid in the layout `flags` names */
  const cryomi_image_params* params; /* n_images entries */
  void* out;                         /* float32 (n_images, box, box) */
  int box;                           /* 64, 128, 192 or 256 */
  int n_images;                      /* >= 0 */
  unsigned flags;                    /* CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_GRID_BRICKED | CRYOMI_GATHER_FILL;
                                        the layout bit must agree across groups */
  int reserved;                      /* 0 */
} cryomi_volume_group;

/* Device workspace for cryomi_simulate_groups with these groups (0 on a bad
 * description). Depends on the "groups.*" tuning values in force. */
size_t cryomi_simulate_groups_workspace_bytes(const cryomi_volume_group* groups, int n_groups);

/* Images of every group, one persistent launch (plus a one-wave queue reset).
 * Replaces a sequence of vmapped AbstractImageModel.simulate() calls
 * (base_image_model.py:68-105), one per volume. `groups` is a HOST array of
 * n_groups (1..CRYOMI_MAX_GROUPS) descriptions, read during the call only. */
int cryomi_simulate_groups(const cryomi_volume_group* groups, int n_groups, void* workspace,
                           size_t workspace_bytes, struct ihipStream_t* stream);

/* Capture of one cryomi_simulate_groups call into an instantiated HIP graph.
 * The graph keeps the device pointers of `groups` and the workspace, not their
 * contents: between replays the caller may rewrite the parameter tables in
 * place (new poses, defocus, offsets: decode-style stepping) and the next
 * replay forms the new images. Replays on one graph must not overlap. */
typedef struct cryomi_graph cryomi_graph;

int cryomi_simulate_groups_graph_create(const cryomi_volume_group* groups, int n_groups, void* workspace,
                                        size_t workspace_bytes, cryomi_graph** graph);
int cryomi_graph_launch(cryomi_graph* graph, struct ihipStream_t* stream);
void cryomi_graph_destroy(cryomi_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* CRYOMI_BATCH_H */
