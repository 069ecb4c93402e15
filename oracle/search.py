"""FP64 restatement of cryoJAX's grid search and of the pose-search NCC objective.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Grid search (reference: cryojax/jax_util/_grid_search/):
  * tree grid = tuple of leaves, each an array whose leading axis indexes grid
    points; the grid is their cartesian product, raveled in C order
    (pytree_manipulation.py:12-158, jnp.unravel_index at :114, which clips
    out-of-range raveled indices to the last point);
  * run_grid_search: fori_loop over grid_size // batch_size batches of
    linspace-built raveled indices, then a remainder loop that reruns the
    batched body grid_size % batch_size times (search_loop.py:108-153);
  * MinimumSearchMethod: state starts at +inf; update keeps value < best
    (strict); batch_update takes argmin over the batch (first occurrence)
    then merges with strict < (search_method.py:187-265). Net rule: the lowest
    raveled index wins exact ties.
Pinned by the reference's own analytic landscape test
(tests/test_jax_util.py:124-167, golden json) and by brute force on tiny grids.

NCC objective (no reference counterpart; the library has no NCC, the survey of
the path defines it for the pose search): for observed image y with rfft
half-plane Y = rfftn(ifftshift(y)) and template T = CTF * slice(R) * shift(t)
(LinearImageModel.compute_fourier_image), over the half-plane bins K without
DC and the three self-conjugate bins, with rfft weights w = 1 on columns 0 and
N/2 and 2 elsewhere:
    NCC = sum_K w Re(Y conj T) / sqrt(sum_K w |Y|^2 * sum_K w |T|^2)
and the search minimises -NCC. Pinned by: NCC(y, y) = 1 for a noise-free
template on the grid, invariance to affine changes of y's intensity, and its
identity with the real-space normalised correlation of the two images after
their DC and self-conjugate frequencies are removed (Parseval; checks the
weights, the conjugation and the real part).
"""

from __future__ import annotations

import math

import numpy as np

from . import imaging as O

# ----------------------------------------------------------------------------
# grid search semantics
# ----------------------------------------------------------------------------


def tree_grid_shape(tree_grid):
    """pytree_manipulation.py:12-75 for a tuple of array leaves."""
    if len(tree_grid) == 0:
        raise ValueError("The pytree passed to `tree_grid_shape` should have at least one leaf.")
    return tuple(int(np.shape(leaf)[0]) for leaf in tree_grid)


def tree_grid_unravel_index(raveled, tree_grid):
    """C-order unravel with jnp.unravel_index's clipping (pytree_manipulation.py:78-118)."""
    shape = tree_grid_shape(tree_grid)
    total = math.prod(shape)
    r = np.clip(np.asarray(raveled, dtype=np.int64), -total, total - 1)
    r = np.where(r < 0, r + total, r)
    return np.unravel_index(r, shape)


def tree_grid_take(tree_grid, index):
    """pytree_manipulation.py:121-158."""
    return tuple(np.asarray(leaf)[i] for leaf, i in zip(tree_grid, index))


class MinimumState:
    def __init__(self, current_minimum_eval, current_best_raveled_index):
        self.current_minimum_eval = current_minimum_eval
        self.current_best_raveled_index = current_best_raveled_index


def run_grid_search_minimum(fn, tree_grid, args, batch_size=None):
    """run_grid_search with MinimumSearchMethod(batch_size) (values, best raveled index).

    `fn(point, args)` gets one grid point (unbatched path) or, with a batch
    size, a tuple of leaves with a leading batch axis and must return the
    stacked values (the vmap of the reference's fn).
    """
    shape = tree_grid_shape(tree_grid)
    grid_size = math.prod(shape)
    p0 = tree_grid_take(tree_grid, tree_grid_unravel_index(0, tree_grid))
    out0 = np.asarray(fn(tuple(x[None] for x in p0), args) if batch_size else fn(p0, args), dtype=np.float64)
    f_shape = out0.shape[1:] if batch_size else out0.shape
    best = np.full(f_shape, np.inf)
    best_idx = np.zeros(f_shape, dtype=np.int64)
    if batch_size is None:
        for it in range(grid_size):
            pt = tree_grid_take(tree_grid, tree_grid_unravel_index(it, tree_grid))
            v = np.asarray(fn(pt, args), dtype=np.float64)
            less = v < best
            best = np.where(less, v, best)
            best_idx = np.where(less, it, best_idx)
        return MinimumState(best, best_idx)
    n_iter = grid_size // batch_size

    def body(it, best, best_idx):
        idx = np.linspace(it * batch_size, (it + 1) * batch_size - 1, batch_size).astype(np.int32)
        pts = tree_grid_take(tree_grid, tree_grid_unravel_index(idx, tree_grid))
        vals = np.asarray(fn(pts, args), dtype=np.float64)
        am = np.argmin(vals, axis=0)
        v = np.min(vals, axis=0)
        ridx = np.take(idx, am).astype(np.int64)
        less = v < best
        return np.where(less, v, best), np.where(less, ridx, best_idx)

    for it in range(n_iter):
        best, best_idx = body(it, best, best_idx)
    if grid_size % batch_size != 0:
        for it in range(n_iter, n_iter + grid_size % batch_size):
            best, best_idx = body(it, best, best_idx)
    return MinimumState(best, best_idx)


# ----------------------------------------------------------------------------
# pose grids
# ----------------------------------------------------------------------------

# psi^4 = psi + 4 (Alexa, "Super-Fibonacci Spirals", CVPR 2022)
SUPER_FIBONACCI_PSI = 1.533751168755204288118041
SUPER_FIBONACCI_PHI = math.sqrt(2.0)


def super_fibonacci_quaternions(n):
    """Deterministic, near-uniform unit quaternions (wxyz) on S^3.

    Alexa 2022, Algorithm 1: s = i + 1/2, r = sqrt(s/n), R = sqrt(1 - s/n),
    alpha = 2 pi s / phi, beta = 2 pi s / psi,
    q = (r sin alpha, r cos alpha, R sin beta, R cos beta).
    """
    i = np.arange(n, dtype=np.float64)
    s = i + 0.5
    r = np.sqrt(s / n)
    rr = np.sqrt(1.0 - s / n)
    alpha = 2 * np.pi * s / SUPER_FIBONACCI_PHI
    beta = 2 * np.pi * s / SUPER_FIBONACCI_PSI
    return np.stack([r * np.sin(alpha), r * np.cos(alpha), rr * np.sin(beta), rr * np.cos(beta)], -1)


def offset_grid(n_side, step):
    """n_side x n_side in-plane offsets (Angstrom), C order over (x, y)."""
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * step
    ox, oy = np.meshgrid(c, c, indexing="ij")
    return np.stack([ox.ravel(), oy.ravel()], -1)


def geodesic_degrees(qa, qb):
    """SO(3) geodesic angle between unit quaternions, arccos((tr(Ra^T Rb) - 1) / 2)."""
    d = np.abs(np.sum(np.asarray(qa) * np.asarray(qb), axis=-1))
    return np.rad2deg(2 * np.arccos(np.clip(d, -1.0, 1.0)))


# ----------------------------------------------------------------------------
# NCC objective
# ----------------------------------------------------------------------------


def rfft_weights(n):
    """Half-plane weights: 1 on columns 0 and n/2, 2 elsewhere; 0 at DC and the
    three self-conjugate bins (zeroed by the shift), even n."""
    m = n // 2
    w = np.full((n, m + 1), 2.0)
    w[:, 0] = 1.0
    w[:, m] = 1.0
    w[0, 0] = 0.0
    w[m, 0] = 0.0
    w[0, m] = 0.0
    w[m, m] = 0.0
    return w


def observed_spectrum(images):
    """Y = rfftn(ifftshift(y)) in float64, cryojax/ndimage/_fft.py:80-99."""
    return O.rfftn(np.asarray(images, dtype=np.float64))


def ncc_scores(grid, y_spec, ctf_params, quats, offsets, pixel_size, voltage_kv=300.0, cs_mm=2.7, ac=0.1,
               wrap_negative_indices=True):
    """FP64 -NCC for every (image, rotation, offset): shape (n_img, n_rot, n_off).

    y_spec: (n_img, N, N/2+1) complex128; ctf_params: dict of per-image arrays
    defocus, astigmatism, astigmatism_angle. Templates follow
    compute_fourier_image (oracle.imaging) with the image's CTF and the
    offset's shift, and the gather's out-of-range rule `wrap_negative_indices`.
    """
    n = grid.shape[0]
    w = rfft_weights(n)
    freq = O.safe_multiply_by_constant(O.make_frequency_grid((n, n), outputs_rfftfreqs=True), 1.0 / pixel_size)
    ctf = O.astigmatic_ctf(freq, voltage_kv, ctf_params["defocus"], ctf_params["astigmatism"],
                           ctf_params["astigmatism_angle"], cs_mm, ac)  # (n_img, N, M+1)
    shifts = O.translation_operator(freq, offsets)  # (n_off, N, M+1)
    slices = O.extract_slice(grid, quats, wrap_negative_indices=wrap_negative_indices)  # (n_rot, N, M+1)
    ynorm = np.sqrt(np.einsum("kl,ikl->i", w, np.abs(y_spec) ** 2))
    # |T|^2 = ctf^2 |P|^2 (|shift| = 1; excluded bins carry w = 0)
    tnorm2 = np.einsum("kl,ikl,rkl->ir", w, ctf**2, np.abs(slices) ** 2, optimize=True)
    out = np.empty((y_spec.shape[0], quats.shape[0], offsets.shape[0]))
    for o in range(offsets.shape[0]):
        yt = w * y_spec * ctf * np.conj(shifts[o])  # fold CTF and shift into the observed side
        num = (np.einsum("ikl,rkl->ir", yt.real, slices.real, optimize=True)
               + np.einsum("ikl,rkl->ir", yt.imag, slices.imag, optimize=True))
        out[:, :, o] = -num / (ynorm[:, None] * np.sqrt(tnorm2))
    return out


def ncc_grid_search(grid, y_spec, ctf_params, quats, offsets, pixel_size, batch_size=None, **kw):
    """run_grid_search over tree_grid = (quats, offsets) of the per-image -NCC.

    Returns (best -NCC per image, best raveled index rot * n_off + off).
    """
    def fn(point, args):
        q, o = point
        if batch_size is None:
            return ncc_scores(grid, y_spec, ctf_params, q[None], o[None], pixel_size, **kw)[:, 0, 0]
        # batched points: evaluate each (q_b, o_b) pair
        vals = [ncc_scores(grid, y_spec, ctf_params, q[b:b + 1], o[b:b + 1], pixel_size, **kw)[:, 0, 0]
                for b in range(q.shape[0])]
        return np.stack(vals)

    st = run_grid_search_minimum(fn, (quats, offsets), None, batch_size)
    return st.current_minimum_eval, st.current_best_raveled_index


def ncc_exhaustive(grid, y_spec, ctf_params, quats, offsets, pixel_size, **kw):
    """Same minimum as ncc_grid_search, evaluated as one score table (lowest
    raveled index wins exact ties)."""
    s = ncc_scores(grid, y_spec, ctf_params, quats, offsets, pixel_size, **kw)
    flat = s.reshape(s.shape[0], -1)
    idx = np.argmin(flat, axis=1)  # first occurrence = lowest raveled index
    return flat[np.arange(flat.shape[0]), idx], idx.astype(np.int64)


# ----------------------------------------------------------------------------
# exhaustive search over a large rotation grid, in parallel chunks
# ----------------------------------------------------------------------------

_WORKER = {}


def _chunk_init(grid, y_spec, ctf_params, offsets, pixel_size, kw):
    from threadpoolctl import threadpool_limits

    threadpool_limits(1)
    _WORKER.update(grid=grid, y_spec=y_spec, ctf_params=ctf_params, offsets=offsets, pixel_size=pixel_size, kw=kw)


def _chunk_best(quats_and_r0):
    quats, r0 = quats_and_r0
    w = _WORKER
    s = ncc_scores(w["grid"], w["y_spec"], w["ctf_params"], quats, w["offsets"], w["pixel_size"], **w["kw"])
    flat = s.reshape(s.shape[0], -1)
    idx = np.argmin(flat, axis=1)  # first occurrence in the chunk
    return flat[np.arange(flat.shape[0]), idx], idx.astype(np.int64) + r0 * w["offsets"].shape[0]


def ncc_exhaustive_chunked(grid, y_spec, ctf_params, quats, offsets, pixel_size, chunk=500, workers=None, **kw):
    """ncc_exhaustive over rotation chunks evaluated by worker processes (spawned, NumPy only),
    merged in raveled-index order with a strict `<`: the lowest raveled index still wins exact
    ties across chunks (search_method.py:247,254). Same result as ncc_exhaustive."""
    import multiprocessing as mp
    import os
    from concurrent.futures import ProcessPoolExecutor

    quats = np.asarray(quats, dtype=np.float64)
    n_rot = quats.shape[0]
    jobs = [(quats[r0:r0 + chunk], r0) for r0 in range(0, n_rot, chunk)]
    if workers is None:
        workers = min(16, len(os.sched_getaffinity(0)))
    workers = max(1, min(workers, len(jobs)))
    init = (np.asarray(grid), np.asarray(y_spec), {k: np.asarray(v) for k, v in ctf_params.items()},
            np.asarray(offsets, dtype=np.float64), float(pixel_size), kw)
    if workers == 1:
        _chunk_init(*init)
        results = [_chunk_best(j) for j in jobs]
    else:
        with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn"), initializer=_chunk_init,
                                 initargs=init) as pool:
            results = list(pool.map(_chunk_best, jobs))
    n_img = np.asarray(y_spec).shape[0]
    best = np.full(n_img, np.inf)
    best_idx = np.zeros(n_img, dtype=np.int64)
    for v, i in results:  # ascending raveled ranges
        less = v < best
        best = np.where(less, v, best)
        best_idx = np.where(less, i, best_idx)
    return best, best_idx
