"""CPU oracle for the cryoJAX image-formation hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product (`cryojax_amd`) may import,
call or execute anything in this package. Only `tests/`,
`__graft_entry__.smoke()` (as the checker) and `bench.py`'s `cpu_baseline`
leg (as the timed CPU restatement) use it.

The oracle is a NumPy restatement of cryoJAX's algorithm, written from the
reference's source text (it was never imported or run; JAX is absent from this
image). Each function cites the reference file:line it follows. By default it
computes in float64/complex128, which stands in for cryoJAX's own
`simulate()` under `jax_enable_x64`; passing `dtype=np.float32` gives the FP32
restatement used as the reported CPU baseline.

Pinning (what checks each part against something other than itself) is listed
in each module's header; functions with no such check say "parity unpinned".
"""
