"""Autotuner logic on the CPU (search order, cache round trip, tuning keys)
and one tuning pass on the GPU."""

import numpy as np
import pytest

from cryojax_amd import _lib, autotune


def test_search_picks_fastest_and_keeps_first_on_ties():
    cands = [{"k": i} for i in range(5)]
    times = {0: 3.0, 1: 2.0, 2: None, 3: 2.0, 4: 5.0}
    res = autotune.search(cands, lambda c: times[c["k"]])
    assert res.best == {"k": 1} and res.best_ms == 2.0
    assert [c["k"] for c, _ in res.table] == [0, 1, 2, 3, 4]
    with pytest.raises(RuntimeError):
        autotune.search(cands, lambda c: None)


def test_cache_round_trip(tmp_path):
    c = autotune.TuningCache(tmp_path / "t.json")
    assert c.get("dev", "sig") is None
    c.put("dev", "sig", {"groups.threads": 512, "groups.wg_per_cu": 2}, 1.5)
    c.put("dev", "sig2", {"a": 1})
    c2 = autotune.TuningCache(tmp_path / "t.json")
    assert c2.get("dev", "sig") == {"groups.threads": 512, "groups.wg_per_cu": 2, "_ms": 1.5}
    assert c2.get("dev", "sig2") == {"a": 1}
    (tmp_path / "bad.json").write_text("not json")
    assert autotune.TuningCache(tmp_path / "bad.json").get("dev", "sig") is None


def test_tuning_keys_roundtrip_and_validation():
    lib = _lib.load()
    try:
        autotune.apply({"groups.threads": 512, "groups.wg_per_cu": 3, "persistent_wg_per_cu.192": 2})
        assert autotune.current(["groups.threads", "groups.wg_per_cu", "persistent_wg_per_cu.192"]) == {
            "groups.threads": 512, "groups.wg_per_cu": 3, "persistent_wg_per_cu.192": 2}
        assert lib.cryomi_get_tuning(b"persistent_wg_per_cu.128") == -1
        for key, val in [(b"groups.threads", 300), (b"groups.wg_per_cu", 65), (b"persistent_wg_per_cu.98", 1),
                         (b"persistent_wg_per_cu.x", 1), (b"nope", 1), (b"persistent_wg_per_cu.64", -2)]:
            assert lib.cryomi_set_tuning(key, val) == _lib.CRYOMI_ERR_INVALID_ARGUMENT, key
    finally:
        autotune.apply({"groups.threads": 256, "groups.wg_per_cu": 0, "persistent_wg_per_cu.192": -1})


@pytest.mark.gpu
def test_tune_groups_on_device(tmp_path):
    import torch

    from cryojax_amd import batching, ops
    from tests.helpers import draw, packed_params

    groups = []
    for i, (n, c) in enumerate([(128, 300), (192, 200), (64, 500)]):
        g = (torch.randn((n, n, n), device="cuda:0") + 1j * torch.randn((n, n, n), device="cuda:0")).to(
            torch.complex64)
        p = ops.params_to_device(packed_params(draw(c, i), 1.0), "cuda:0")
        groups.append(batching.VolumeGroup(ops.grid_to_bricks(g), n, p))
    cache = autotune.TuningCache(tmp_path / "t.json")
    try:
        res = autotune.tune_groups(groups, wg_per_cu=(0, 2), reps=2, warmup=1, cache=cache)
        print("tuned", res.best, f"{res.best_ms:.3f} ms", [(c, None if m is None else round(m, 3)) for c, m in res.table])
        assert len(res.table) == 4 and np.isfinite(res.best_ms)
        assert autotune.current(["groups.threads", "groups.wg_per_cu"]) == res.best
        hit = autotune.tune_groups(groups, cache=cache)  # served from the cache
        assert hit.best == res.best and hit.table == []
        box = 128
        r2 = autotune.tune_forward(groups[0].grid, box, groups[0].params,
                                   _lib.APPLY_CTF | _lib.APPLY_TRANSLATION | _lib.GRID_BRICKED, wg_per_cu=(0, 1, 2),
                                   reps=2, warmup=1, cache=cache)
        assert f"persistent_wg_per_cu.{box}" in r2.best
    finally:
        autotune.apply({"groups.threads": 256, "groups.wg_per_cu": 0, "persistent_wg_per_cu.128": -1})
