"""Launch-shape autotuner for the persistent image kernels.

The reference leaves launch shapes to XLA's compiler; here they are tuning
values of the C ABI (cryomi_set_tuning): resident workgroups per CU of the
persistent kernels, per box, and the grouped kernel's workgroup size (which
fixes its column/row tile shapes) and workgroups per CU. The tuner times each
candidate on the caller's own workload with HIP events, checks that every
candidate forms bit-identical images, keeps the fastest, and remembers the
choice per (device, workload) in a JSON cache.
"""

from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from pathlib import Path

from . import _lib

DEFAULT_CACHE = Path(os.environ.get("CRYOMI_TUNING_CACHE", Path.home() / ".cache" / "cryojax_amd" / "tuning.json"))


@dataclass
class TuneResult:
    best: dict
    best_ms: float
    table: list = field(default_factory=list)  # (config, ms) for every candidate


def search(candidates, measure):
    """The candidate with the smallest measure(candidate) (ms); ties keep the
    earlier candidate. `measure` returns None for a candidate that cannot run."""
    table = []
    best, best_ms = None, float("inf")
    for cand in candidates:
        ms = measure(cand)
        table.append((dict(cand), ms))
        if ms is not None and ms < best_ms:
            best, best_ms = dict(cand), ms
    if best is None:
        raise RuntimeError("no tuning candidate could run")
    return TuneResult(best, best_ms, table)


def apply(config):
    """Set every key of a tuning dict through cryomi_set_tuning."""
    lib = _lib.load()
    for key, value in config.items():
        _lib.check(lib.cryomi_set_tuning(key.encode(), int(value)), f"cryomi_set_tuning({key})")


def current(keys):
    lib = _lib.load()
    return {k: lib.cryomi_get_tuning(k.encode()) for k in keys}


class TuningCache:
    """JSON map: device key -> workload signature -> tuning dict."""

    def __init__(self, path=DEFAULT_CACHE):
        self.path = Path(path)

    def _read(self):
        try:
            with open(self.path) as f:
                data = json.load(f)
            return data if isinstance(data, dict) else {}
        except (OSError, ValueError):
            return {}

    def get(self, device_key, signature):
        return self._read().get(device_key, {}).get(signature)

    def put(self, device_key, signature, config, ms=None):
        data = self._read()
        entry = dict(config)
        if ms is not None:
            entry["_ms"] = float(ms)
        data.setdefault(device_key, {})[signature] = entry
        self.path.parent.mkdir(parents=True, exist_ok=True)
        tmp = self.path.with_suffix(".tmp")
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def device_key(device=None):
    import torch

    p = torch.cuda.get_device_properties(device if device is not None else torch.cuda.current_device())
    return f"{p.name}|{getattr(p, 'gcnArchName', '')}|cu{p.multi_processor_count}"


def _time_ms(fn, reps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def groups_signature(groups):
    parts = sorted((g.box, int(g.params.shape[0]), int(bool(g.bricked))) for g in groups)
    return "groups:" + ",".join(f"{b}x{n}{'b' if br else 's'}" for b, n, br in parts)


def tune_groups(groups, threads=(256, 512), wg_per_cu=(0, 1, 2, 3, 4), reps=5, warmup=2, cache=None,
                use_cache=True):
    """Tune "groups.threads" / "groups.wg_per_cu" for cryomi_simulate_groups on
    these groups, apply the best and return the TuneResult."""
    import torch

    from . import batching

    sig = groups_signature(groups)
    dev = groups[0].grid.device
    cache = cache if cache is not None else TuningCache()
    key = device_key(dev)
    if use_cache:
        hit = cache.get(key, sig)
        if hit is not None:
            cfg = {k: v for k, v in hit.items() if not k.startswith("_")}
            apply(cfg)
            return TuneResult(cfg, hit.get("_ms", float("nan")), [])
    before = current(["groups.threads", "groups.wg_per_cu"])
    apply({"groups.threads": 256, "groups.wg_per_cu": 0})
    reference = [o.clone() for o in batching.simulate_groups(groups)]
    ws = torch.empty(0, dtype=torch.uint8, device=dev)

    def measure(cfg):
        nonlocal ws
        apply(cfg)
        lib = _lib.load()
        arr, _ = batching._table(groups)
        import ctypes

        need = lib.cryomi_simulate_groups_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), len(groups))
        if ws.numel() < need:
            ws = torch.empty(int(need), dtype=torch.uint8, device=dev)
        batching.simulate_groups(groups, ws)
        if not all(torch.equal(a, g.out) for a, g in zip(reference, groups)):
            raise RuntimeError(f"tuning candidate {cfg} changed the images")
        return _time_ms(lambda: batching.simulate_groups(groups, ws), reps, warmup)

    cands = [{"groups.threads": t, "groups.wg_per_cu": w} for t in threads for w in wg_per_cu]
    try:
        res = search(cands, measure)
    except Exception:
        apply(before)
        raise
    apply(res.best)
    cache.put(key, sig, res.best, res.best_ms)
    return res


def tune_forward(grid, box, params, flags, wg_per_cu=(0, 1, 2, 3, 4), reps=5, warmup=2, cache=None,
                 use_cache=True):
    """Tune "persistent_wg_per_cu.<box>" for cryomi_simulate on this workload."""
    import torch

    from . import ops

    n = int(params.shape[0])
    sig = f"simulate:{box}x{n}:flags{int(flags)}"
    key_name = f"persistent_wg_per_cu.{int(box)}"
    cache = cache if cache is not None else TuningCache()
    key = device_key(grid.device)
    if use_cache:
        hit = cache.get(key, sig)
        if hit is not None:
            cfg = {k: v for k, v in hit.items() if not k.startswith("_")}
            apply(cfg)
            return TuneResult(cfg, hit.get("_ms", float("nan")), [])
    before = current([key_name])
    apply({key_name: 0})
    out = ops.simulate_images(grid, params, flags, box=box)
    reference = out.clone()
    lib = _lib.load()
    ws = torch.empty(0, dtype=torch.uint8, device=grid.device)

    def measure(cfg):
        nonlocal ws
        apply(cfg)
        need = lib.cryomi_simulate_workspace_bytes(int(box), n, int(flags))
        if ws.numel() < need:
            ws = torch.empty(int(need), dtype=torch.uint8, device=grid.device)
        ops.simulate_images(grid, params, flags, out=out, workspace=ws, box=box)
        if not torch.equal(out, reference):
            raise RuntimeError(f"tuning candidate {cfg} changed the images")
        return _time_ms(lambda: ops.simulate_images(grid, params, flags, out=out, workspace=ws, box=box), reps,
                        warmup)

    try:
        res = search([{key_name: w} for w in wg_per_cu], measure)
    except Exception:
        apply(before)
        raise
    apply(res.best)
    cache.put(key, sig, res.best, res.best_ms)
    return res
