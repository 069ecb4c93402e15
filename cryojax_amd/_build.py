"""Build the gfx950 HIP library `cryojax_amd/_lib/libcryomi.so` with hipcc.

Each source under csrc/ compiles to its own object in parallel (one hipcc per
file), then one link step makes the shared library.
"""

from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libcryomi.so"
OBJDIR = LIBDIR / "obj"
ROOT = PKG.parent

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 (v_pk_*_f32) stays off in every device code object
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def sources():
    return sorted(CSRC.glob("*.hip"))


def _filter(text: str) -> str:
    # the host pass does not know the device feature named above
    return "\n".join(line for line in text.splitlines() if "packed-fp32-ops" not in line)


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    return res.returncode, _filter(res.stdout + res.stderr)


def build(verbose: bool = False) -> Path:
    """Compile every HIP source under csrc/ into one shared library (always rebuilds)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    LIBDIR.mkdir(exist_ok=True)
    OBJDIR.mkdir(exist_ok=True)
    srcs = sources()
    objs = [OBJDIR / (s.stem + ".o") for s in srcs]
    cmds = [
        [hipcc, *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-c", "-o", str(o), str(s)] for s, o in zip(srcs, objs)
    ]
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(_run, cmds))
    failed = []
    for cmd, (rc, out) in zip(cmds, results):
        if verbose:
            print(" ".join(cmd), flush=True)
            if out.strip():
                print(out)
        if rc != 0:
            failed.append(f"{cmd[-1]} ({rc}):\n{out}")
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
    tmp = LIB.with_suffix(".so.tmp")
    link = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    rc, out = _run(link)
    if verbose:
        print(" ".join(link), flush=True)
        if out.strip():
            print(out)
    if rc != 0:
        raise RuntimeError(f"hipcc link failed ({rc}):\n{out}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
