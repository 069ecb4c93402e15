"""Minimal MRC (mode 2, float32) reader for the copied fixture volume."""

import struct

import numpy as np


def read_mrc(path):
    with open(path, "rb") as f:
        header = f.read(1024)
        nx, ny, nz, mode = struct.unpack("<4i", header[:16])
        mx, my, mz = struct.unpack("<3i", header[28:40])
        cella = struct.unpack("<3f", header[40:52])
        nsymbt = struct.unpack("<i", header[92:96])[0]
        if mode != 2:
            raise ValueError(f"only MRC mode 2 (float32) is supported, got {mode}")
        f.seek(1024 + nsymbt)
        data = np.frombuffer(f.read(4 * nx * ny * nz), dtype="<f4").reshape(nz, ny, nx)
    voxel = cella[0] / mx if mx else 1.0
    return np.array(data, dtype=np.float32), float(voxel)
