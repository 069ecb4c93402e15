"""MRC2014 and STAR file I/O (host side of the particle-stack writer).

Mirrors cryojax.io (cryojax/io/_mrc.py:14-228 read_array_from_mrc,
write_image_to_mrc, write_image_stack_to_mrc, write_volume_to_mrc;
cryojax/io/_starfile.py:12-38 read_starfile, write_starfile). The reference
delegates the formats to the third-party `mrcfile` and `starfile` packages,
which this image does not have; both formats are implemented here from their
published specifications:

* MRC2014 (CCP-EM): a 1024-byte little-endian header of 56 words plus ten
  80-byte labels, an optional extended header of `nsymbt` bytes, then the
  data in (z, y, x) C order. Files are written as mrcfile writes them (the
  reference's test fixtures hold such headers): mode 2 (float32), map "MAP ",
  machine stamp 0x44 0x44 0 0, nversion 20141, ispg 0 with mz = 1 for an
  image or image stack and ispg 1 with mz = nz for a volume, cell lengths =
  (mx, my, mz) * voxel size, and dmin / dmax / dmean / rms (population
  standard deviation) of the data.
* STAR (RELION's flavour): `data_<name>` blocks holding either one `loop_`
  table (`_rlnName #i` headers, whitespace-separated rows) or `_rlnName value`
  pairs; written as the starfile package writes them (floats "%.6f",
  tab-separated rows, two blank lines between blocks).
"""

from __future__ import annotations

import datetime
import pathlib
import re
import shlex

import numpy as np
import pandas as pd

# ---------------------------------------------------------------------------
# MRC2014
# ---------------------------------------------------------------------------

_MODE_DTYPES = {0: np.int8, 1: np.int16, 2: np.float32, 4: np.complex64, 6: np.uint16, 12: np.float16}
_DTYPE_MODES = {np.dtype(v).newbyteorder("<"): k for k, v in _MODE_DTYPES.items()}
HEADER_BYTES = 1024


class MrcHeader:
    """The fields of an MRC2014 header this module reads and writes."""

    def __init__(self, raw: bytes):
        if len(raw) < HEADER_BYTES:
            raise ValueError("MRC file is shorter than its 1024-byte header")
        stamp = raw[212:216]
        self.byteorder = ">" if stamp[0] == 0x11 else "<"  # 0x44 0x44 (or 0x44 0x41): little-endian
        i4 = np.frombuffer(raw[:HEADER_BYTES], dtype=self.byteorder + "i4")
        f4 = np.frombuffer(raw[:HEADER_BYTES], dtype=self.byteorder + "f4")
        self.nx, self.ny, self.nz = (int(v) for v in i4[0:3])
        self.mode = int(i4[3])
        self.mx, self.my, self.mz = (int(v) for v in i4[7:10])
        self.cella = tuple(float(v) for v in f4[10:13])
        self.mapcrs = tuple(int(v) for v in i4[16:19])
        self.dmin, self.dmax, self.dmean = (float(v) for v in f4[19:22])
        self.ispg = int(i4[22])
        self.nsymbt = int(i4[23])
        self.nversion = int(i4[27])
        self.map = raw[208:212]
        self.rms = float(f4[54])
        nlabl = max(0, min(int(i4[55]), 10))
        self.labels = [raw[224 + 80 * k:304 + 80 * k].rstrip(b" \x00").decode(errors="replace") for k in range(nlabl)]

    @property
    def voxel_size(self):
        """cella / (mx, my, mz) per axis (x, y, z), 0 where the sampling is 0."""
        m = (self.mx, self.my, self.mz)
        return tuple(c / s if s else 0.0 for c, s in zip(self.cella, m))

    def data_shape(self):
        if self.nz == 1 and self.ispg == 0:
            return (self.ny, self.nx)
        if self.ispg >= 401 and self.mz > 0 and self.nz % self.mz == 0:
            return (self.nz // self.mz, self.mz, self.ny, self.nx)
        return (self.nz, self.ny, self.nx)


def _header_bytes(data: np.ndarray, voxel_size, ispg: int, mz: int) -> bytes:
    nz, ny, nx = (1, *data.shape) if data.ndim == 2 else data.shape
    i4 = np.zeros(256, dtype="<i4")
    f4 = i4.view("<f4")
    i4[0:3] = (nx, ny, nz)
    i4[3] = _DTYPE_MODES[data.dtype.newbyteorder("<")]
    i4[7:10] = (nx, ny, mz)
    vs = np.broadcast_to(np.asarray(voxel_size, dtype=np.float64), (3,))
    f4[10:13] = np.array([nx, ny, mz], dtype=np.float64) * vs
    f4[13:16] = 90.0
    i4[16:19] = (1, 2, 3)
    if data.size:
        d = np.abs(data) if np.iscomplexobj(data) else data  # reductions in the data's own precision, as mrcfile
        f4[19:22] = (d.min(), d.max(), d.mean())
        f4[54] = d.std()
    else:
        f4[19:22] = (0.0, -1.0, -2.0)  # MRC2014 "not well determined"
        f4[54] = -1.0
    i4[22] = ispg
    i4[27] = 20141
    raw = bytearray(i4.tobytes())
    raw[208:212] = b"MAP "
    raw[212:216] = bytes([0x44, 0x44, 0x00, 0x00])
    i4_55 = np.array([1], dtype="<i4").tobytes()
    raw[220:224] = i4_55
    stamp = datetime.datetime.now().strftime("%Y-%m-%d %H:%M:%S")
    label = "Created by cryojax_amd".ljust(60) + stamp
    raw[224:304] = label.encode()[:80].ljust(80)
    return bytes(raw)


def _validate_filename_and_return_suffix(filename):
    """_mrc.py:214-223: the leading suffix must be .mrc or .mrcs."""
    suffixes = pathlib.Path(filename).suffixes
    if len(suffixes) == 0 or suffixes[0] not in [".mrc", ".mrcs"]:
        raise OSError(f"Filename should include .mrc or .mrcs suffix. Got filename {filename}.")
    return suffixes[0]


def read_mrc(filename, mmap=False):
    """(header, data) of an MRC file; data is a read-only memmap when mmap."""
    path = pathlib.Path(filename)
    with open(path, "rb") as fh:
        head = MrcHeader(fh.read(HEADER_BYTES))
    if head.mode not in _MODE_DTYPES:
        raise ValueError(f"unsupported MRC mode {head.mode} in {filename}")
    dtype = np.dtype(_MODE_DTYPES[head.mode]).newbyteorder(head.byteorder)
    shape = head.data_shape()
    offset = HEADER_BYTES + max(head.nsymbt, 0)
    count = int(np.prod(shape))
    if mmap:
        data = np.memmap(path, dtype=dtype, mode="r", offset=offset, shape=shape)
    else:
        with open(path, "rb") as fh:
            fh.seek(offset)
            data = np.fromfile(fh, dtype=dtype, count=count)
        if data.size != count:
            raise ValueError(f"MRC file {filename} holds fewer values than its header declares")
        data = data.reshape(shape)
    return head, data


def read_array_from_mrc(filename, *, loads_grid_spacing=False, mmap=False, permissive=False):
    """_mrc.py:45-114: the data (float64 unless mmap) and optionally its grid spacing."""
    _validate_filename_and_return_suffix(filename)
    head, array = read_mrc(filename, mmap=mmap)
    if not permissive and head.map != b"MAP ":
        raise ValueError(f"{filename} is not a valid MRC file (map field {head.map!r})")
    ndim = array.ndim
    vs = head.voxel_size
    spacing = np.asarray(vs[:2] if ndim == 2 or (ndim == 3 and head.ispg == 0) else vs, dtype=float)
    if not mmap:
        array = array.astype(np.float64)
    if loads_grid_spacing:
        if not np.all(spacing != 0):
            raise AssertionError("Mrcfile.voxel_size must be set if reading the grid spacing. Found that "
                                 "Mrcfile.voxel_size = (0.0, 0.0, 0.0)")
        if not np.all(spacing == spacing[0]):
            raise AssertionError("Mrcfile.voxel_size must be same in all dimensions.")
        return array, float(spacing[0])
    return array


def _as_host_float32(x):
    if hasattr(x, "detach"):  # torch tensor (device or host)
        x = x.detach().cpu().numpy()
    a = np.asarray(x)
    if a.dtype == np.float64 or a.dtype.kind in "iub" and a.dtype not in (np.int8, np.int16, np.uint16):
        a = a.astype(np.float32)
    return np.ascontiguousarray(a)


def _write_mrc(data, voxel_size, filename, overwrite, ispg, mz_of):
    path = pathlib.Path(filename)
    if path.exists() and not overwrite:
        raise ValueError(f"File '{path}' already exists; set overwrite=True to overwrite it")
    vs = voxel_size.detach().cpu().numpy() if hasattr(voxel_size, "detach") else np.asarray(voxel_size)
    vs = np.asarray(vs, dtype=np.float64).ravel()
    vs = vs[:3] if vs.size >= 3 else np.full(3, vs[0] if vs.size else 0.0)
    mz = mz_of(data)
    with open(path, "wb") as fh:
        fh.write(_header_bytes(data, vs, ispg, mz))
        data.astype(data.dtype.newbyteorder("<"), copy=False).tofile(fh)


def _no_compression(compression):
    if compression is not None:
        raise NotImplementedError("compressed MRC files are not supported")


def write_image_to_mrc(image, pixel_size, filename, overwrite=False, compression=None):
    """_mrc.py:117-149."""
    suffix = _validate_filename_and_return_suffix(filename)
    if suffix != ".mrc":
        raise OSError(f"The suffix for an image in MRC format must be .mrc. Instead, got {suffix}.")
    if image.ndim != 2:
        raise ValueError("image.ndim must be equal to 2.")
    _no_compression(compression)
    _write_mrc(_as_host_float32(image), pixel_size, filename, overwrite, 0, lambda d: 1)


def write_image_stack_to_mrc(image_stack, pixel_size, filename, overwrite=False, compression=None):
    """_mrc.py:152-184."""
    suffix = _validate_filename_and_return_suffix(filename)
    if suffix != ".mrcs":
        raise OSError(f"The suffix for an image stack in MRC format must be .mrcs. Instead, got {suffix}.")
    if image_stack.ndim != 3:
        raise ValueError("image_stack.ndim must be equal to 3.")
    _no_compression(compression)
    _write_mrc(_as_host_float32(image_stack), pixel_size, filename, overwrite, 0, lambda d: 1)


def write_volume_to_mrc(voxel_grid, voxel_size, filename, overwrite=False, compression=None):
    """_mrc.py:187-211."""
    suffix = _validate_filename_and_return_suffix(filename)
    if suffix != ".mrc":
        raise OSError(f"The suffix for a volume in MRC format must be .mrc. Instead, got {suffix}.")
    if voxel_grid.ndim != 3:
        raise ValueError("voxel_grid.ndim must be equal to 3.")
    _no_compression(compression)
    _write_mrc(_as_host_float32(voxel_grid), voxel_size, filename, overwrite, 1, lambda d: d.shape[0])


# ---------------------------------------------------------------------------
# STAR
# ---------------------------------------------------------------------------


def _validate_star_filename(filename, mode):
    """_starfile.py:31-38."""
    suffixes = pathlib.Path(filename).suffixes
    if not (len(suffixes) == 1 and suffixes[0] == ".star"):
        raise OSError(
            f"Tried to {('write' if mode == 'w' else 'read')} STAR file, but the filename does not include a "
            f"'.star' suffix. Got filename '{filename}'."
        )


def _typed_column(values):
    s = pd.Series(values, dtype=object)
    for kind in ("int", "float"):
        try:
            conv = pd.to_numeric(s, errors="raise")
        except (ValueError, TypeError):
            continue
        if kind == "int" and not pd.api.types.is_integer_dtype(conv):
            continue
        return conv
    return s.astype(object)


def _tokens(line):
    return shlex.split(line, comments=False, posix=True) if ('"' in line or "'" in line) else line.split()


def parse_star(text):
    """Blocks of a STAR file as {name: DataFrame}; key-value blocks become one-row frames."""
    blocks = {}
    lines = text.splitlines()
    i, n = 0, len(lines)

    def skip(line):
        s = line.strip()
        return not s or s.startswith("#")

    while i < n:
        line = lines[i].strip()
        i += 1
        if not line.startswith("data_"):
            continue
        name = line[5:]
        while i < n and skip(lines[i]):
            i += 1
        if i < n and lines[i].strip() == "loop_":
            i += 1
            cols = []
            while i < n and (lines[i].strip().startswith("_") or skip(lines[i])):
                s = lines[i].strip()
                if s.startswith("_"):
                    cols.append(s.split()[0][1:])
                i += 1
            rows = []
            while i < n:
                s = lines[i].strip()
                if s.startswith("data_") or s == "loop_":
                    break
                if not skip(lines[i]):
                    toks = _tokens(s)
                    if len(toks) != len(cols):
                        raise ValueError(f"STAR block data_{name}: row with {len(toks)} values for {len(cols)} columns")
                    rows.append(toks)
                i += 1
            frame = pd.DataFrame({c: _typed_column([r[k] for r in rows]) for k, c in enumerate(cols)})
            blocks[name] = frame
        else:
            pairs = {}
            while i < n:
                s = lines[i].strip()
                if s.startswith("data_"):
                    break
                if s.startswith("_"):
                    toks = _tokens(s)
                    pairs[toks[0][1:]] = toks[1] if len(toks) > 1 else ""
                i += 1
            blocks[name] = pd.DataFrame({k: _typed_column([v]) for k, v in pairs.items()})
    return blocks


def read_starfile(filename, **kwargs):
    """_starfile.py:12-19: every block of the file as {name: DataFrame}."""
    _validate_star_filename(filename, mode="r")
    return parse_star(pathlib.Path(filename).read_text())


def _format_value(v, float_format):
    if isinstance(v, (float, np.floating)):
        return "nan" if np.isnan(v) else float_format % v
    if isinstance(v, (bool, np.bool_)):
        return str(int(v))
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if v is None or (isinstance(v, float) and np.isnan(v)) or v is pd.NA:
        return "nan"
    s = str(v)
    return f'"{s}"' if (not s or re.search(r"\s", s)) else s


def format_star(blocks, float_format="%.6f", sep="\t"):
    out = [f"# Created by cryojax_amd at {datetime.datetime.now().strftime('%H:%M:%S on %d/%m/%Y')}", "", ""]
    for name, frame in blocks.items():
        if isinstance(frame, dict):
            frame = pd.DataFrame({k: [v] for k, v in frame.items()})
        out += [f"data_{name}", "", "loop_"]
        out += [f"_{c} #{k + 1}" for k, c in enumerate(frame.columns)]
        cols = [frame[c].tolist() for c in frame.columns]
        for r in range(len(frame)):
            out.append(sep.join(_format_value(col[r], float_format) for col in cols))
        out += ["", ""]
    return "\n".join(out) + "\n"


def write_starfile(starfile_data, filename, float_format="%.6f", sep="\t", **kwargs):
    """_starfile.py:22-28: {name: DataFrame} as loop_ blocks."""
    _validate_star_filename(filename, mode="w")
    pathlib.Path(filename).write_text(format_star(starfile_data, float_format=float_format, sep=sep))
