"""MRC2014 and STAR I/O on the CPU: the reference's own fixture files (an
mrcfile-written particle stack and volume, a starfile-written STAR file) and
write -> read round trips."""

import os

import numpy as np
import pandas as pd
import pytest

from cryojax_amd import io
from tests.mrc import read_mrc as minimal_read

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
STACK = os.path.join(GOLDEN, "relion_000000.mrcs")
VOLUME = os.path.join(GOLDEN, "3j9g_potential_ps4_4.mrc")
STAR = os.path.join(GOLDEN, "relion_particles.star")


def test_read_fixture_stack_and_header_statistics():
    head, data = io.read_mrc(STACK)
    assert data.shape == (3, 16, 16) and data.dtype == np.float32
    assert head.ispg == 0 and head.mz == 1 and head.nversion == 20141 and head.map == b"MAP "
    d = data.astype(np.float64)
    # the statistics mrcfile stored: min, max, mean, population std
    np.testing.assert_allclose([head.dmin, head.dmax, head.dmean, head.rms], [d.min(), d.max(), d.mean(), d.std()],
                               rtol=1e-6)
    arr = io.read_array_from_mrc(STACK)
    assert arr.dtype == np.float64 and arr.shape == (3, 16, 16)
    np.testing.assert_array_equal(arr, d)
    with pytest.raises(AssertionError):  # this fixture's x and y voxel sizes differ (1 and 12)
        io.read_array_from_mrc(STACK, loads_grid_spacing=True)


def test_read_fixture_volume_matches_independent_reader():
    arr, spacing = io.read_array_from_mrc(VOLUME, loads_grid_spacing=True)
    ref, ref_spacing = minimal_read(VOLUME)
    assert arr.shape == (64, 64, 64)
    np.testing.assert_array_equal(arr, ref.astype(np.float64))
    assert spacing == pytest.approx(ref_spacing) and spacing == pytest.approx(4.4)
    mm = io.read_array_from_mrc(VOLUME, mmap=True)
    assert isinstance(mm, np.memmap) and mm.dtype == np.float32
    np.testing.assert_array_equal(np.asarray(mm), ref)


def header_words(path):
    raw = open(path, "rb").read(1024)
    return np.frombuffer(raw[:220], dtype="<i4"), raw


def test_written_stack_header_matches_mrcfile_layout(tmp_path):
    _, data = io.read_mrc(STACK)
    out = tmp_path / "stack.mrcs"
    io.write_image_stack_to_mrc(data, 12.0, out)
    mine, raw_mine = header_words(out)
    ref, raw_ref = header_words(STACK)
    same = list(range(0, 10)) + list(range(13, 55))  # every word but the cell lengths (the fixture's are 16, 192, 12)
    np.testing.assert_array_equal(mine[same], ref[same])
    assert raw_mine[208:216] == raw_ref[208:216]
    head, back = io.read_mrc(out)
    np.testing.assert_array_equal(back, data)
    assert head.voxel_size == pytest.approx((12.0, 12.0, 12.0))
    with pytest.raises(ValueError):
        io.write_image_stack_to_mrc(data, 12.0, out)  # exists
    io.write_image_stack_to_mrc(data * 2, 12.0, out, overwrite=True)
    np.testing.assert_array_equal(io.read_array_from_mrc(out), 2.0 * data.astype(np.float64))


def test_volume_and_image_round_trips(tmp_path, rng):
    vol = rng.standard_normal((8, 10, 12))
    io.write_volume_to_mrc(vol, 2.5, tmp_path / "v.mrc")
    head, back = io.read_mrc(tmp_path / "v.mrc")
    assert head.ispg == 1 and head.mz == 8 and back.shape == (8, 10, 12) and back.dtype == np.float32
    np.testing.assert_array_equal(back, vol.astype(np.float32))
    arr, spacing = io.read_array_from_mrc(tmp_path / "v.mrc", loads_grid_spacing=True)
    assert spacing == pytest.approx(2.5)
    img = rng.standard_normal((9, 7)).astype(np.float32)
    io.write_image_to_mrc(img, 1.5, tmp_path / "i.mrc")
    arr, spacing = io.read_array_from_mrc(tmp_path / "i.mrc", loads_grid_spacing=True)
    assert arr.shape == (9, 7) and spacing == pytest.approx(1.5)
    np.testing.assert_array_equal(arr, img.astype(np.float64))


def test_suffix_and_rank_errors(tmp_path):
    z = np.zeros((2, 4, 4), dtype=np.float32)
    with pytest.raises(OSError):
        io.write_image_stack_to_mrc(z, 1.0, tmp_path / "s.mrc")
    with pytest.raises(OSError):
        io.write_volume_to_mrc(z, 1.0, tmp_path / "v.mrcs")
    with pytest.raises(OSError):
        io.read_array_from_mrc(tmp_path / "x.txt")
    with pytest.raises(ValueError):
        io.write_image_to_mrc(z, 1.0, tmp_path / "i.mrc")
    with pytest.raises(NotImplementedError):
        io.write_volume_to_mrc(z, 1.0, tmp_path / "c.mrc", compression="gzip")


def test_read_fixture_star():
    data = io.read_starfile(STAR)
    assert list(data) == ["particles", "optics"]
    p, o = data["particles"], data["optics"]
    assert len(p) == 3 and len(o) == 1
    assert p["rlnDefocusU"].tolist() == pytest.approx([10061.054688, 12291.794922, 13190.399414])
    assert p["rlnAnglePsi"].iloc[2] == pytest.approx(-161.981888)
    assert p["rlnImageName"].tolist() == ["0000001@000000.mrcs", "0000002@000000.mrcs", "0000003@000000.mrcs"]
    assert pd.api.types.is_integer_dtype(p["rlnClassNumber"]) and int(o["rlnImageSize"].iloc[0]) == 16
    assert float(o["rlnImagePixelSize"].iloc[0]) == 12.0 and float(o["rlnVoltage"].iloc[0]) == 300.0


def test_star_round_trip_and_key_value_blocks(tmp_path):
    data = io.read_starfile(STAR)
    io.write_starfile(data, tmp_path / "out.star")
    back = io.read_starfile(tmp_path / "out.star")
    for k in data:
        pd.testing.assert_frame_equal(back[k], data[k], check_exact=False, rtol=1e-6)
    text = "data_general\n\n_rlnImageSize   64\n_rlnComment \"a b\"\n\ndata_x\nloop_\n_rlnA #1\n_rlnB #2\n1 x\n2 y\n"
    blocks = io.parse_star(text)
    assert int(blocks["general"]["rlnImageSize"].iloc[0]) == 64 and blocks["general"]["rlnComment"].iloc[0] == "a b"
    assert blocks["x"]["rlnA"].tolist() == [1, 2] and blocks["x"]["rlnB"].tolist() == ["x", "y"]
    with pytest.raises(OSError):
        io.write_starfile(data, tmp_path / "out.txt")
