"""The commands' own .mif reader and writer (mrtrix3_amd/csrc/host/mif.cpp), compiled on the CPU
with tests/native/mif_copy.cpp: images in every supported data type and byte order, with flipped
and permuted axes and intensity scaling, read into canonical order and written back in both
layouts, compared with numpy (core/formats/mrtrix.cpp, core/formats/mrtrix_utils.h: layout
"+r"/"-r" gives the rank of an axis' stride and its direction)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mrtrix3_amd import mif

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "mrtrix3_amd", "csrc")


@pytest.fixture(scope="module")
def mif_copy(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("bin") / "mif_copy")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", SRC, "-o", exe, os.path.join(HERE, "native", "mif_copy.cpp"),
                    os.path.join(SRC, "host", "mif.cpp"), "-lpthread"], check=True)
    return exe


def write_raw(path, data, layout, dtype, scaling=None):
    """Stores data [x, y, z, ...] with the given signed layout ranks (e.g. [-1, +2, +0, +3])."""
    nd = data.ndim
    ranks = [int(r.lstrip("+-")) for r in layout]
    filed = data
    for a in range(nd):
        if str(layout[a]).startswith("-"):
            filed = np.flip(filed, a)
    order = sorted(range(nd), key=lambda a: ranks[a])      # fastest first
    filed = np.ascontiguousarray(np.transpose(filed, list(reversed(order)))).astype(dtype)
    lay = ",".join(("-" if str(layout[a]).startswith("-") else "+") + str(ranks[a]) for a in range(nd))
    head = "mrtrix image\ndim: %s\nvox: %s\nlayout: %s\ndatatype: %s\n" % (
        ",".join(str(d) for d in data.shape), ",".join(["1.5"] * nd), lay, dtype_name(dtype))
    if scaling is not None:
        head += "scaling: %r,%r\n" % scaling
    head += "file: . 512\nEND\n"
    raw = head.encode()
    with open(path, "wb") as f:
        f.write(raw + b"\0" * (512 - len(raw)) + filed.tobytes())


def dtype_name(dt):
    dt = np.dtype(dt)
    base = {"f": "Float", "i": "Int", "u": "UInt"}[dt.kind] + str(8 * dt.itemsize)
    if dt.itemsize == 1:
        return base
    return base + ("BE" if dt.byteorder == ">" else "LE")


CASES = [
    ("<f4", ["+1", "+2", "+3", "+0"], None),
    ("<f4", ["+0", "+1", "+2", "+3"], None),
    ("<f4", ["-2", "+0", "-3", "+1"], None),
    (">f4", ["+1", "-2", "+3", "+0"], (0.5, 2.0)),
    ("<f8", ["-0", "+1", "+2", "+3"], None),
    (">i2", ["+3", "-1", "+0", "-2"], (1.0, 0.25)),
    ("<u2", ["+0", "+1", "+2", "+3"], (-3.0, 1.5)),
    ("i1", ["-3", "-2", "-1", "-0"], None),
    ("u1", ["+2", "+3", "+1", "+0"], (2.0, 3.0)),
    (">i4", ["+1", "+0", "+3", "-2"], None),
    ("<u4", ["-1", "+2", "+0", "+3"], None),
]


@pytest.mark.parametrize("dt,layout,scaling", CASES)
def test_native_read_write_matches_numpy(mif_copy, tmp_path, dt, layout, scaling):
    rng = np.random.default_rng(7)
    shape = (9, 7, 5, 6)
    if np.dtype(dt).kind == "f":
        data = rng.standard_normal(shape)
    else:
        info = np.iinfo(np.dtype(dt))
        data = rng.integers(max(info.min, -1000), min(info.max, 1000), size=shape, endpoint=True)
    src = str(tmp_path / "in.mif")
    write_raw(src, data, layout, dt, scaling)
    want = np.asarray(data).astype(dt).astype(np.float64)   # the values the file holds
    if scaling is not None:
        want = scaling[0] + scaling[1] * want
    want = want.astype(np.float32)
    for vc in ("vc", "xc"):
        out = str(tmp_path / ("out_%s.mif" % vc))
        subprocess.run([mif_copy, src, out, vc], check=True, timeout=60)
        back, h = mif.read(out)
        assert h["dim"] == list(shape)
        assert np.array_equal(back, want), (dt, layout, scaling, vc)
        assert h["datatype"].lower() == "float32le"


def test_native_large_image_threads(mif_copy, tmp_path):
    # enough rows (axis 1 and up) to split over several host threads; odd sizes so ranges are ragged
    rng = np.random.default_rng(3)
    data = rng.standard_normal((13, 61, 47, 5)).astype(np.float32)
    src = str(tmp_path / "in.mif")
    write_raw(src, data, ["-1", "+2", "-3", "+0"], "<f4")
    for vc in ("vc", "xc"):
        out = str(tmp_path / ("out_%s.mif" % vc))
        subprocess.run([mif_copy, src, out, vc], check=True, timeout=60)
        back, _ = mif.read(out)
        assert np.array_equal(back, data)


def test_native_3d_and_truncated(mif_copy, tmp_path):
    data = np.arange(4 * 3 * 2, dtype=np.float32).reshape(4, 3, 2)
    src = str(tmp_path / "m.mif")
    write_raw(src, data, ["-0", "+1", "-2"], "<f4")
    out = str(tmp_path / "o.mif")
    subprocess.run([mif_copy, src, out], check=True, timeout=60)
    back, _ = mif.read(out)
    assert np.array_equal(back, data)
    raw = open(src, "rb").read()
    with open(src, "wb") as f:
        f.write(raw[:-4])
    r = subprocess.run([mif_copy, src, out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "too small" in r.stderr


@pytest.mark.parametrize("shape", [(13, 61, 47, 5), (70, 3, 2, 130), (1, 1, 1, 3)])
def test_native_transpose_to_voxel_rows(mif_copy, tmp_path, shape):
    # ragged 64 x 64 tiles on both sides; the commands' image <-> voxel-row conversion
    data = np.random.default_rng(11).standard_normal(shape).astype(np.float32)
    src, out = str(tmp_path / "t.mif"), str(tmp_path / "t.raw")
    write_raw(src, data, ["+1", "+2", "+3", "+0"], "<f4")
    subprocess.run([mif_copy, "--transpose", src, out], check=True, timeout=60)
    rows = np.fromfile(out, dtype=np.float32).reshape(-1, shape[3])
    assert np.array_equal(rows, data.transpose(2, 1, 0, 3).reshape(-1, shape[3]))
