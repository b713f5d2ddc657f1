"""The network file reader behind mrwarm_load / dwi2fod -warm_net (host/npz.cpp, no GPU needed):
the .npz files numpy writes (stored and deflated members, .npy format versions 1 and 2), and clear
errors for files that are not networks."""
import ctypes as C
import os
import zipfile

import numpy as np
import pytest

from mrtrix3_amd import _lib

NET = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "warm_net_single_shell_brain.npz")


def file_dims(path):
    lib = _lib.load()
    lib.mrwarm_file_dims.argtypes = [C.c_char_p] + [C.POINTER(C.c_int)] * 4 + [C.POINTER(C.c_size_t)]
    v = [C.c_int() for _ in range(4)]
    n = C.c_size_t()
    rc = lib.mrwarm_file_dims(path.encode(), *[C.byref(x) for x in v], C.byref(n))
    if rc:
        raise ValueError(lib.mrcsd_last_error().decode())
    return [x.value for x in v] + [n.value]


def test_reads_the_committed_network():
    with np.load(NET, allow_pickle=False) as z:
        dims, n = z["dims"].tolist(), z["params"].size
    assert file_dims(NET) == dims + [n] == [64, 256, 45, 300, 236889]


@pytest.mark.parametrize("writer", ["savez", "savez_compressed"])
def test_stored_and_deflated_members(tmp_path, writer):
    path = str(tmp_path / "net.npz")
    getattr(np, writer)(path, dims=np.array([60, 256, 45, 300], dtype=np.int64),
                        params=np.arange(1000, dtype=np.float32))
    assert file_dims(path) == [60, 256, 45, 300, 1000]


def test_npy_version_2_header(tmp_path):
    path = str(tmp_path / "v2.npz")
    with zipfile.ZipFile(path, "w") as zf:
        for name, arr in (("dims", np.array([8, 256, 6, 300], dtype=np.int32)), ("params", np.ones(7, np.float32))):
            with zf.open(name + ".npy", "w") as f:
                np.lib.format.write_array(f, arr, version=(2, 0))
    assert file_dims(path) == [8, 256, 6, 300, 7]


def test_malformed_files(tmp_path):
    bad = str(tmp_path / "bad.npz")
    open(bad, "wb").write(b"not a zip file at all, but long enough to look for a directory")
    with pytest.raises(ValueError, match="malformed"):
        file_dims(bad)
    other = str(tmp_path / "other.npz")
    np.savez(other, weights=np.zeros(3, np.float32))
    with pytest.raises(ValueError, match="'dims' and 'params'"):
        file_dims(other)
    f64 = str(tmp_path / "f64.npz")
    np.savez(f64, dims=np.array([1, 2, 3, 4]), params=np.zeros(3))
    with pytest.raises(ValueError, match="float32"):
        file_dims(f64)
    with pytest.raises(ValueError, match="cannot open"):
        file_dims(str(tmp_path / "missing.npz"))
