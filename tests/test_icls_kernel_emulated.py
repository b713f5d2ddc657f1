"""The ICLS kernel's own source (mrtrix3_amd/csrc/kernels/icls_solve.hip) compiled for the CPU
against a wavefront emulation (tests/native/emu: each of the 64 lanes a fiber, every
__syncthreads, ballot, readlane, readfirstlane and shuffle a collective that all lanes must reach
at the same call site) and run through the C ABI on the masked multi-shell case of
tests/test_gpu_icls.py: every voxel's iterations, active-set size and solution against the oracle.
This checks the kernel's control flow, indexing and lane protocol without a GPU; arithmetic is the
host compiler's (no fused multiply-adds), so solutions are compared to 1e-6."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mrtrix3_amd import icls, phantom
from oracle import orc

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "mrtrix3_amd", "csrc")
EMU = os.path.join(HERE, "native", "emu")


@pytest.fixture(scope="module")
def emu_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    tmp_path = tmp_path_factory.mktemp("emu")
    # the kernel file without its launcher (the harness launches the emulated workgroups)
    src = open(os.path.join(CSRC, "kernels", "icls_solve.hip")).read()
    cut = src.index("// ---- HIP launches ----")
    prep = tmp_path / "icls_kernel_emu.hip"
    prep.write_text(src[:cut] + "}  // namespace\n}  // namespace mrcsd\n")
    exe = str(tmp_path / "icls_emu")
    host = [os.path.join(CSRC, f) for f in ("capi_icls.cpp", "capi_common.cpp")] + \
           [os.path.join(CSRC, "host", f) for f in ("icls.cpp", "msmt.cpp", "sh.cpp", "gradient.cpp", "dense.cpp", "csd.cpp")]
    cmd = ["g++", "-std=c++17", "-O1", "-w", "-I" + os.path.join(EMU, "include"), "-I" + CSRC,
           "-I" + os.path.join(CSRC, "kernels"), "-I" + os.path.join(ROOT, "include"),
           '-DKERNEL_SOURCE="%s"' % prep, "-x", "c++", os.path.join(EMU, "harness.cpp"),
           os.path.join(EMU, "emu_runtime.cpp")] + host + ["-o", exe]
    subprocess.run(cmd, check=True, timeout=600)
    return exe


@pytest.mark.parametrize("capacity", [0, 16])
def test_icls_kernel_under_wavefront_emulation(tmp_path, emu_exe, capacity):
    """capacity 0: the default first pass (48 slots for 47 parameters); 16: every voxel outgrows
    the first pass and is queued for the 64-slot pass, which must give the same answers."""
    exe = emu_exe

    ph = phantom.multi_shell(seed=6, shape=(4, 4, 4), semi=None)
    nv = 24
    valid = np.ones(nv, dtype=np.uint8)
    valid[::5] = 0
    resp = [np.atleast_2d(np.asarray(r, dtype=np.float64)) for r in ph.response]
    d = tmp_path / "case"
    d.mkdir()
    with open(d / "case.txt", "w") as f:
        f.write("%d %d %d\n%d\n" % (ph.grad.shape[0], nv, ph.dw.shape[1], len(resp)))
        for r in resp:
            f.write("%d %d\n" % r.shape)
    np.ascontiguousarray(ph.grad, dtype=np.float64).tofile(d / "grad.bin")
    np.concatenate([r.ravel() for r in resp]).tofile(d / "resp.bin")
    np.ascontiguousarray(ph.dw[:nv], dtype=np.float32).tofile(d / "dw.bin")
    valid.tofile(d / "valid.bin")
    r = subprocess.run([exe, str(d), str(capacity)], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "EMU:" not in r.stderr, r.stderr[-3000:]
    if capacity == 16:
        import re
        queued = int(re.search(r"pass 0: capacity 16, queued for the next pass (\d+)", r.stderr).group(1))
        assert queued >= 10   # the valid voxels whose active set passes 16 constraints

    s = icls.MSMTShared(ph.grad, ph.response)
    o = orc.MsmtOracle(ph.grad, ph.response)
    n = s.nparams
    x = np.fromfile(d / "x.bin", dtype=np.float32).reshape(nv, n)
    stats = np.loadtxt(d / "iters.txt", dtype=np.int64).reshape(nv, 2)
    for v in range(nv):
        if not valid[v]:
            assert stats[v, 0] == 0xFFFF and np.all(x[v] == 0)
            continue
        xo, it, act = o.solve(ph.dw[v].astype(np.float64))
        assert stats[v, 0] == it and stats[v, 1] == int(act.sum())
        assert np.linalg.norm(x[v] - xo) <= 1e-6 * np.linalg.norm(xo)


def test_two_slots_per_lane_under_emulation(tmp_path, emu_exe):
    """Non-negative least squares with 100 unknowns and 71-83 active bounds: passes of 16, 64 and
    128 slots (two slots per lane in the last), against the oracle."""
    rng = np.random.default_rng(21)
    n, m, nvox = 100, 160, 6
    H = rng.standard_normal((m, n))
    A = np.eye(n)
    xt = rng.standard_normal((nvox, n)) - 0.9
    B = (xt @ H.T + 0.01 * rng.standard_normal((nvox, m))).astype(np.float32)
    d = tmp_path / "generic"
    d.mkdir()
    (d / "generic.txt").write_text("%d %d %d %d %.17g\n" % (m, n, n, nvox, 1e-10))
    H.astype(np.float64).tofile(d / "H.bin")
    A.astype(np.float64).tofile(d / "A.bin")
    B.tofile(d / "b.bin")
    r = subprocess.run([emu_exe, str(d), "16"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "EMU:" not in r.stderr, r.stderr[-3000:]
    assert "pass 2: capacity 128" in r.stderr
    x = np.fromfile(d / "x.bin", dtype=np.float32).reshape(nvox, n)
    stats = np.loadtxt(d / "iters.txt", dtype=np.int64).reshape(nvox, 2)
    o = orc.IclsOracle(H, A, None, con_reg=1e-10)
    for v in range(nvox):
        xo, it, act = o.solve(B[v].astype(np.float64))
        assert stats[v, 0] == it and stats[v, 1] == int(act.sum()) and stats[v, 1] > 64
        assert np.linalg.norm(x[v] - xo) <= 1e-6 * np.linalg.norm(xo)
