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


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_icls_kernel_under_wavefront_emulation(tmp_path):
    # the kernel file without its launcher (the harness launches the emulated workgroups)
    src = open(os.path.join(CSRC, "kernels", "icls_solve.hip")).read()
    cut = src.index("__global__ void reset_counter_kernel")
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
    r = subprocess.run([exe, str(d)], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "EMU:" not in r.stderr, r.stderr[-3000:]

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
