"""The per-voxel classes with the reference's call surface (host/csd.h CSD, host/msmt.h MSMT_CSD,
host/icls.h ICLS::Solver), driven by tests/native/voxel_classes.cpp in the shape of dwi2fod's
CSD_Processor / MSMT_Processor (cmd/dwi2fod.cpp:120-231): several threads, one solver copy each.

Bars: the CSD loop gives the batched solver's FODs and loop counters bit for bit (same kernel and
arithmetic); FOD() after c iterate() calls equals the batched solver built with niter = c (the
reference's F mid-way); MSMT_CSD, whose solve takes and returns double precision, equals the
single-precision batch path after rounding its result to float32, with identical niter, and the
FP64 oracle to 1e-6 (two FP64 active-set solvers whose sums differ in order); ICLS in double reproduces the reference's Matlab goldens at its own 1e-6."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from mrtrix3_amd import csd, icls, phantom
from oracle import orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "native", "_bin", "voxel_classes")
GOLD = np.load(os.path.join(HERE, "golden", "icls_golden.npz"))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def run_driver(*args):
    assert os.path.exists(EXE), "voxel_classes is built by __graft_entry__.build()"
    say("[voxel classes] %s" % " ".join(str(a) for a in args))
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr + r.stdout
    return r


def write_csd_input(path, grad, response, dw, lmax=0):
    with open(path, "wb") as f:
        np.array([grad.shape[0], dw.shape[0], np.size(response), lmax], dtype=np.int32).tofile(f)
        np.ascontiguousarray(grad, dtype=np.float64).tofile(f)
        np.ascontiguousarray(np.ravel(response), dtype=np.float64).tofile(f)
        np.ascontiguousarray(dw, dtype=np.float32).tofile(f)


def csd_sample():
    ph = phantom.single_shell_small()
    dw = np.ascontiguousarray(ph.dw[:3000]).copy()
    dw[5, 7] = np.nan          # skipped voxels (zero FOD)
    dw[11, :] = np.inf
    dw[20:40, 10:20] *= -1.0   # negative signals (taken as 0)
    return ph, dw


def test_csd_processor_loop_matches_batch(torch_cuda):
    ph, dw = csd_sample()
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        write_csd_input(fin, ph.grad, ph.response, dw)
        run_driver("csd", fin, fout, 4)
        raw = np.fromfile(fout, dtype=np.uint8)
    nsh = int(raw[:4].view(np.int32)[0])
    nvox = dw.shape[0]
    fod = raw[4:4 + 4 * nvox * nsh].view(np.float32).reshape(nvox, nsh)
    n = raw[4 + 4 * nvox * nsh:].view(np.int32)
    ref, it, _ = csd.CSD(csd.CSDShared(ph.grad, ph.response)).run(dw)
    it = it.astype(np.int64)
    it[it == 0xFFFF] = -1
    say("[voxel classes] csd: %d voxels, %d skipped, loop counters %d..%d" % (nvox, int((n < 0).sum()), n.min(), n.max()))
    assert np.array_equal(n, it)
    assert np.array_equal(fod, ref)
    assert (n[[5, 11]] == -1).all() and not fod[[5, 11]].any()


def test_fod_midway_is_the_reference_trajectory(torch_cuda):
    ph, dw = csd_sample()
    sub = np.ascontiguousarray(dw[100:124])
    cmax = 6
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        write_csd_input(fin, ph.grad, ph.response, sub)
        run_driver("trajectory", fin, fout, cmax)
        raw = np.fromfile(fout, dtype=np.uint8)
    nsh = int(raw[:4].view(np.int32)[0])
    traj = raw[4:].view(np.float32).reshape(sub.shape[0], cmax + 1, nsh)
    for c in range(cmax + 1):
        ref, _, _ = csd.CSD(csd.CSDShared(ph.grad, ph.response, niter=c)).run(sub)
        assert np.array_equal(traj[:, c], ref), c
    # the trajectory moves (F after 0, 1, 2 solves differ) and settles
    assert not np.array_equal(traj[:, 0], traj[:, 1]) and not np.array_equal(traj[:, 1], traj[:, 2])
    say("[voxel classes] trajectory: %d voxels x %d budgets identical to the batch solver" % (sub.shape[0], cmax + 1))


def msmt_sample(nvox=400):
    ph = phantom.multi_shell(seed=4, shape=(16, 16, 8), semi=None)
    dw = np.ascontiguousarray(ph.dw[:nvox], dtype=np.float32)
    return ph, dw


def test_msmt_processor_loop_matches_batch_and_oracle(torch_cuda):
    ph, dw = msmt_sample()
    resp = ph.response
    lmax = [8, 0, 0]
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            rows = [np.atleast_2d(r).shape[0] for r in resp]
            cols = [np.atleast_2d(r).shape[1] for r in resp]
            np.array([ph.grad.shape[0], dw.shape[0], len(resp)], dtype=np.int32).tofile(f)
            np.array(rows + cols + lmax, dtype=np.int32).tofile(f)
            np.ascontiguousarray(ph.grad, dtype=np.float64).tofile(f)
            for r in resp:
                np.ascontiguousarray(np.atleast_2d(r), dtype=np.float64).tofile(f)
            dw.tofile(f)
        run_driver("msmt", fin, fout, 4)
        raw = np.fromfile(fout, dtype=np.uint8)
    n = int(raw[:4].view(np.int32)[0])
    nvox = dw.shape[0]
    x = raw[4:4 + 8 * nvox * n].view(np.float64).reshape(nvox, n)
    niter = raw[4 + 8 * nvox * n:].view(np.int32)
    shared = icls.MSMTShared(ph.grad, resp, lmax=lmax)
    x32, it32, _ = shared.problem().solve(dw)
    assert np.array_equal(niter, it32.astype(np.int32))
    assert np.array_equal(x.astype(np.float32), x32)
    o = orc.MsmtOracle(ph.grad, resp, lmax=lmax)
    worst = 0.0
    for v in range(0, nvox, 7):
        xo, ito, _ = o.solve(dw[v].astype(np.float64))
        # with tol 0 a rounding-level constraint value can add one no-op outer iteration
        assert abs(int(niter[v]) - int(ito)) <= 1
        worst = max(worst, np.linalg.norm(x[v] - xo) / np.linalg.norm(xo))
    say("[voxel classes] msmt: %d voxels, niter %d..%d, worst relative error vs oracle %.2e" %
        (nvox, niter.min(), niter.max(), worst))
    assert worst <= 1e-6


def test_icls_double_precision_goldens(torch_cuda):
    """The reference's unit-test problems (testing_unit_tests_icls.cpp:46-131) with the double
    right-hand side: the Matlab quadprog solutions at the reference's isApprox(.., 1e-6)."""
    H, Ai, Ae = GOLD["H"], GOLD["A_ineq"], GOLD["A_eq"]
    ti, te = GOLD["t_ineq"], GOLD["t_eq"]
    A = np.vstack([Ai, Ae])
    t = np.concatenate([ti, te])
    cases = [(dict(H=H, A=A, t=t, num_eq=10), GOLD["x_eq"]), (dict(H=H, A=A, t=None, num_eq=10), GOLD["x_eq_zero"]),
             (dict(H=H, A=Ai, t=ti), GOLD["x_no_eq"]), (dict(H=H, A=Ai, t=None), GOLD["x_no_eq_zero"])]
    for kw, ref in cases:
        p = icls.IclsProblem(**kw)
        b = np.tile(GOLD["b"].astype(np.float64), (2, 1))
        x, it, na = p.solve(b)
        assert x.dtype == np.float64
        xo, ito, acto = orc.IclsOracle(**kw).solve(b[0])
        for r in range(2):
            err = np.linalg.norm(x[r] - ref) / np.linalg.norm(ref)
            assert err <= 1e-6, err
            assert np.linalg.norm(x[r] - xo) <= 1e-8 * np.linalg.norm(xo)
            assert abs(int(it[r]) - int(ito)) <= 1 and na[r] == acto.sum()
    say("[voxel classes] ICLS goldens in double: within 1e-6 of Matlab")
