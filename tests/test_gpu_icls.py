"""GPU ICLS / MSMT-CSD (libmrcsd.so through include/mricls.h) against the reference's golden
vectors and the CPU oracle."""
import os
import sys

import numpy as np
import pytest

from mrtrix3_amd import icls, phantom
from oracle import orc

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "icls_golden.npz"))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def is_approx(x, y, prec):
    return np.linalg.norm(x - y) <= prec * min(np.linalg.norm(x), np.linalg.norm(y))


def golden_cases():
    H, Ai, Ae = GOLD["H"], GOLD["A_ineq"], GOLD["A_eq"]
    ti, te = GOLD["t_ineq"], GOLD["t_eq"]
    A = np.vstack([Ai, Ae])
    t = np.concatenate([ti, te])
    return [
        ("test1_eq", dict(H=H, A=A, t=t, num_eq=10), GOLD["x_eq"]),
        ("test2_eq_zero_t", dict(H=H, A=A, t=None, num_eq=10), GOLD["x_eq_zero"]),
        ("test3_no_eq", dict(H=H, A=Ai, t=ti), GOLD["x_no_eq"]),
        ("test4_no_eq_zero_t", dict(H=H, A=Ai, t=None), GOLD["x_no_eq_zero"]),
    ]


@pytest.mark.parametrize("name,kw,ref", golden_cases(), ids=[c[0] for c in golden_cases()])
def test_icls_golden_on_gpu(torch_cuda, name, kw, ref):
    say("[gpu icls] golden %s" % name)
    p = icls.IclsProblem(**kw)
    b = np.tile(GOLD["b"].astype(np.float32), (3, 1))     # a batch of three identical right-hand sides
    x, it, na = p.solve(b)
    o = orc.IclsOracle(**kw)
    xo, ito, acto = o.solve(b[0].astype(np.float64))       # the oracle on the same float32 input
    for r in range(3):
        # with tol = 0 the reference takes one extra no-op outer iteration whenever rounding leaves
        # an active constraint at c = -1e-17 instead of +1e-17, so niter may differ by one
        assert abs(int(it[r]) - int(ito)) <= 1 and na[r] == acto.sum()
        # the float32 measurement rounding moves the answer by ~1e-7; the reference bar is 1e-6
        assert is_approx(x[r].astype(np.float64), ref, 2e-6), np.linalg.norm(x[r] - ref) / np.linalg.norm(ref)
        assert is_approx(x[r].astype(np.float64), xo, 1e-6)


def test_icls_random_vs_oracle(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    nmeas, n, m = 60, 20, 120
    H = rng.standard_normal((nmeas, n))
    A = rng.standard_normal((m, n))
    t = A @ rng.standard_normal(n) - rng.uniform(0.1, 1.0, m)
    p = icls.IclsProblem(H, A, t, con_reg=1e-10)
    o = orc.IclsOracle(H, A, t, con_reg=1e-10)
    B = (rng.standard_normal((500, nmeas)) * 3).astype(np.float32)
    say("[gpu icls] random problem, 500 right-hand sides")
    x, it, na = p.solve(torch.from_numpy(B).cuda())
    torch.cuda.synchronize()
    x = x.cpu().numpy()
    it = it.cpu().numpy().astype(np.int64)
    na = na.cpu().numpy().astype(np.int64)
    bad = 0
    for v in range(500):
        xo, ito, acto = o.solve(B[v].astype(np.float64))
        if not (np.linalg.norm(x[v] - xo) <= 1e-5 * max(np.linalg.norm(xo), 1e-12)
                and abs(it[v] - ito) <= 1 and na[v] == acto.sum()):
            bad += 1
    say("[gpu icls] mismatches %d / 500" % bad)
    assert bad <= 1


def test_msmt_small_volume(torch_cuda):
    torch = torch_cuda
    ph = phantom.multi_shell(seed=4, shape=(16, 16, 8), semi=None)
    s = icls.MSMTShared(ph.grad, ph.response)
    p = s.problem()
    say("[gpu msmt] %d voxels, %d params, %d constraints" % (ph.nvox, s.nparams, s.ncon))
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    x, it, na = p.solve(torch.from_numpy(ph.dw).cuda(), stats=stats)
    torch.cuda.synchronize()
    x = x.cpu().numpy().astype(np.float64)
    it = it.cpu().numpy().astype(np.int64)
    say("[gpu msmt] stats", stats.cpu().numpy().tolist())
    o = orc.MsmtOracle(ph.grad, ph.response)
    rel, same_it = [], []
    for v in range(0, ph.nvox, 3):
        xo, ito, _ = o.solve(ph.dw[v].astype(np.float64))
        rel.append(np.linalg.norm(x[v] - xo) / np.linalg.norm(xo))
        same_it.append(it[v] == ito)
    rel = np.array(rel)
    say("[gpu msmt] max_rel=%.3e p99.9=%.3e same_iters=%.4f" % (rel.max(), np.quantile(rel, 0.999),
                                                                 np.mean(same_it)))
    assert np.mean(rel <= 1e-4) >= 0.999
    assert np.all(it < 0xFFFE)
    wm, gm, csf = s.split(x)
    assert wm.shape[1] == 45 and gm.shape[1] == 1 and csf.shape[1] == 1


def test_msmt_masked_voxels(torch_cuda):
    torch = torch_cuda
    ph = phantom.multi_shell(seed=6, shape=(4, 4, 4), semi=None)
    s = icls.MSMTShared(ph.grad, ph.response)
    p = s.problem()
    valid = np.ones(ph.nvox, dtype=np.uint8)
    valid[::5] = 0
    x, it, na = p.solve(torch.from_numpy(ph.dw).cuda(), valid=torch.from_numpy(valid).cuda())
    torch.cuda.synchronize()
    x = x.cpu().numpy()
    it = it.cpu().numpy().view(np.uint16)
    assert np.all(x[::5] == 0) and np.all(it[::5] == 0xFFFF)
    assert np.all(it[1::5] < 0xFFFE)


def test_capacity_passes_give_identical_answers(torch_cuda):
    """The multi-shell volume with a 16-slot first pass: every voxel outgrows it and is solved again
    with 64 slots.  The arithmetic does not depend on the capacity, so the answers are the default
    pass's bit for bit."""
    torch = torch_cuda
    ph = phantom.multi_shell(seed=4, shape=(16, 16, 8), semi=None)
    p = icls.MSMTShared(ph.grad, ph.response).problem()
    assert p.capacity == 48
    b = torch.from_numpy(ph.dw).cuda()
    valid = torch.from_numpy((np.arange(ph.nvox) % 7 != 3).astype(np.uint8)).cuda()
    x0, it0, na0 = p.solve(b, valid=valid)
    p.capacity = 16
    st = torch.zeros(8, dtype=torch.int64, device="cuda")
    x1, it1, na1 = p.solve(b, valid=valid, stats=st)
    torch.cuda.synchronize()
    say("[gpu icls] capacity 16 -> 64: stats", st.cpu().numpy().tolist())
    assert torch.equal(x0, x1) and torch.equal(it0, it1) and torch.equal(na0, na1)
    assert int(st[0]) == int(valid.sum()) and int(st[7]) == 0
    assert int(na1.max()) > 16
    p.capacity = 0
    assert p.capacity == 48


def test_large_active_sets_against_oracle(torch_cuda):
    """Non-negative least squares with 100 unknowns whose answers keep about 80 of the 100
    constraints active: the 16-slot first pass hands every voxel to the 64-slot pass and most on to
    the 128-slot pass.  Against the oracle, which has no capacity."""
    torch = torch_cuda
    rng = np.random.default_rng(21)
    n, m, nvox = 100, 160, 96
    H = rng.standard_normal((m, n))
    A = np.eye(n)
    xt = rng.standard_normal((nvox, n)) - 0.9            # mostly negative: most bounds end active
    B = (xt @ H.T + 0.01 * rng.standard_normal((nvox, m))).astype(np.float32)
    p = icls.IclsProblem(H, A, con_reg=1e-10)
    assert p.capacity == 128
    p.capacity = 16
    x, it, na = p.solve(torch.from_numpy(B).cuda())
    torch.cuda.synchronize()
    x, it, na = x.cpu().numpy(), it.cpu().numpy().view(np.uint16), na.cpu().numpy()
    o = orc.IclsOracle(H, A, None, con_reg=1e-10)
    bad = 0
    for v in range(nvox):
        xo, ito, acto = o.solve(B[v].astype(np.float64))
        if not (np.linalg.norm(x[v] - xo) <= 1e-5 * np.linalg.norm(xo) and it[v] == ito and na[v] == acto.sum()):
            bad += 1
    say("[gpu icls] 100 unknowns: active %d..%d, mismatches %d / %d" % (na.min(), na.max(), bad, nvox))
    assert na.max() > 64 and bad == 0
