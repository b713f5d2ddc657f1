"""Launch-shape autotuners and captured solve steps (include/mrtune.h) on the GPU.

Bars: a tuned launch shape changes no result (the persistent grid and the ICLS capacity passes
run the same per-voxel arithmetic: identical FODs, solutions and iteration counts); a captured
step replayed several times, with the program's own eager work and synchronisations between the
replays, gives the eager step's output bit for bit on every replay (plain and warm start)."""
import os
import sys

import numpy as np
import pytest

from mrtrix3_amd import csd, icls, phantom, tune, warm

pytestmark = pytest.mark.gpu

NET = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "warm_net_single_shell_brain.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def brain(torch_cuda):
    torch = torch_cuda
    ph = phantom.single_shell_brain()
    shared = csd.CSDShared(ph.grad, ph.response)
    dev = torch.device("cuda:0")
    dw = torch.from_numpy(np.ascontiguousarray(ph.dw[:65536])).to(dev)
    dw2 = torch.from_numpy(np.ascontiguousarray(ph.dw[65536:131072])).to(dev)
    dwis = torch.from_numpy(shared.dwis.astype(np.int32)).to(dev)
    b0 = torch.from_numpy(np.nonzero(shared.matrix("grad")[:, 3] <= 10.0)[0].astype(np.int32)).to(dev)
    return shared, dw, dw2, dwis, b0


def test_csd_autotune_keeps_results(torch_cuda, brain):
    torch = torch_cuda
    shared, dw, _, dwis, b0 = brain
    solver = csd.CSD(shared)
    net = warm.WarmNet.load(NET)
    f0, it0, _ = solver.run(dw)
    _, neg = net.infer(dw, dwis, b0)
    fw0, itw0, _ = solver.run(dw, warm_neg=neg, warm_budget=6)
    rp = tune.autotune_csd(solver, dw)
    rw = tune.autotune_csd(solver, dw, warm_neg=neg)
    say("[tune] plain grid: %s" % rp)
    say("[tune] warm grid: %s" % rw)
    assert len(rp["candidates"]) == 5 and len(rw["candidates"]) == 5
    assert rp["chosen"] in [c["value"] for c in rp["candidates"]] and rp["candidates"][0]["value"] == rp["full"]
    f1, it1, _ = solver.run(dw)
    fw1, itw1, _ = solver.run(dw, warm_neg=neg, warm_budget=6)
    torch.cuda.synchronize()
    assert torch.equal(f0, f1) and torch.equal(it0, it1)
    assert torch.equal(fw0, fw1) and torch.equal(itw0, itw1)


def test_icls_autotune_keeps_results(torch_cuda):
    torch = torch_cuda
    ph = phantom.multi_shell(seed=4, shape=(32, 32, 16), semi=None)
    p = icls.MSMTShared(ph.grad, ph.response).problem()
    b = torch.from_numpy(ph.dw).cuda()
    x0, it0, na0 = p.solve(b)
    rep = tune.autotune_icls(p, b)
    say("[tune] icls capacity: %s" % rep)
    assert p.capacity == rep["chosen"] and rep["chosen"] >= 48
    x1, it1, na1 = p.solve(b)
    torch.cuda.synchronize()
    assert torch.equal(x0, x1) and torch.equal(it0, it1) and torch.equal(na0, na1)


@pytest.mark.parametrize("with_net", [False, True], ids=["plain", "warm"])
def test_graph_replays_match_eager(torch_cuda, brain, with_net):
    torch = torch_cuda
    shared, dw, dw2, dwis, b0 = brain
    solver = csd.CSD(shared)
    net = warm.WarmNet.load(NET) if with_net else None
    if with_net:
        _, neg = net.infer(dw, dwis, b0)
        ref, _, _ = solver.run(dw, warm_neg=neg, warm_budget=6)
    else:
        ref, _, _ = solver.run(dw)
    torch.cuda.synchronize()
    fod = torch.full_like(ref, float("nan"))
    step = tune.SolveGraph(solver, dw, fod, net=net, dwis=dwis, b0=b0)
    assert torch.equal(fod, ref)          # the eager step inside graph creation
    other = torch.empty_like(ref)
    for r in range(3):
        fod.fill_(float("nan"))
        # eager work of the program between replays: another solve, torch arithmetic, a sync
        solver.run(dw2, fod=other)
        junk = (dw2 * 2.0).sum()
        torch.cuda.synchronize()
        step.launch()
        torch.cuda.synchronize()
        assert torch.equal(fod, ref), "replay %d" % r
        assert torch.isfinite(junk)
    say("[tune] %s graph: 3 replays identical to the eager step" % ("warm" if with_net else "plain"))
