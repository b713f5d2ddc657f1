"""The learned warm start on the whole-brain workload: 646,128 voxels of the single-shell phantom
(8 b=0 + 64 b=3000 volumes, lmax 8, 300 constraint directions) solved from the proposals of the
network trained for that acquisition protocol (64 inputs, 256-256-256 hidden, 45 coefficients +
300 set logits) on 2.58M plain-solver voxels of four further phantoms of the protocol (seeds
100-103, held out from the evaluated one).  The network is tests/golden/warm_net_single_shell_brain.npz,
written by `python bench.py --save-net tests/golden/warm_net_single_shell_brain.npz` (bench.py's
default training: 64 epochs, batch 8192, peak learning rate 8e-3, set-loss weight 0.3).

Bar: identical final negative sets and <= 1e-4 relative SH error against the plain solver in every
voxel; the plain and warm FODs against the oracle on a 2,048-voxel sample; at most 2.7 solves per
voxel (the plain solver takes 4.6)."""
import os
import sys

import numpy as np
import pytest

from mrtrix3_amd import csd, phantom, warm
from oracle import orc

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


def test_warm_start_on_the_whole_brain_workload(torch_cuda):
    torch = torch_cuda
    say("[c5] generating the 646,128-voxel phantom")
    ph = phantom.single_shell_brain()
    shared = csd.CSDShared(ph.grad, ph.response)
    solver = csd.CSD(shared)
    net = warm.WarmNet.load(NET)
    assert (net.n_in, net.hidden, net.nsh, net.nlog) == (64, 256, 45, 300) == (shared.n_dw, 256, shared.n_sh,
                                                                               shared.n_hr)
    dev = torch.device("cuda:0")
    dwis = torch.from_numpy(shared.dwis.astype(np.int32)).to(dev)
    b0 = torch.from_numpy(np.nonzero(shared.matrix("grad")[:, 3] <= 10.0)[0].astype(np.int32)).to(dev)
    dw = torch.from_numpy(ph.dw).to(dev)
    st_p = torch.zeros(6, dtype=torch.int64, device=dev)
    st_w = torch.zeros(6, dtype=torch.int64, device=dev)
    say("[c5] plain and warm-started solves of %d voxels" % ph.nvox)
    f_p, _, _, nm_p = solver.run(dw, stats=st_p, negmask=True)
    f_w, _, _, nm_w = warm.warm_solve(net, solver, dw, dwis, b0, budget=6, stats=st_w, negmask=True)
    torch.cuda.synchronize()
    a = f_p.double()
    rel = ((f_w.double() - a).norm(dim=1) / a.norm(dim=1).clamp_min(1e-30)).cpu().numpy()
    same = (nm_p == nm_w).all(dim=1).cpu().numpy()
    sp, sw = st_p.cpu().numpy(), st_w.cpu().numpy()
    say("[c5] max rel %.3e, identical sets %.6f, solves/voxel plain %.3f warm %.3f, fallback %.5f"
        % (rel.max(), same.mean(), sp[1] / sp[0], sw[1] / sw[0], sw[4] / sw[0]))
    assert same.all() and rel.max() <= 1e-4
    assert sw[1] / sw[0] <= 2.7 < sp[1] / sp[0]
    # both against the oracle on a sample spread over the volume
    idx = np.sort(np.random.default_rng(5).choice(ph.nvox, 2048, replace=False))
    o = orc.CsdOracle(ph.grad, ph.response)
    ref, _ = o.run_volume(np.ascontiguousarray(ph.dw[idx]), (64, 32, 1), nthreads=8)
    ref = ref.astype(np.float64)
    for name, f in (("plain", f_p), ("warm", f_w)):
        g = f.cpu().numpy()[idx].astype(np.float64)
        r = np.linalg.norm(g - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
        say("[c5] %s vs oracle on %d voxels: max rel %.3e" % (name, len(idx), r.max()))
        assert r.max() <= 1e-4
