"""amp2response on the GPU (include/mrresp.h) against the float64 oracle (oracle/resp_oracle.py):
the accumulated normal equations (cmd/amp2response.cpp:152-191) to 1e-12 and the per-shell
responses (:336-361) to 1e-7 (the constrained solve takes b and returns the response in double, as
Problem<default_type> does; its active constraints -- amplitudes at elevations a degree apart --
are nearly dependent, so two FP64 active-set solvers that round differently agree to ~2e-8)."""
import sys

import numpy as np
import pytest

from mrtrix3_amd import phantom, resp
from oracle import resp_oracle as R

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def shell(n, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("lmax", [0, 4, 10, 12])
def test_accumulate_matches_oracle(torch_cuda, lmax):
    torch = torch_cuda
    rng = np.random.default_rng(lmax)
    nvox, ld = 700, 70
    dirs = shell(60, 1)
    vols = np.arange(5, 65)
    amp = rng.uniform(0.05, 1.0, (nvox, ld)).astype(np.float32)
    fib = rng.standard_normal((nvox, 4)).astype(np.float32)
    fib[3, :3] = 0.0                     # a zero fibre direction
    mask = (rng.random(nvox) < 0.8).astype(np.uint8)
    mask[3] = 0
    say("[gpu resp] accumulate lmax %d, %d voxels" % (lmax, nvox))
    acc = resp.Amp2Response(dirs, vols, lmax)
    d_amp, d_fib = torch.from_numpy(amp).cuda(), torch.from_numpy(fib).cuda()
    # with the zero direction in the mask the elevation is acos(0 / 0): NaN in every sum with an
    # l > 0 term
    mz = mask.copy()
    mz[3] = 1
    Mz, bz, _ = acc.accumulate(d_amp, d_fib, torch.from_numpy(mz).cuda())
    Mzo, bzo, _ = R.accumulate(amp[:10], fib[:10], mz[:10], dirs, vols, lmax)
    assert np.array_equal(np.isnan(Mz), np.isnan(Mzo)) and np.array_equal(np.isnan(bz), np.isnan(bzo))
    assert np.isfinite(Mz[0, 0]) and (lmax == 0 or np.isnan(Mz[1:]).all())
    M, b, count = acc.accumulate(d_amp, d_fib, torch.from_numpy(mask).cuda())
    Mo, bo, co = R.accumulate(amp, fib, mask, dirs, vols, lmax)
    say("[gpu resp] lmax %d: M rel %.2e, b rel %.2e" % (lmax, rel(M, Mo), rel(b, bo)))
    assert count == co == int(mask.sum())
    assert rel(M, Mo) <= 1e-12 and rel(b, bo) <= 1e-12
    M2, b2, _ = acc.accumulate(d_amp, d_fib, torch.from_numpy(mask).cuda())
    assert np.array_equal(M, M2) and np.array_equal(b, b2)   # fixed summation order


@pytest.mark.parametrize("lmax,constrained", [(0, True), (6, False), (10, True), (10, False)])
def test_solve_matches_oracle(torch_cuda, lmax, constrained):
    rng = np.random.default_rng(100 + lmax)
    dirs = shell(64, 2)
    # a noisy tensor-like signal around 30 random fibres: the least-squares response violates
    # the constraints, so the solver adds and drops some (11 iterations at lmax 10)
    f = rng.standard_normal((30, 3))
    f /= np.linalg.norm(f, axis=1, keepdims=True)
    c2 = (f @ dirs.T) ** 2
    amp = np.exp(-3000 * (0.3e-3 + 1.4e-3 * c2)) + 0.3 * rng.standard_normal(c2.shape)
    M, b, _ = R.accumulate(amp, f, None, dirs, np.arange(64), lmax)
    say("[gpu resp] solve lmax %d constrained %s" % (lmax, constrained))
    rf, it = resp.solve(lmax, M, b, constrained)
    ro, ito = R.solve(lmax, M, b, constrained)
    say("[gpu resp] rf rel %.2e, iterations %d (oracle %d)" % (rel(rf, ro), it, ito))
    assert rel(rf, ro) <= 1e-7
    assert it == ito
    if constrained and lmax:
        assert it > 0


def test_brain_single_fibre_voxels(torch_cuda):
    """The phantom's b=3000 shell over its 646,128 voxels with fibre directions from the GPU CSD
    and sh2peaks (the first peak), accumulated at lmax 10; a 3,000-voxel sample against the oracle
    and the whole volume's response against the oracle's solve of the GPU's sums."""
    torch = torch_cuda
    from mrtrix3_amd import csd, peaks
    ph = phantom.single_shell_brain()
    sh = csd.CSDShared(ph.grad, ph.response)
    fod, _, _ = csd.CSD(sh).run(torch.from_numpy(ph.dw).cuda(), iters=False, nneg=False)
    pk = peaks.SH2Peaks(sh.n_sh).run(fod)
    fib = pk[:, :3].contiguous()
    ok = torch.isfinite(fib).all(dim=1).to(torch.uint8)
    vols = np.nonzero(ph.grad[:, 3] > 100)[0]
    dirs = ph.grad[vols, :3] / np.linalg.norm(ph.grad[vols, :3], axis=1, keepdims=True)
    dw = torch.from_numpy(ph.dw).cuda()
    say("[gpu resp] brain: %d voxels, %d directions" % (ph.nvox, len(vols)))
    acc = resp.Amp2Response(dirs, vols, 10)
    M, b, count = acc.accumulate(dw, fib, ok)
    assert count == int(ok.sum().item())
    idx = np.random.default_rng(3).choice(ph.nvox, 3000, replace=False)
    sub_mask = np.zeros(ph.nvox, dtype=np.uint8)
    sub_mask[idx] = ok.cpu().numpy()[idx]
    Ms, bs, cs = acc.accumulate(dw, fib, torch.from_numpy(sub_mask).cuda())
    Mo, bo, co = R.accumulate(ph.dw[idx], fib.cpu().numpy()[idx], sub_mask[idx], dirs, vols, 10)
    say("[gpu resp] brain sample: M rel %.2e b rel %.2e" % (rel(Ms, Mo), rel(bs, bo)))
    assert cs == co and rel(Ms, Mo) <= 1e-12 and rel(bs, bo) <= 1e-12
    rf, it = resp.solve(10, M, b, True)
    ro, ito = R.solve(10, M, b, True)
    say("[gpu resp] brain response", np.round(rf, 3).tolist(), "rel %.2e" % rel(rf, ro))
    assert rel(rf, ro) <= 1e-7 and it == ito
