"""Global tractography's external energy on the GPU (include/mrgt.h) against the oracle
(oracle/gt_oracle.py): per-voxel energies, isotropic fractions and their sum, on a three-shell
scheme with two isotropic tissues; the closed-form case of a noise-free model signal; nf = 0.

Bar: energies and fractions within 1e-5 relative of the oracle's (float32 inputs and outputs, FP64
arithmetic), the total within 1e-9 relative of the sum of the oracle's energies (summation order
differs).  Up to three isotropic tissues the fused kernel enumerates the NNLS active sets; four
take the batched ICLS kernel; absent tissues in the data make active bounds in both."""
import sys

import numpy as np
import pytest

from mrtrix3_amd import gt, phantom
from oracle import gt_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def setup(seed, nf=2, lmax=8):
    """Three shells (b = 0, 1000, 3000) for up to two isotropic tissues; five (b = 0 to 4000) for
    more: the WM column and the isotropic columns of Ak are independent only with at least as many
    shells as unknowns."""
    rng = np.random.default_rng(seed)
    if nf <= 2:
        grad = phantom.gradient_table([(0, 4), (1000, 20), (3000, 40)], rng)
        wm = np.array([[1000.0, 0, 0, 0, 0], [600.0, -120.0, 30.0, -5.0, 0.5], [350.0, -150.0, 60.0, -15.0, 2.5]])
        iso = [[800.0, 350.0, 100.0], [1000.0, 200.0, 5.0]][:nf]
    else:
        grad = phantom.gradient_table([(0, 4), (1000, 20), (2000, 30), (3000, 40), (4000, 40)], rng)
        wm = np.array([[1000.0, 0, 0, 0, 0], [600.0, -120.0, 30.0, -5.0, 0.5], [470.0, -140.0, 45.0, -10.0, 1.5],
                       [350.0, -150.0, 60.0, -15.0, 2.5], [260.0, -140.0, 70.0, -20.0, 4.0]])
        iso = [[800.0, 350.0, 180.0, 100.0, 60.0], [1000.0, 200.0, 40.0, 5.0, 1.0], [600.0, 420.0, 330.0, 300.0, 250.0],
               [900.0, 100.0, 12.0, 1.0, 0.1]][:nf]
    return rng, grad, wm, iso, lmax


@pytest.mark.parametrize("nf", [2, 0, 3, 4])   # nf <= 3: the fused kernel; 4: residual + ICLS + energy kernels
def test_energy_matches_oracle(torch_cuda, nf):
    torch = torch_cuda
    rng, grad, wm, iso, lmax = setup(11, nf)
    E = gt.ExternalEnergy(grad, wm, iso, lmax=lmax, weight=0.8, ppot=0.05)
    K, Ak = E.matrix("K"), E.matrix("Ak")
    O = gt_oracle.Energy(K, Ak, mu=0.05 * np.sqrt(4 * np.pi))
    nvox = 3000
    tod = (rng.standard_normal((nvox, K.shape[1])) * 0.05).astype(np.float32)
    tod[:, 0] = np.abs(tod[:, 0]) + 0.1
    f = rng.uniform(0, 0.5, (nvox, nf + 1))
    f[:, 1:][rng.random((nvox, nf)) < 0.3] = 0.0     # absent tissues: active bounds of the NNLS
    dwi = (tod.astype(np.float64) @ K.T + f @ Ak.T + rng.standard_normal((nvox, K.shape[0])) * 20.0).astype(np.float32)
    say("[gpu gt] %d voxels, %d volumes, %d coefficients, nf %d" % (nvox, K.shape[0], K.shape[1], nf))
    e, fi, total = E.energy(torch.from_numpy(dwi).cuda(), torch.from_numpy(tod).cuda())
    torch.cuda.synchronize()
    e = e.cpu().numpy()
    total = float(total.cpu().numpy()[0])
    eo = np.empty(nvox)
    fo = np.empty((nvox, nf + 1))
    for v in range(nvox):
        eo[v], fo[v] = O.voxel(dwi[v].astype(np.float64), tod[v].astype(np.float64))
    rel = np.abs(e - eo) / np.abs(eo)
    say("[gpu gt] energy max rel %.2e, total %.10g vs %.10g" % (rel.max(), total, eo.sum()))
    assert rel.max() <= 1e-5
    assert total == pytest.approx(eo.astype(np.float32).astype(np.float64).sum(), rel=1e-6)
    assert total == pytest.approx(eo.sum(), rel=1e-9)
    if nf:
        fi = fi.cpu().numpy()
        scale = np.abs(fo[:, 1:]).max()
        assert np.abs(fi - fo[:, 1:]).max() <= 1e-5 * scale
        assert (fi >= -1e-9 * scale).all()   # active bounds hold up to rounding, as in the oracle
        active = fo[:, 1:] <= 1e-9 * scale
        say("[gpu gt] %d active bounds among %d fractions" % (int(active.sum()), active.size))
        assert active.sum() > 0 and np.abs(fi[active]).max() <= 1e-5 * scale
    else:
        assert fi is None


def test_model_signal_closed_form(torch_cuda):
    torch = torch_cuda
    rng, grad, wm, iso, lmax = setup(12)
    E = gt.ExternalEnergy(grad, wm, iso, lmax=lmax, weight=1.0, ppot=0.1)
    K, Ak = E.matrix("K"), E.matrix("Ak")
    nvox = 512
    tod = (rng.standard_normal((nvox, K.shape[1])) * 0.05).astype(np.float32)
    tod[:, 0] = np.abs(tod[:, 0]) + 0.1
    f = rng.uniform(0, 1.0, (nvox, 3))
    dwi = (tod.astype(np.float64) @ K.T + f @ Ak.T).astype(np.float32)
    e, fi, total = E.energy(dwi, tod)          # host-pointer entry
    # |Ak[:, 0]|^2 f0^2 / nrows + mu t0, up to the float32 rounding of the signal
    want = (Ak[:, 0] @ Ak[:, 0]) * f[:, 0] ** 2 / K.shape[0] + 0.1 * np.sqrt(4 * np.pi) * tod[:, 0]
    np.testing.assert_allclose(e, want, rtol=2e-4)
    np.testing.assert_allclose(fi, f[:, 1:], atol=1e-4)
    say("[gpu gt] closed form: max rel %.2e" % float(np.max(np.abs(e - want) / want)))


def test_empty_batch(torch_cuda):
    rng, grad, wm, iso, lmax = setup(13)
    E = gt.ExternalEnergy(grad, wm, iso, lmax=lmax)
    e, fi, total = E.energy(np.zeros((0, E.nrows), np.float32), np.zeros((0, E.ncols), np.float32))
    assert e.shape == (0,) and total == 0.0
