"""The external-energy oracle (oracle/gt_oracle.py) pinned without the reference: the kernel matrix
against the SH transform of the volume directions (another route to the same basis values), a
closed-form energy for a noise-free signal of the model, and its NNLS (the oracle's ICLS on
Problem(Ak, I)) against scipy's NNLS (Lawson-Hanson)."""
import numpy as np
import pytest
from scipy.optimize import nnls

from mrtrix3_amd import phantom
from oracle import gt_oracle, orc


def scheme(seed=3):
    rng = np.random.default_rng(seed)
    return phantom.gradient_table([(0, 4), (1000, 20), (3000, 40)], rng)


def responses():
    wm = np.array([[1000.0, 0, 0, 0, 0], [600.0, -120.0, 30.0, -5.0, 0.5], [350.0, -150.0, 60.0, -15.0, 2.5]])
    return wm, [[800.0, 350.0, 100.0], [1000.0, 200.0, 5.0]]


def test_kernel_matrix_is_the_convolved_sh_transform():
    grad = scheme()
    wm, iso = responses()
    K, Ak = gt_oracle.kernels(grad, wm, iso, 8, 0.7)
    g, _ = orc.normalise_grad(grad)
    shell_of, nshell = orc.msmt_shells(g)
    assert nshell == 3
    for r in range(g.shape[0]):
        u = g[r, :3]
        if np.linalg.norm(u) > 0:
            az, el = np.arctan2(u[1], u[0]), np.arccos(np.clip(u[2] / np.linalg.norm(u), -1, 1))
        else:
            az, el = 0.0, np.pi / 2        # SH::delta of the zero vector: cp = 1, sp = 0, z = 0
        Y = orc.sh_transform(np.array([[az, el]]), 8)[0]
        rh = orc.zsh2rh(np.concatenate([wm[shell_of[r]], np.zeros(0)]))
        want = np.concatenate([Y[l * (l + 1) // 2 - l:l * (l + 1) // 2 + l + 1] * rh[l // 2] for l in range(0, 9, 2)])
        np.testing.assert_allclose(K[r], 0.7 * want, rtol=1e-12, atol=1e-12)
        assert Ak[r, 0] == pytest.approx(wm[shell_of[r], 0] / np.sqrt(4 * np.pi), rel=1e-15)
        assert Ak[r, 1] == iso[0][shell_of[r]] and Ak[r, 2] == iso[1][shell_of[r]]


def test_closed_form_energy_of_a_model_signal():
    """dwi = K t + Ak f with f >= 0: the fit recovers f exactly, the residual left after removing the
    isotropic part is Ak[:, 0] f0, so e = |Ak[:, 0]|^2 f0^2 / nrows + mu t0 (the reference subtracts
    only the isotropic columns, :250)."""
    grad = scheme()
    wm, iso = responses()
    K, Ak = gt_oracle.kernels(grad, wm, iso, 8, 1.0)
    E = gt_oracle.Energy(K, Ak, mu=0.37)
    rng = np.random.default_rng(4)
    for _ in range(20):
        t = rng.standard_normal(K.shape[1]) * 0.1
        t[0] = abs(t[0]) + 0.2
        f = rng.uniform(0.0, 1.0, 3)
        f[rng.integers(0, 3)] = 0.0            # an active constraint
        e, fk = E.voxel(K @ t + Ak @ f, t)
        np.testing.assert_allclose(fk, f, atol=1e-9)
        want = (Ak[:, 0] @ Ak[:, 0]) * f[0] ** 2 / K.shape[0] + 0.37 * t[0]
        assert e == pytest.approx(want, rel=1e-9, abs=1e-12)


def test_nnls_matches_lawson_hanson():
    grad = scheme()
    wm, iso = responses()
    K, Ak = gt_oracle.kernels(grad, wm, iso, 8, 1.0)
    E = gt_oracle.Energy(K, Ak, mu=0.0)
    rng = np.random.default_rng(5)
    for _ in range(50):
        y = rng.standard_normal(Ak.shape[0]) * 300.0
        f, _, _ = E.nnls.solve(y)
        ref, _ = nnls(Ak, y)
        np.testing.assert_allclose(f, ref, rtol=1e-8, atol=1e-9 * np.abs(ref).max() + 1e-12)


def test_shell_count_errors():
    grad = scheme()
    wm, iso = responses()
    with pytest.raises(ValueError, match="WM kernel size"):
        gt_oracle.kernels(grad, wm[:2], iso, 8, 1.0)
    with pytest.raises(ValueError, match="Isotropic kernel size"):
        gt_oracle.kernels(grad, wm, [[1.0, 2.0]], 8, 1.0)


def test_library_host_setup_matches_oracle():
    from mrtrix3_amd import gt
    grad = scheme(7)
    wm, iso = responses()
    for lmax, weight in ((8, 1.0), (4, 0.3), (10, 2.0)):
        K, Ak = gt.kernels(grad, wm, iso, lmax, weight)
        Ko, Ako = gt_oracle.kernels(grad, wm, iso, lmax, weight)
        np.testing.assert_allclose(K, Ko, rtol=1e-12, atol=1e-12)
        np.testing.assert_array_equal(Ak, Ako)
    with pytest.raises(Exception, match="WM kernel size"):
        gt.kernels(grad, wm[:2], iso, 8, 1.0)
