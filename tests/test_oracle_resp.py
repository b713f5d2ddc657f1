"""The amp2response oracle (oracle/resp_oracle.py) pinned by what the mathematics fixes: the
Legendre values against scipy, the elevation-derivative transform against a finite difference of
the amplitude transform, exact recovery of a known zonal response by both solves, and invariance
to the fibre's sign and to the random axes of the rotation (amp2response.cpp:72-90)."""
import math

import numpy as np
import pytest
from scipy import special

from oracle import resp_oracle as R


def test_zonal_legendre_against_scipy():
    for x in (-0.9, -0.3, 0.0, 0.41, 0.77, 1.0):
        p = R.plm_sph_column(12, 0, x)
        for l in range(0, 13):
            ref = math.sqrt((2 * l + 1) / (4 * math.pi)) * special.eval_legendre(l, x)
            assert abs(p[l] - ref) <= 1e-13


def test_derivative_transform_is_d_del_of_amplitudes():
    els = np.linspace(0.05, math.pi / 2 - 0.05, 17)
    h = 1e-6
    fd = (R.amp_transform(els + h, 10) - R.amp_transform(els - h, 10)) / (2 * h)
    assert np.allclose(R.deriv_transform(els, 10), fd, atol=1e-7)


def shell(n=64, seed=1):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def tensor_response(lmax, bval=3000.0, dpar=1.7e-3, dperp=0.3e-3):
    """Zonal coefficients of exp(-b D(el)), D = dperp + (dpar - dperp) cos^2 el, fitted on a fine
    grid: a smooth, non-negative response that increases with elevation."""
    els = np.linspace(0, math.pi / 2, 400)
    amp = np.exp(-bval * (dperp + (dpar - dperp) * np.cos(els) ** 2))
    T = R.amp_transform(els, lmax)
    return np.linalg.lstsq(T, amp, rcond=None)[0]


def synth(nvox, dirs, r, lmax, seed=2):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((nvox, 3))
    f /= np.linalg.norm(f, axis=1, keepdims=True)
    amp = np.zeros((nvox, len(dirs)))
    for v in range(nvox):
        c = np.abs(dirs @ f[v])
        amp[v] = R.amp_transform(np.arccos(np.clip(c, -1, 1)), lmax) @ r
    return amp, f


@pytest.mark.parametrize("constrained", [False, True])
def test_recovers_a_known_response(constrained):
    lmax = 10
    dirs = shell()
    r = tensor_response(lmax)
    assert np.all(R.constraints(lmax) @ r >= -1e-9)   # the constraints hold for it
    amp, f = synth(300, dirs, r, lmax)
    M, b, count = R.accumulate(amp, f, None, dirs, np.arange(len(dirs)), lmax)
    assert count == 300
    rf, _ = R.solve(lmax, M, b, constrained)
    assert np.allclose(rf, r, rtol=1e-6, atol=1e-9 * np.abs(r).max())


def test_fibre_sign_and_random_axes_do_not_matter():
    lmax = 8
    dirs = shell(40, 3)
    rng = np.random.default_rng(4)
    amp = rng.uniform(0.1, 1.0, (50, 40))
    f = rng.standard_normal((50, 3))
    M1, b1, _ = R.accumulate(amp, f, None, dirs, np.arange(40), lmax, seed=0)
    M2, b2, _ = R.accumulate(amp, -3.0 * f, None, dirs, np.arange(40), lmax, seed=7)
    assert np.allclose(M1, M2, rtol=1e-12) and np.allclose(b1, b2, rtol=1e-12)


def test_mask_and_volume_selection():
    lmax = 4
    dirs = shell(20, 5)
    rng = np.random.default_rng(6)
    amp = rng.uniform(0.1, 1.0, (30, 25))
    vols = np.arange(3, 23)
    f = rng.standard_normal((30, 3))
    mask = (np.arange(30) % 3 != 0).astype(np.uint8)
    M, b, count = R.accumulate(amp, f, mask, dirs, vols, lmax)
    Ms, bs, cs = R.accumulate(amp[mask == 1], f[mask == 1], None, dirs, vols, lmax)
    assert count == cs == 20 and np.allclose(M, Ms, rtol=1e-13) and np.allclose(b, bs, rtol=1e-13)


def test_zero_fibre_direction_gives_nan_but_at_lmax_0():
    """A zero fibre direction leaves the rotation zero (Eigen's normalize keeps zero vectors), so
    cartesian2spherical takes acos(0 / 0): every sum with an l > 0 term is NaN; the l = 0 value
    does not depend on the elevation."""
    dirs = shell(10, 8)
    amp = np.ones((1, 10))
    f = np.zeros((1, 3))
    M, b, _ = R.accumulate(amp, f, None, dirs, np.arange(10), 4)
    assert np.isnan(M[1:]).all() and np.isnan(b[1:]).all() and np.isfinite(M[0, 0]) and np.isfinite(b[0])
    M0, b0, _ = R.accumulate(amp, f, None, dirs, np.arange(10), 0)
    assert np.isfinite(M0).all() and abs(M0[0, 0] - 10 * R.Y00 ** 2) < 1e-15
