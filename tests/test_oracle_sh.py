"""Pins the oracle's spherical-harmonic set-up (core/math/legendre.h, SH.h, ZSH.h,
least_squares.h) against closed forms and invariants, not against a retyped formula."""
import math

import numpy as np
import pytest

from oracle import orc


def dirs_from_xyz(v):
    v = v / np.linalg.norm(v, axis=1, keepdims=True)
    return np.stack([np.arctan2(v[:, 1], v[:, 0]), np.arccos(np.clip(v[:, 2], -1, 1))], axis=1)


def test_y00_and_degree2_closed_forms():
    rng = np.random.default_rng(0)
    d = dirs_from_xyz(rng.standard_normal((50, 3)))
    Y = orc.sh_transform(d, 2)
    az, el = d[:, 0], d[:, 1]
    st, ct = np.sin(el), np.cos(el)
    assert np.allclose(Y[:, 0], 1.0 / (2.0 * math.sqrt(math.pi)), atol=1e-15)
    # index(l,m) = l(l+1)/2 + m: degree 2 occupies columns 1..5 for m = -2..2
    assert np.allclose(Y[:, 3], math.sqrt(5 / (16 * math.pi)) * (3 * ct ** 2 - 1), atol=1e-14)
    # real basis with Condon-Shortley phase and sqrt(2): m=+1 -> cos, m=-1 -> sin
    c1 = -math.sqrt(15 / (4 * math.pi)) * st * ct
    assert np.allclose(Y[:, 4], c1 * np.cos(az), atol=1e-14)
    assert np.allclose(Y[:, 2], c1 * np.sin(az), atol=1e-14)
    c2 = math.sqrt(15 / (16 * math.pi)) * st ** 2
    assert np.allclose(Y[:, 5], c2 * np.cos(2 * az), atol=1e-14)
    assert np.allclose(Y[:, 1], c2 * np.sin(2 * az), atol=1e-14)


@pytest.mark.parametrize("lmax", [2, 4, 8, 12])
def test_quadrature_orthonormality(lmax):
    # Gauss-Legendre in cos(theta) x uniform azimuth integrates degree <= 2*lmax exactly
    nz, nphi = lmax + 2, 2 * lmax + 3
    x, w = np.polynomial.legendre.leggauss(nz)
    phi = 2 * math.pi * np.arange(nphi) / nphi
    el = np.repeat(np.arccos(x), nphi)
    az = np.tile(phi, nz)
    wt = np.repeat(w, nphi) * (2 * math.pi / nphi)
    Y = orc.sh_transform(np.stack([az, el], axis=1), lmax)
    G = (Y * wt[:, None]).T @ Y
    assert np.allclose(G, np.eye(Y.shape[1]), atol=1e-12)


@pytest.mark.parametrize("lmax", [4, 8])
def test_addition_theorem(lmax):
    rng = np.random.default_rng(lmax)
    u = rng.standard_normal((20, 3))
    v = rng.standard_normal((20, 3))
    Yu = orc.sh_transform(dirs_from_xyz(u), lmax)
    Yv = orc.sh_transform(dirs_from_xyz(v), lmax)
    cosg = np.sum(u / np.linalg.norm(u, axis=1)[:, None] * v / np.linalg.norm(v, axis=1)[:, None], axis=1)
    for l in range(0, lmax + 1, 2):
        cols = slice(l * (l - 1) // 2, l * (l + 1) // 2 + l + 1)
        lhs = np.sum(Yu[:, cols] * Yv[:, cols], axis=1)
        rhs = (2 * l + 1) / (4 * math.pi) * np.polynomial.legendre.legval(cosg, [0] * l + [1])
        assert np.allclose(lhs, rhs, atol=1e-12)


def test_antipodal_symmetry():
    rng = np.random.default_rng(3)
    u = rng.standard_normal((30, 3))
    assert np.allclose(orc.sh_transform(dirs_from_xyz(u), 8), orc.sh_transform(dirs_from_xyz(-u), 8),
                       atol=1e-12)


def test_poles_give_zonal_only():
    Y = orc.sh_transform(np.array([[0.3, 0.0], [1.1, math.pi]]), 8)
    for l in range(0, 9, 2):
        for m in range(-l, l + 1):
            c = l * (l + 1) // 2 + m
            if m == 0:
                assert abs(Y[0, c] - math.sqrt((2 * l + 1) / (4 * math.pi))) < 1e-13
            else:
                assert Y[0, c] == 0.0 and Y[1, c] == 0.0


def test_zsh2rh_closed_form():
    z = np.array([3.0, -2.0, 1.5, 0.25, -0.125])
    rh = orc.zsh2rh(z)
    want = z * np.sqrt(4 * math.pi / (2 * np.arange(0, 9, 2) + 1))
    assert np.allclose(rh, want, rtol=1e-14)


def test_lforn_matches_single_precision_formula():
    # SH.h:70-73 evaluates sqrt(1+8N) in float; spot values and the inverse of NforL
    for L in range(0, 21, 2):
        n = (L + 1) * (L + 2) // 2
        assert orc.lforn(n) == L
        assert orc.lforn(n + 1) == L
    assert orc.lforn(60) == 8 and orc.lforn(64) == 8 and orc.lforn(45) == 8 and orc.lforn(44) == 6


@pytest.mark.parametrize("shape", [(60, 45), (45, 91), (10, 4)])
def test_pinv_moore_penrose(shape):
    rng = np.random.default_rng(sum(shape))
    M = rng.standard_normal(shape)
    P = orc.pinv(M)
    assert np.allclose(M @ P @ M, M, atol=1e-9)
    assert np.allclose(P @ M @ P, P, atol=1e-9)
    assert np.allclose(P, np.linalg.pinv(M), atol=1e-9)
