"""Test infrastructure only (imported by tests/, never by mrtrix3_amd/): a float64 numpy
restatement of amp2response (cmd/amp2response.cpp), the checker for include/mrresp.h.

accumulate() follows Accumulator::operator() (amp2response.cpp:152-191) step for step, the
rotation included: R's third row is the unit fibre direction and its first two rows are built from
a random normal vector (gen_rotation_matrix, :72-90), the rotated directions go to azimuth /
elevation by Math::Sphere::cartesian2spherical (core/math/sphere.h:83-90), elevations above pi/2
are folded (:170-179), and the zonal amplitude transform is ZSH::init_amp_transform
(core/math/ZSH.h:66-77) on the orthonormal Legendre values of core/math/legendre.h:100-123.
solve() is the per-shell step (:336-361) on the oracle's ICLS restatement (oracle/csd_oracle.c).
"""
import math

import numpy as np

from . import orc

Y00 = 0.28209479177387814


def plm_sph_column(lmax, m, x):
    """Orthonormal P_l^m(x), l = m..lmax, at out[l] (legendre.h:100-123; seed by legendre.h:66-69)."""
    out = np.zeros(lmax + 2)
    x2 = x * x
    if m and x2 >= 1.0:
        return out
    seed = Y00
    if m:
        prod = 1.0
        for k in range(1, m + 1):
            prod = ((1.0 - x2) * (2 * k - 1)) / (2 * k) * prod
        seed *= math.sqrt((2 * m + 1) * prod)
        if m & 1:
            seed = -seed
    out[m] = seed
    if lmax == m:
        return out
    f = math.sqrt(2 * m + 3)
    out[m + 1] = x * f * out[m]
    for l in range(m + 2, lmax + 1):
        v = x * out[l - 1] - out[l - 2] / f
        f = math.sqrt((4 * l * l - 1) / (l * l - m * m))
        out[l] = v * f
    return out


def amp_transform(els, lmax):
    """ZSH::init_amp_transform (ZSH.h:66-77): [len(els), lmax/2 + 1]."""
    T = np.zeros((len(els), lmax // 2 + 1))
    for i, el in enumerate(els):
        p = plm_sph_column(lmax, 0, math.cos(el))
        T[i] = p[0:lmax + 1:2]
    return T


def deriv_transform(els, lmax):
    """ZSH::init_deriv_transform (ZSH.h:87-99)."""
    T = np.zeros((len(els), lmax // 2 + 1))
    for i, el in enumerate(els):
        p = plm_sph_column(lmax, 1, math.cos(el))
        for l in range(2, lmax + 1, 2):
            T[i, l // 2] = p[l] * math.sqrt(l * (l + 1))
    return T


def rotation(d, rng):
    """gen_rotation_matrix (amp2response.cpp:72-90)."""
    def normalized(v):   # Eigen's normalize(): a zero vector is left as it is
        n = np.linalg.norm(v)
        return v / n if n > 0 else v

    R = np.zeros((3, 3))
    R[2] = d
    v2 = normalized(np.cross(d, rng.standard_normal(3)))
    R[0] = v2
    R[1] = normalized(np.cross(d, v2))
    return R


def accumulate(amp, fibre, mask, dirs, volumes, lmax, seed=0):
    """Sum of T^T T and T^T a over the masked voxels; amp [nvox, ld], fibre [nvox, >=3]."""
    rng = np.random.default_rng(seed)
    n = lmax // 2 + 1
    M = np.zeros((n, n))
    b = np.zeros(n)
    count = 0
    dirs = np.asarray(dirs, dtype=np.float64)
    for v in range(len(amp)):
        if mask is not None and not mask[v]:
            continue
        count += 1
        d = np.asarray(fibre[v, :3], dtype=np.float64)
        nrm = np.linalg.norm(d)
        if nrm > 0:
            d = d / nrm
        R = rotation(d, rng)
        rot = dirs @ R.T
        r = np.sqrt((rot ** 2).sum(axis=1))
        az = np.arctan2(rot[:, 1], rot[:, 0])
        with np.errstate(invalid="ignore", divide="ignore"):
            el = np.arccos(rot[:, 2] / r)   # NaN for a zero fibre direction, as the reference
        fold = el > math.pi / 2
        az = np.where(fold, np.where(az > math.pi, az - math.pi, az + math.pi), az)
        el = np.where(fold, math.pi - el, el)
        T = amp_transform(el, lmax)
        a = np.asarray(amp[v, volumes], dtype=np.float64)
        b += T.T @ a
        M += T.T @ T
    return M, b, count


def constraints(lmax):
    """amp2response.cpp:343-354: amplitudes then derivatives at 0, 1, ..., 90 degrees."""
    els = np.arange(91) * math.pi / 180.0
    return np.concatenate([amp_transform(els, lmax), deriv_transform(els, lmax)])


def solve(lmax, M, b, constrained=True):
    """amp2response.cpp:336-361; returns (rf, iterations)."""
    if not lmax or not constrained:
        return np.linalg.solve(M, b), 0
    o = orc.IclsOracle(M, constraints(lmax), None, 0, 1e-10, 1e-10, 0, 0.0, standard_form=True)
    x, it, _ = o.solve(b)
    return x, it
