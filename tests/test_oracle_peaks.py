"""The single-precision restatement of sh2peaks (oracle/peaks_oracle.c), pinned without a GPU:
Legendre values against the double-precision oracle (itself pinned by closed forms), amplitude
and the five derivatives against the double SH transform and its central finite differences,
peak positions against FODs whose peak is known exactly (axially symmetric lobes), a dense-grid
search on random FODs, and the per-voxel rules of sh2peaks.cpp (de-duplication, threshold,
ordering, NaN / empty input, matching to given directions)."""
import numpy as np
import pytest

from oracle import orc

LMAX = 8
NSH = 45


def amp_double(sh, el, az):
    """sum_j Y_j(el, az) sh_j in double precision, from the pinned SH transform."""
    Y = orc.sh_transform(np.array([[az, el]]), LMAX)
    return float(Y[0] @ np.asarray(sh, dtype=np.float64))


def unit(az, el):
    return np.array([np.cos(az) * np.sin(el), np.sin(az) * np.sin(el), np.cos(el)])


def angles(v):
    return np.arctan2(v[1], v[0]), np.arctan2(np.hypot(v[0], v[1]), v[2])


def delta_fod(dirs, weights=None):
    """SH coefficients of sum_k w_k * delta(u_k), truncated at LMAX: coefficient j = sum w_k Y_j(u_k).
    A single lobe is axially symmetric about u_k, so its maximum is exactly at +-u_k."""
    dirs = np.atleast_2d(dirs)
    w = np.ones(len(dirs)) if weights is None else np.asarray(weights)
    ang = np.array([angles(u) for u in dirs])
    Y = orc.sh_transform(ang, LMAX)
    return (w @ Y).astype(np.float32)


def test_legendre_float_matches_double():
    rng = np.random.default_rng(0)
    for _ in range(40):
        el = rng.uniform(0.0, np.pi)
        x, s = np.cos(el), np.sin(el)
        for m in range(LMAX + 1):
            f = orc.plm_sph_f(LMAX, m, x, s)[m:]
            d = orc.plm(LMAX, m, x)[m:]
            assert np.allclose(f, d, rtol=2e-5, atol=2e-6), (el, m)


def test_amplitude_and_derivatives_against_finite_differences():
    rng = np.random.default_rng(1)
    sh = rng.standard_normal(NSH).astype(np.float32)
    h = 1e-4
    for _ in range(25):
        el, az = rng.uniform(0.2, np.pi - 0.2), rng.uniform(-np.pi, np.pi)
        a, d_el, d_az, d2_el, d2_elaz, d2_az = orc.sh_derivatives_f(sh, LMAX, el, az).astype(np.float64)
        A = lambda e, z: amp_double(sh, e, z)
        s = np.sin(el)
        fd = [A(el, az),
              (A(el + h, az) - A(el - h, az)) / (2 * h),
              (A(el, az + h) - A(el, az - h)) / (2 * h) / s,
              (A(el + h, az) - 2 * A(el, az) + A(el - h, az)) / h ** 2,
              (A(el + h, az + h) - A(el + h, az - h) - A(el - h, az + h) + A(el - h, az - h)) / (4 * h * h) / s,
              (A(el, az + h) - 2 * A(el, az) + A(el, az - h)) / h ** 2 / s ** 2]
        got = [a, d_el, d_az, d2_el, d2_elaz, d2_az]
        scale = np.abs(sh).sum()
        for g, f, tol in zip(got, fd, [1e-5, 1e-4, 1e-4, 5e-3, 5e-3, 5e-3]):
            assert abs(g - f) <= tol * scale, (el, az, got, fd)


def test_peak_of_a_single_lobe_is_its_axis():
    rng = np.random.default_rng(2)
    for _ in range(20):
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        sh = delta_fod(u)
        # a seed 30 degrees off the axis
        w = np.cross(u, rng.standard_normal(3))
        w /= np.linalg.norm(w)
        seed = np.cos(0.5) * u + np.sin(0.5) * w
        a, d, it = orc.sh_get_peak_f(sh, LMAX, seed / np.linalg.norm(seed))
        assert np.isfinite(a) and it < 50
        assert abs(abs(d @ u) - 1.0) < 1e-6          # within ~1.4e-3 rad of the axis
        assert abs(a - amp_double(sh, *angles(u)[::-1])) <= 1e-4 * abs(a)


def test_two_lobes_deduplicated_and_ordered():
    u1 = unit(0.3, 1.1)
    u2 = unit(0.3 + 1.4, 1.0)     # about 75 degrees apart
    sh = delta_fod(np.stack([u1, u2]), [1.0, 0.6])
    out, _ = orc.sh2peaks(sh[None, :], npeaks=3)
    p = out[0].reshape(3, 3)
    assert np.all(np.isfinite(p[:2]))
    a = np.linalg.norm(p[:2], axis=1)
    assert a[0] > a[1]                               # decreasing amplitude
    # a delta truncated at lmax 8 rings: a third maximum, if any, is a small side lobe
    assert np.all(np.isnan(p[2])) or np.linalg.norm(p[2]) < 0.25 * a[0]
    d = p[:2] / a[:, None]
    assert abs(abs(d[0] @ u1) - 1) < 1e-3 and abs(abs(d[1] @ u2) - 1) < 1e-3
    # -num 1 keeps only the largest
    out1, _ = orc.sh2peaks(sh[None, :], npeaks=1)
    assert np.allclose(out1[0], out[0][:3])
    # a threshold between the two amplitudes keeps one
    outt, _ = orc.sh2peaks(sh[None, :], npeaks=3, threshold=float(a.mean()))   # side lobes fall below it too
    assert np.all(np.isfinite(outt[0][:3])) and np.all(np.isnan(outt[0][3:]))


def test_peaks_are_local_maxima_on_random_fods():
    rng = np.random.default_rng(3)
    grid_el = np.linspace(0.01, np.pi - 0.01, 60)
    grid_az = np.linspace(-np.pi, np.pi, 120, endpoint=False)
    E, Z = np.meshgrid(grid_el, grid_az, indexing="ij")
    Yg = orc.sh_transform(np.stack([Z.ravel(), E.ravel()], axis=1), LMAX)
    for _ in range(5):
        dirs = rng.standard_normal((3, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        sh = delta_fod(dirs, rng.uniform(0.3, 1.0, 3))
        sh[1:] += (0.02 * rng.standard_normal(NSH - 1)).astype(np.float32)
        out, _ = orc.sh2peaks(sh[None, :], npeaks=6)
        peaks = out[0].reshape(6, 3)
        peaks = peaks[np.isfinite(peaks[:, 0])]
        assert len(peaks) >= 1
        vals = Yg @ sh.astype(np.float64)
        top = vals.max()
        # the largest peak found is the global maximum of the dense grid, up to the grid spacing
        assert np.linalg.norm(peaks[0]) >= top - 1e-3 * abs(top)
        for p in peaks:
            a = np.linalg.norm(p)
            az, el = angles(p / a)
            # a local maximum: small moves in any direction do not increase the amplitude
            for de, dz in [(1e-2, 0), (-1e-2, 0), (0, 1e-2), (0, -1e-2)]:
                assert amp_double(sh, el + de, az + dz / max(np.sin(el), 1e-3)) <= a + 1e-5 * abs(a)


def test_input_rules():
    sh = delta_fod(unit(0.2, 0.9))
    vols = np.stack([sh, sh, sh, sh])
    vols[1, 7] = np.nan                   # NaN coefficient -> all NaN
    vols[2, 1:] = 0.0                     # nothing past the first coefficient -> all NaN
    mask = np.array([1, 1, 1, 0], dtype=np.uint8)
    out, _ = orc.sh2peaks(vols, npeaks=2, mask=mask)
    assert np.all(np.isfinite(out[0, :3]))
    assert np.all(np.isnan(out[1])) and np.all(np.isnan(out[2])) and np.all(np.isnan(out[3]))


def test_matching_given_directions():
    u1, u2 = unit(0.3, 1.1), unit(1.7, 1.0)
    sh = delta_fod(np.stack([u1, u2]), [1.0, 0.6])
    # -direction: report the peak closest to each given direction, in the given order
    match = np.stack([u2, u1]).astype(np.float32)
    out, _ = orc.sh2peaks(sh[None, :], directions=match)
    p = out[0].reshape(2, 3)
    assert abs(abs(p[0] / np.linalg.norm(p[0]) @ u2) - 1) < 1e-3
    assert abs(abs(p[1] / np.linalg.norm(p[1]) @ u1) - 1) < 1e-3
    # -peaks: per-voxel directions; a NaN first value gives an all-NaN voxel
    per_vox = np.stack([np.concatenate([u1, u2]), np.full(6, np.nan)]).astype(np.float32)
    out2, _ = orc.sh2peaks(np.stack([sh, sh]), peaks_image=per_vox)
    assert np.allclose(out2[0].reshape(2, 3)[0], p[1], atol=1e-6) and np.all(np.isnan(out2[1]))


def test_threaded_run_equals_single_voxels():
    rng = np.random.default_rng(4)
    vols = np.stack([delta_fod(rng.standard_normal((2, 3)) / 1.7, rng.uniform(0.4, 1, 2)) for _ in range(70)])
    a, ita = orc.sh2peaks(vols, nthreads=4)
    b, itb = orc.sh2peaks(vols, nthreads=0)
    assert np.array_equal(a, b, equal_nan=True) and ita == itb


def test_fast_table_against_legendre_closed_form():
    # -fast table rows (SH.h:418-434): the SH-normalised associated Legendre functions with the
    # Condon-Shortley phase, sqrt((2l+1)/4pi (l-m)!/(l+m)!) P_l^m(cos el), at el = n pi / 511
    from math import factorial
    from scipy.special import lpmv
    for lmax in (2, 8, 12):
        tab = orc.peaks_fast_table(lmax)
        assert tab.shape == (512, (lmax // 2 + 1) ** 2)
        x = np.cos(np.arange(512) * np.pi / 511)
        for l in range(0, lmax + 1, 2):
            for m in range(0, l + 1):
                want = np.sqrt((2 * l + 1) / (4 * np.pi) * factorial(l - m) / factorial(l + m)) * lpmv(m, l, x)
                got = tab[:, l * l // 4 + m].astype(np.float64)
                # 1e-4: near the poles 1 - x^2 is formed in single precision, as the reference's
                # one-argument Plm_sph does (up to 2.1e-5 at lmax 12); a wrong sign, index or
                # normalisation is off by orders more
                assert np.max(np.abs(got - want)) <= 1e-4 * max(1.0, np.max(np.abs(want))), (lmax, l, m)
        # m > 0 rows vanish at the poles (x^2 >= 1 in single precision)
        assert np.all(tab[0, [l * l // 4 + m for l in range(2, lmax + 1, 2) for m in range(1, l + 1)]] == 0.0)


def test_fast_run_close_to_exact_run():
    # the table's linear interpolation moves the peaks by far less than the search tolerance of the
    # de-duplication (acos 0.99): the same peaks, directions within 5e-3 rad, amplitudes within 1e-3
    # lobes at least 45 degrees apart: closer ones merge, and the second reported peak is then one
    # of a ring of equal sidelobe maxima, whichever a search happens to reach
    rng = np.random.default_rng(5)
    dirs = rng.standard_normal((400, 2, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    dirs = dirs[np.abs(np.sum(dirs[:, 0] * dirs[:, 1], axis=1)) < np.cos(np.pi / 4)][:200]
    sh = np.stack([delta_fod(d, [1.0, 0.6]) for d in dirs]).astype(np.float32)
    exact, _ = orc.sh2peaks(sh, npeaks=2, nthreads=4)
    fast, _ = orc.sh2peaks(sh, npeaks=2, nthreads=4, fast=True)
    assert not np.array_equal(exact, fast)    # the table is in use
    ok = 0
    for v in range(len(sh)):
        e, f = exact[v].reshape(2, 3), fast[v].reshape(2, 3)
        if not np.array_equal(np.isfinite(e[:, 0]), np.isfinite(f[:, 0])):
            continue
        good = True
        for i in np.nonzero(np.isfinite(e[:, 0]))[0]:
            ae, af = np.linalg.norm(e[i]), np.linalg.norm(f[i])
            c = abs(e[i] @ f[i]) / (ae * af)
            good &= np.arccos(min(1.0, c)) < 5e-3 and abs(ae - af) < 1e-3 * ae
        ok += good
    assert len(sh) == 200 and ok >= 0.995 * len(sh)
