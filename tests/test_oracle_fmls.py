"""The FMLS segmentation oracle (oracle/fmls_oracle.c, src/dwi/fmls.cpp) pinned by construction.

* Lobe growth rules (fmls.cpp:180-253) on small direction graphs whose answer is worked out by hand:
  visiting order by decreasing |amplitude|, a direction adjacent to one lobe joins it, one
  adjacent to several is assigned after the loop to the first of them (and stays unlabelled
  meanwhile, so that a direction reached only through it starts a lobe of its own), lobes merge
  when the bridge amplitude over the smaller adjacent peak exceeds the ratio.
* The integration weights integrate the constant to 4 pi and every other even SH to ~0.
* Real FODs: a single fibre gives one lobe whose Newton-revised peak is the fibre axis; two fibres
  at 90 degrees give two lobes whose peaks are local maxima near the axes, ordered by peak
  amplitude; the lobes of the constant FOD cover the sphere, their integrals adding up to
  2 sqrt(pi) c.
* The fast direction lookup returns the nearest direction (brute force over all directions).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import orc


@pytest.fixture(scope="module")
def fm():
    return orc.FmlsOracle(8)


def graph_segment(values, edges, weights=None, merge_ratio=1.0):
    """Segments values on a direction graph given by its edges (no SH: the amplitude transform is
    the identity, peaks are not revised)."""
    n = len(values)
    adj = -np.ones((n, orc.FMLS_MAX_ADJ), dtype=np.int32)
    nb = [[] for _ in range(n)]
    for a, b in edges:
        nb[a].append(b)
        nb[b].append(a)
    for i in range(n):
        v = sorted(nb[i])
        adj[i, :len(v)] = v
    L = orc.lib()
    f = L.orc_fmls_segment
    f.restype = C.c_int
    f.argtypes = [orc.D, C.c_int, orc.D, C.c_int, orc.D, orc.I, orc.D, C.c_void_p, orc.D, C.c_int,
                  C.c_double, C.c_double, C.c_double, orc.D, C.c_int, orc.I]
    sh = np.ascontiguousarray(values, dtype=np.float64)
    lmax = 0
    while (lmax + 1) * (lmax + 2) // 2 < n:
        lmax += 2
    nsh = (lmax + 1) * (lmax + 2) // 2
    sh = np.r_[sh, np.zeros(nsh - n)]
    SHT = np.zeros((n, nsh))
    SHT[np.arange(n), np.arange(n)] = 1.0
    unit = np.tile([0.0, 0.0, 1.0], (n, 1))
    w = np.ones(n) if weights is None else np.asarray(weights, dtype=np.float64)
    out = np.zeros((32, 8))
    lut = np.zeros(n, dtype=np.int32)
    k = f(orc._d(sh), lmax, orc._d(SHT), n, orc._d(unit), adj.ctypes.data_as(orc.I), orc._d(w), None, None, 0,
          0.0, 0.0, merge_ratio, orc._d(out), 32, lut.ctypes.data_as(orc.I))
    return out[:k], lut


def test_path_graph_retrospective_assignment():
    # a - b - c - d - e with a=5, b=2, c=4, d=3, e=1: visiting a, c, d, b, e.  a and c start lobes, d
    # joins c, b touches both (assigned to the first, a's, after the loop), e joins d's lobe.
    lobes, lut = graph_segment([5, 2, 4, 3, 1], [(0, 1), (1, 2), (2, 3), (3, 4)])
    assert len(lobes) == 2
    assert list(lut) == [0, 0, 1, 1, 1]
    assert lobes[0, 6] == 7 and lobes[1, 6] == 8       # integrals with unit weights
    assert lobes[0, 7] == 5 and lobes[1, 7] == 4       # seed amplitudes (peaks not revised)


def test_saddle_stays_unlabelled_during_the_loop():
    # f (=1.5) hangs off b only: when f is visited, b is still waiting for its retrospective
    # assignment, so f starts a third lobe
    lobes, lut = graph_segment([5, 2, 4, 3, 1, 1.5], [(0, 1), (1, 2), (2, 3), (3, 4), (1, 5)])
    assert len(lobes) == 3
    assert list(lut) == [0, 0, 1, 1, 1, 2]
    assert lobes[2, 6] == 1.5


def test_negative_directions_never_join_positive_lobes_and_weights_apply():
    # c < 0 separates a and e; negative lobes are dropped and negative directions map to "none"
    lobes, lut = graph_segment([5, 2, -4, 3, 1], [(0, 1), (1, 2), (2, 3), (3, 4)], weights=[1, 2, 1, 0.5, 4])
    assert len(lobes) == 2
    assert list(lut) == [0, 0, 2, 1, 1]
    assert lobes[0, 6] == 5 + 2 * 2 and lobes[1, 6] == 3 * 0.5 + 1 * 4


def test_lobes_merge_above_the_ratio():
    # with ratio 0.3 the bridge b (2 / 4 = 0.5 of the smaller adjacent peak) merges c's lobe into a's
    lobes, lut = graph_segment([5, 2, 4, 3, 1], [(0, 1), (1, 2), (2, 3), (3, 4)], merge_ratio=0.3)
    assert len(lobes) == 1 and list(lut) == [0] * 5
    assert lobes[0, 6] == 15 and lobes[0, 7] == 5
    # with 0.6 the same bridge does not merge
    lobes, _ = graph_segment([5, 2, 4, 3, 1], [(0, 1), (1, 2), (2, 3), (3, 4)], merge_ratio=0.6)
    assert len(lobes) == 2


def test_non_positive_maximum_gives_no_lobes():
    # the largest |amplitude| is negative: no lobe at all, even with positive directions elsewhere
    lobes, lut = graph_segment([1, -6, 2, 3], [(0, 1), (1, 2), (2, 3)])
    assert len(lobes) == 0


def test_adjacency_is_the_antipodal_convex_hull(fm):
    # every direction has 5 or 6 neighbours on the 1281 tessellation (12 icosahedral vertices come
    # in 6 antipodal pairs), adjacency is symmetric, and every neighbour is among the 8 nearest
    deg = (fm.adj >= 0).sum(axis=1)
    assert set(np.unique(deg)) == {5, 6} and np.sum(deg == 5) == 6
    for i in range(fm.n):
        for j in fm.adj[i][fm.adj[i] >= 0]:
            assert i in fm.adj[j]
        near = np.argsort(-np.abs(fm.unit @ fm.unit[i]))[1:9]
        assert set(fm.adj[i][fm.adj[i] >= 0]) <= set(near)


def test_integration_weights(fm):
    azel = np.stack([np.arctan2(fm.unit[:, 1], fm.unit[:, 0]), np.arccos(fm.unit[:, 2])], axis=1)
    S = orc.sh_transform(azel, 8)
    integ = fm.weights @ S
    # a least-squares fit over the basis up to lmax 50 (1326 functions, 1281 weights): the constant
    # integrates to 4 pi and the lmax-8 basis to zero up to its residual (about 1e-8 of 4 pi)
    assert abs(fm.weights.sum() - 4 * np.pi) < 1e-7
    assert abs(integ[0] - 2 * np.sqrt(np.pi)) < 1e-7 and np.max(np.abs(integ[1:])) < 1e-6
    assert np.all(fm.weights > 0)


def test_lookup_returns_the_nearest_direction(fm):
    rng = np.random.default_rng(3)
    p = rng.normal(size=(3000, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    for q in p:
        assert fm.select(q) == int(np.argmax(np.abs(fm.unit @ q)))


def stick_fod(dirs, weights, lmax=8):
    az = np.arctan2(dirs[:, 1], dirs[:, 0])
    el = np.arccos(np.clip(dirs[:, 2], -1, 1))
    return np.asarray(weights) @ orc.sh_transform(np.stack([az, el], axis=1), lmax)


def test_single_fibre_lobe(fm):
    u = np.array([[0.3, -0.5, 0.81]])
    u /= np.linalg.norm(u)
    lobes, lut = fm.segment(stick_fod(u, [1.0]))
    assert len(lobes) >= 1
    peak = lobes[0, 3:6]
    assert abs(abs(peak @ u[0]) - 1.0) < 1e-7            # the Newton peak is the axis
    # the band-limited delta peaks at sum_l (2l+1)/(4 pi) = 45 / (4 pi) on its axis
    assert abs(lobes[0, 7] - 45 / (4 * np.pi)) < 1e-3    # table interpolation in the peak search
    assert abs(abs(lobes[0, 0:3] @ u[0]) - 1.0) < 1e-3   # mean direction close to the axis
    assert np.all(np.isin(lut, np.arange(len(lobes) + 1)))


def test_two_crossing_fibres(fm):
    u = np.array([[1.0, 0.0, 0.2], [0.0, 1.0, -0.1]])
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    lobes, _ = fm.segment(stick_fod(u, [0.6, 0.4]))
    big = lobes[lobes[:, 7] > 0.5]
    assert len(big) == 2
    assert big[0, 7] > big[1, 7]                          # reference order: decreasing peak
    # each fibre's tails shift the other's maximum slightly off its axis: the peaks lie within a
    # few milliradians of the axes and are local maxima of the FOD (the exact transform around them)
    sh = stick_fod(u, [0.6, 0.4])
    for k in range(2):
        p = big[k, 3:6]
        assert np.arccos(min(1.0, abs(p @ u[k]))) < 0.01
        e1 = np.cross(p, [0.3, 0.5, 0.8]); e1 /= np.linalg.norm(e1)
        e2 = np.cross(p, e1)
        ring = np.array([np.cos(a) * e1 + np.sin(a) * e2 for a in np.linspace(0, 2 * np.pi, 16, endpoint=False)])
        pts = p[None, :] + 2e-3 * ring
        pts /= np.linalg.norm(pts, axis=1, keepdims=True)
        amp = stick_fod(np.vstack([p[None, :], pts]), np.eye(17)) @ sh
        assert np.all(amp[1:] < amp[0]) and abs(amp[0] - big[k, 7]) < 1e-4 * amp[0]


def test_constant_fod_integrates_to_the_sphere(fm):
    # every direction is positive with the same amplitude c / (2 sqrt pi): visiting order follows
    # rounding noise and the sphere splits into many lobes, but every direction lands in one of them,
    # so the lobe integrals add up to the integral of the FOD over the sphere, 2 sqrt(pi) c
    c = 0.7
    sh = np.zeros(45)
    sh[0] = c
    lobes, lut = fm.segment(sh, max_out=1024)
    assert np.all(lut < len(lobes))
    assert abs(lobes[:, 6].sum() - 2 * np.sqrt(np.pi) * c) < 1e-7


def test_empty_and_thresholds(fm):
    assert len(fm.segment(np.zeros(45))[0]) == 0
    sh = np.zeros(45)
    sh[0] = np.nan
    assert len(fm.segment(sh)[0]) == 0
    weak = stick_fod(np.array([[0.0, 0.0, 1.0]]), [0.02])   # peak 0.07 < 0.1
    assert len(fm.segment(weak)[0]) == 0
    nothr = orc.FmlsOracle(8, integral=0.0, peak_value=0.0)
    assert len(nothr.segment(weak)[0]) >= 1
