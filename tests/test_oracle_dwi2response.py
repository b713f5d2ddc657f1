"""The dwi2response oracle (oracle/dwi2response_oracle.py: tournier and tax) pinned on the CPU: its
image arithmetic against independent forms (scipy's binary dilation, a sorted selection, float64,
closed-form statistics), its loop control against scripted sequences of steps, and each whole
loop on a phantom whose single-fibre voxels are known."""
import math

import numpy as np
import pytest
from scipy import ndimage

from mrtrix3_amd import phantom
from oracle import dwi2response_oracle as T


def test_dilate_matches_scipy_face_connectivity():
    rng = np.random.default_rng(0)
    shape = (9, 7, 5)
    st = ndimage.generate_binary_structure(3, 1)
    for npass in (1, 2, 3):
        m = rng.random(np.prod(shape)) < 0.05
        want = ndimage.binary_dilation(m.reshape(shape[::-1]), structure=st, iterations=npass).reshape(-1)
        assert np.array_equal(T.dilate(m, shape, npass), want)
    assert np.array_equal(T.dilate(np.zeros(np.prod(shape), bool), shape, 2), np.zeros(np.prod(shape), bool))


def test_threshold_top_is_the_n_largest_with_ties_and_without_nan():
    rng = np.random.default_rng(1)
    cf = rng.random(500).astype(np.float32)
    cf[::17] = np.nan
    for n in (1, 10, 100, int(np.sum(~np.isnan(cf)))):
        sel = T.threshold_top(cf, n)
        vals = np.sort(cf[~np.isnan(cf)])[::-1]
        assert np.array_equal(sel, ~np.isnan(cf) & (cf >= vals[n - 1]))
        assert sel.sum() == n          # distinct values: exactly n
    ties = np.array([3, 1, 2, 2, 2, np.nan, 0], dtype=np.float32)
    assert T.threshold_top(ties, 2).tolist() == [True, False, True, True, True, False, False]   # ties kept (GE)
    inf = np.array([np.inf, 1, 2], dtype=np.float32)
    assert T.threshold_top(inf, 1).tolist() == [True, False, False]    # +inf is a value, NaN is not
    with pytest.raises(ValueError, match="less than number of voxels"):
        T.threshold_top(ties, 7)
    with pytest.raises(ValueError, match="No valid input data"):
        T.threshold_top(np.array([np.nan], dtype=np.float32), 1)


def test_cost_in_single_precision():
    rng = np.random.default_rng(2)
    p1 = rng.uniform(0.1, 2.0, 1000).astype(np.float32)
    p2 = (p1 * rng.uniform(0, 1, 1000)).astype(np.float32)
    c = T.cost(p1, p2)
    assert c.dtype == np.float32
    ref = np.sqrt(p1.astype(np.float64)) * (1.0 - p2.astype(np.float64) / p1) ** 2
    np.testing.assert_allclose(c, ref, rtol=3e-6, atol=1e-7)
    # a single fibre costs sqrt(p1); equal peaks cost 0; no fixel (p1 = 0) is NaN
    assert T.cost([4.0], [0.0])[0] == 2.0 and T.cost([4.0], [4.0])[0] == 0.0
    assert np.isnan(T.cost([0.0], [0.0])[0])


def test_loop_control_against_scripted_steps():
    """Termination when an iteration selects no voxel the previous one did not (mrstats max of the
    difference is 0), the response of that iteration, the first iteration's lmax 4 and '1 -1 1'
    response, later iterations' -lmax, and the iteration cap."""
    shape = (4, 4, 1)
    mask = np.ones(16, bool)
    calls = []

    def scripted(selections):
        def fn(dw, grad, shp, mask_in, rf_in, amp_lmax, number, nthreads=8):
            k = len(calls)
            calls.append((list(np.ravel(rf_in)), amp_lmax, mask_in.copy()))
            sf = np.zeros(16, bool)
            sf[selections[min(k, len(selections) - 1)]] = True
            cf = np.where(sf, 2.0, np.arange(16) / 100.0).astype(np.float32)   # distinct costs
            return dict(rf=np.array([100.0 + k]), cf=cf, count=16, sf=sf)
        return fn

    rf, sf, its, conv, hist = T.run(None, None, shape, mask, number=2, iter_voxels=3, lmax=6,
                                    step_fn=scripted([[0, 1], [1, 2], [1, 2]]))
    assert conv and its == 3 and rf[0] == 102.0 and sorted(np.nonzero(sf)[0]) == [1, 2]
    assert calls[0][:2] == ([1.0, -1.0, 1.0], 4) and calls[1][:2] == ([100.0], 6) and calls[2][:2] == ([101.0], 6)
    # the next mask: the top iter_voxels = 3 costs (voxels 0, 1 and 15) and their face neighbours
    assert np.nonzero(calls[1][2])[0].tolist() == [0, 1, 2, 4, 5, 11, 14, 15]
    # a selection that only loses voxels also terminates
    calls.clear()
    _, _, its, conv, _ = T.run(None, None, shape, mask, number=2, step_fn=scripted([[0, 1, 5], [0, 1]]))
    assert conv and its == 2
    # no convergence: the last iteration's response and selection
    calls.clear()
    rf, sf, its, conv, _ = T.run(None, None, shape, mask, number=2, max_iters=4,
                                 step_fn=scripted([[0], [1], [2], [3], [4]]))
    assert not conv and its == 4 and rf[0] == 103.0 and np.nonzero(sf)[0].tolist() == [3]
    with pytest.raises(ValueError, match="at least 2"):
        T.run(None, None, shape, mask, max_iters=1)
    with pytest.raises(ValueError, match="iter_voxels"):
        T.run(None, None, shape, mask, number=10, iter_voxels=5)


def test_loop_selects_single_fibre_voxels_of_a_phantom():
    """On 2,048 voxels of one to three sticks (SNR 30), the loop converges on voxels that are almost
    all single-fibre, and its response has the generating response's signs and shape (a cost of
    the wrong sign or a mis-indexed peak would select crossings and flatten it)."""
    ph = phantom.single_shell_small()
    n = 2048
    keep = ph.grad[:, 3] > 10
    dw = np.ascontiguousarray(ph.dw[:n][:, keep])
    rf, sf, its, conv, hist = T.run(dw, ph.grad[keep], (16, 16, 8), np.ones(n, bool), number=100, iter_voxels=600)
    assert conv and its <= 10 and sf.sum() == 100
    single = ph.extra["count"][:n] == 1
    assert single[sf].mean() >= 0.9 > single.mean()
    truth = np.ravel(ph.response)
    np.testing.assert_allclose(rf[:5], truth, rtol=0.15, atol=0.03 * truth[0])
    assert abs(rf[5]) < 0.01 * truth[0]       # lmax 10 term of a response generated to lmax 8


def test_printed6_is_ostream_default_formatting():
    assert T.printed6(123.456789) == 123.457
    assert T.printed6(0.000123456789) == 0.000123457
    assert T.printed6(1234567.0) == 1.23457e6
    assert T.printed6(2.0) == 2.0


def test_tax_initial_response_closed_forms():
    """A constant signal in the mask: s = 0, rf = [c sqrt(4 pi), 0, 0]; values outside the mask and
    non-finite values do not count; a two-valued signal has mean (a + b)/2 and sample standard
    deviation |a - b|/2 sqrt(n / (n - 1)), each as mrstats prints it (6 digits)."""
    dw = np.full((10, 4), 7.0, dtype=np.float32)
    mask = np.ones(10, bool)
    dw[3] = 1000.0
    mask[3] = False
    dw[5, 2] = np.nan
    rf = T.tax_init_rf(dw, mask)
    assert rf[0] == 7.0 * math.sqrt(4 * math.pi) and rf[1] == 0.0 and rf[2] == 0.0
    a, b = 100.0, 130.0
    dw2 = np.array([[a, b]] * 6, dtype=np.float32)
    n = 12
    m = T.printed6((a + b) / 2) * math.sqrt(4 * math.pi)
    sd = T.printed6(abs(a - b) / 2 * math.sqrt(n / (n - 1))) * math.sqrt(4 * math.pi)
    np.testing.assert_allclose(T.tax_init_rf(dw2, np.ones(6, bool)), [m, -0.5 * sd, 0.25 * sd * sd / m], rtol=1e-15)


def test_tax_loop_control_against_scripted_steps():
    """Convergence when every coefficient's half-difference is within the percentage of its mean
    (a coefficient whose mean is negative never asks for more), the selection as the next mask,
    and the iteration cap."""
    assert T.tax_converged([100.0, 10.0], [100.4, 10.04], 0.5)
    assert not T.tax_converged([100.0, 10.0], [100.4, 10.2], 0.5)
    assert T.tax_converged([100.0, -1.0], [100.0, -0.5], 0.5)      # negative mean: the script's quirk
    shape = (4, 4, 1)
    calls = []

    def scripted(rfs, sels):
        def fn(dw, grad, shp, mask_in, rf_in, amp_lmax, peak_ratio, nthreads=8):
            k = len(calls)
            calls.append((np.array(rf_in, dtype=float), amp_lmax, mask_in.copy(), peak_ratio))
            sf = np.zeros(16, bool)
            sf[sels[min(k, len(sels) - 1)]] = True
            return dict(rf=np.array(rfs[min(k, len(rfs) - 1)], dtype=float), sf=sf)
        return fn

    rf, sf, its, conv, hist = T.tax_run(None, None, shape, np.ones(16, bool), lmax=8, init_rf=[1.0, 2.0, 3.0],
                                        step_fn=scripted([[10.0, 5.0], [11.0, 5.0], [11.01, 5.0]],
                                                         [[0, 1, 2], [1, 2], [2]]))
    assert conv and its == 3 and rf[0] == 11.01 and np.nonzero(sf)[0].tolist() == [2]
    assert calls[0][0].tolist() == [1.0, 2.0, 3.0] and calls[1][0].tolist() == [10.0, 5.0]
    assert np.nonzero(calls[1][2])[0].tolist() == [0, 1, 2] and np.nonzero(calls[2][2])[0].tolist() == [1, 2]
    assert all(c[1] == 8 and c[3] == 0.1 for c in calls)
    calls.clear()
    rf, sf, its, conv, _ = T.tax_run(None, None, shape, np.ones(16, bool), max_iters=3, init_rf=[1.0],
                                     step_fn=scripted([[10.0], [20.0], [40.0], [80.0]], [[0]]))
    assert not conv and its == 3 and rf[0] == 40.0


def test_tax_loop_on_a_phantom():
    """On 2,048 voxels of one to three sticks (SNR 30) the tax loop converges on a selection of
    mostly single-fibre voxels (second peak under a tenth of the first) and a response of the
    generating response's shape."""
    ph = phantom.single_shell_small()
    n = 2048
    keep = ph.grad[:, 3] > 10
    dw = np.ascontiguousarray(ph.dw[:n][:, keep])
    rf, sf, its, conv, hist = T.tax_run(dw, ph.grad[keep], (16, 16, 8), np.ones(n, bool))
    assert conv and its <= 20
    single = ph.extra["count"][:n] == 1
    assert single[sf].mean() >= 0.7 > 0.4 > single.mean()
    assert all(h["sf"].sum() <= (hist[i - 1]["sf"].sum() if i else n) for i, h in enumerate(hist))
    truth = np.ravel(ph.response)
    np.testing.assert_allclose(rf[:5], truth, rtol=0.2, atol=0.03 * truth[0])
