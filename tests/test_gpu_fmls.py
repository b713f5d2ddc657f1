"""GPU fod2fixel segmentation (include/mrfixel.h) against the FP64 oracle (oracle/fmls_oracle.c) on
the same float32 FODs.

Bar: identical fixel counts in every voxel (integer work: the segmentation follows the same
amplitudes, accumulated in the same order, through the same decisions), and every fixel's
direction, integral and peak amplitude equal to the oracle's rounded to float32 within 2e-6
relative (the Newton peak search and the mean direction differ from the oracle only by the last
bits of the device's sin / cos / atan2 and by fused multiply-adds)."""
import ctypes as C
import sys

import numpy as np
import pytest

from mrtrix3_amd import fixel, phantom
from oracle import orc

pytestmark = pytest.mark.gpu

RTOL, ATOL = 2e-6, 2e-7


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def stick_fod(dirs, weights, lmax=8):
    dirs = np.atleast_2d(np.asarray(dirs, dtype=np.float64))
    dirs = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    az = np.arctan2(dirs[:, 1], dirs[:, 0])
    el = np.arccos(np.clip(dirs[:, 2], -1, 1))
    return np.asarray(weights, dtype=np.float64) @ orc.sh_transform(np.stack([az, el], axis=1), lmax)


def synthetic(n, seed, lmax=8, noise=0.02):
    """n FODs of one to three fibres with SH noise."""
    rng = np.random.default_rng(seed)
    nsh = (lmax + 1) * (lmax + 2) // 2
    out = np.zeros((n, nsh))
    for v in range(n):
        nf = 1 + v % 3
        out[v] = stick_fod(rng.normal(size=(nf, 3)), rng.uniform(0.2, 0.6, nf), lmax)
        out[v] += rng.normal(scale=noise * (1 + v % 4), size=nsh)
        out[v, 0] = abs(out[v, 0]) + 0.05
    return out.astype(np.float32)


def edge_cases():
    const = np.zeros(45); const[0] = 0.7                      # over 127 lobes: the 16-bit growth pass
    neg = -stick_fod([[1, 0, 0]], [0.5]); neg[0] = 0.1        # largest |amplitude| negative
    z = stick_fod([[0, 1, 0]], [0.5]); z[0] = 0.0             # c0 = 0
    nan = stick_fod([[0, 1, 0]], [0.5]); nan[0] = np.nan      # c0 NaN
    weak = stick_fod([[0.2, 0.1, 1]], [0.02])                # below the peak threshold
    zs = stick_fod([[0, 0, 1]], [0.5])                        # axially symmetric: near-equal amplitudes
    xs = stick_fod([[1, 0, 0]], [0.5])
    return np.array([const, neg, z, nan, weak, zs, xs], dtype=np.float32)


def expected(o, fods, mask=None, maxnum=0, dirpeak=False):
    out = []
    for v in range(fods.shape[0]):
        if mask is not None and not mask[v]:
            out.append(np.zeros((0, 5), dtype=np.float32))
            continue
        lobes, _ = o.segment(fods[v].astype(np.float64), max_out=2048)
        if maxnum:
            lobes = lobes[:maxnum]
        d = lobes[:, 3:6] if dirpeak else lobes[:, 0:3]
        out.append(np.concatenate([d, lobes[:, 6:8]], axis=1).astype(np.float32))
    return out


def check(res, exp, label):
    count = np.asarray(res.count).astype(np.int64)
    offset = np.asarray(res.offset).astype(np.int64)
    fix = np.asarray(res.fixels)
    want_counts = np.array([len(e) for e in exp])
    bad = np.nonzero(count != want_counts)[0]
    assert len(bad) == 0, "%s: %d voxels with another fixel count, first %s: %s vs %s" % (
        label, len(bad), bad[:5], count[bad[:5]], want_counts[bad[:5]])
    assert fix.shape[0] == count.sum()
    run = 0
    worst = 0.0
    for v in range(len(exp)):
        assert offset[v] == (run if count[v] else 0), "%s: voxel %d offset" % (label, v)
        got = fix[run:run + count[v]]
        np.testing.assert_allclose(got, exp[v], rtol=RTOL, atol=ATOL, err_msg="%s: voxel %d" % (label, v))
        if count[v]:
            worst = max(worst, float(np.max(np.abs(got - exp[v]) / (np.abs(exp[v]) + 1e-6))))
        run += count[v]
    say("%s: %d voxels, %d fixels, worst relative difference %.2e" % (label, len(exp), run, worst))


@pytest.mark.parametrize("cfg", [
    dict(),
    dict(maxnum=2, dir_from_peak=True, lobe_merge_ratio=0.3),
    dict(integral_threshold=0.02, peak_value_threshold=0.0, lobe_merge_ratio=0.6),
    dict(no_thresholds=True),
], ids=["defaults", "merges_dirpeak_maxnum", "integral_threshold", "no_thresholds"])
def test_synthetic_fods_against_oracle(torch_cuda, cfg):
    torch = torch_cuda
    fods = np.concatenate([synthetic(300, seed=5), edge_cases()])
    mask = np.ones(len(fods), dtype=np.uint8)
    mask[7::23] = 0
    say("synthetic FODs, options %s" % cfg)
    seg = fixel.Segmenter(45, **cfg)
    res = seg.segment(torch.from_numpy(fods).cuda(), torch.from_numpy(mask).cuda(), summary=True)
    torch.cuda.synchronize()
    res = fixel.Fixels(res.count.cpu().numpy(), res.offset.cpu().numpy(), res.fixels.cpu().numpy(),
                       res.summary.cpu().numpy())
    ithr = cfg.get("integral_threshold", 0.0)
    pthr = cfg.get("peak_value_threshold", 0.1)
    if cfg.get("no_thresholds"):
        ithr = pthr = 0.0
    o = orc.FmlsOracle(8, integral=ithr, peak_value=pthr, merge_ratio=cfg.get("lobe_merge_ratio", 1.0))
    exp = expected(o, fods, mask, cfg.get("maxnum", 0), cfg.get("dir_from_peak", False))
    check(res, exp, "synthetic %s" % cfg)
    # the summary: peak amplitude of fixels 0 and 1, direction of fixel 0
    for v in range(len(exp)):
        p = [exp[v][i, 4] if i < len(exp[v]) else 0.0 for i in range(2)]
        np.testing.assert_allclose(res.summary[v, :2], p, rtol=RTOL, atol=ATOL)
        if len(exp[v]):
            np.testing.assert_allclose(res.summary[v, 2:], exp[v][0, :3], rtol=RTOL, atol=ATOL)
        else:
            assert np.all(np.isnan(res.summary[v, 2:]))


@pytest.fixture(scope="module")
def phantom_fods():
    """FODs of 4,096 voxels of the small single-shell phantom, from the CSD oracle (lmax 8)."""
    ph = phantom.single_shell_small()
    o = orc.CsdOracle(ph.grad, ph.response)
    dw = np.ascontiguousarray(ph.dw[:4096])
    fod, _ = o.run_volume(dw, (64, 64, 1), nthreads=8)
    return np.ascontiguousarray(fod, dtype=np.float32)


def test_phantom_fods_against_oracle(torch_cuda, phantom_fods):
    say("phantom FODs: %d voxels" % len(phantom_fods))
    seg = fixel.Segmenter(45)
    res = seg.segment(phantom_fods)          # host pointers through mrfixel_segment_host
    exp = expected(orc.FmlsOracle(8), phantom_fods)
    check(res, exp, "phantom")
    assert res.count.max() >= 2 and np.mean(res.count) > 1.0


def test_lmax_and_fod_image_layout(torch_cuda):
    # lmax 6 and 10 series, rows with padding after the coefficients (ld > n_sh)
    torch = torch_cuda
    for lmax in (6, 10):
        nsh = (lmax + 1) * (lmax + 2) // 2
        f = synthetic(120, seed=lmax, lmax=lmax, noise=0.01)
        padded = np.zeros((len(f), nsh + 3), dtype=np.float32)
        padded[:, :nsh] = f
        padded[:, nsh:] = 1e9      # never read
        say("lmax %d with row stride %d" % (lmax, nsh + 3))
        seg = fixel.Segmenter(nsh)
        res = seg.segment(torch.from_numpy(padded).cuda())
        torch.cuda.synchronize()
        res = fixel.Fixels(res.count.cpu().numpy(), res.offset.cpu().numpy(), res.fixels.cpu().numpy())
        check(res, expected(orc.FmlsOracle(lmax), f), "lmax %d" % lmax)


def test_capacity_shortfall_reports_the_total(torch_cuda):
    torch = torch_cuda
    fods = synthetic(200, seed=9)
    seg = fixel.Segmenter(45)
    full = seg.segment(fods)
    n = int(full.count.sum())
    say("capacity shortfall: %d fixels, capacity %d" % (n, n // 3))
    lib = fixel._lib.load()
    count = np.zeros(200, dtype=np.uint32)
    offset = np.zeros(200, dtype=np.uint32)
    fix = np.zeros((n // 3, 5), dtype=np.float32)
    total = C.c_uint64()
    fixel._lib.check(lib.mrfixel_segment_host(seg._h, fods.ctypes.data, 45, None, 200, count.ctypes.data,
                                              offset.ctypes.data, fix.ctypes.data, n // 3, None, C.byref(total)))
    assert total.value == n and np.array_equal(count, full.count)
    # offsets are complete; the fixels up to the capacity are written
    assert np.array_equal(offset, full.offset)
    np.testing.assert_array_equal(fix, full.fixels[:n // 3])
    # nothing to segment
    d = torch.zeros((0, 45), dtype=torch.float32, device="cuda")
    r = seg.segment(d)
    assert r.count.numel() == 0 and r.fixels.shape[0] == 0


def test_large_batch_properties(torch_cuda):
    """300,000 voxels (several amplitude chunks): offsets are the exclusive scan of the counts,
    directions unit vectors, integrals positive, peaks above the threshold; 400 voxels spread over
    the batch against the oracle."""
    torch = torch_cuda
    base = synthetic(997, seed=21)
    n = 300_000
    rng = np.random.default_rng(4)
    pick = rng.integers(0, len(base), n)
    fods = base[pick] * rng.uniform(0.5, 2.0, (n, 1)).astype(np.float32)
    say("large batch: %d voxels" % n)
    seg = fixel.Segmenter(45)
    t = torch.from_numpy(fods).cuda()
    res = seg.segment(t, summary=True)
    torch.cuda.synchronize()
    count = res.count.cpu().numpy().astype(np.int64)
    offset = res.offset.cpu().numpy().astype(np.int64)
    fix = res.fixels.cpu().numpy()
    scan = np.concatenate([[0], np.cumsum(count)[:-1]])
    assert np.array_equal(offset, np.where(count > 0, scan, 0))
    assert fix.shape[0] == count.sum()
    np.testing.assert_allclose(np.linalg.norm(fix[:, :3], axis=1), 1.0, atol=1e-6)
    assert np.all(fix[:, 3] > 0) and np.all(fix[:, 4] >= 0.1 * (1 - 1e-6))
    idx = np.sort(rng.choice(n, 400, replace=False))
    o = orc.FmlsOracle(8)
    for v in idx:
        e = expected(o, fods[v:v + 1])[0]
        assert count[v] == len(e), v
        np.testing.assert_allclose(fix[offset[v]:offset[v] + count[v]], e, rtol=RTOL, atol=ATOL)
