"""GPU sh2peaks (include/mrpeaks.h) against the single-precision oracle (oracle/peaks_oracle.c) on
the same float32 FODs: the reference's command (cmd/sh2peaks.cpp) and its options.

Bar: in >= 99.9% of voxels (99.8% in the option cases, on 512 voxels) the same number of reported
peaks, each within 1e-3 rad of the oracle's direction and with an amplitude within 1e-4 of
max(its amplitude, the voxel's FOD scale): the two differ only in the last bits of the
single-precision sin / cos / atan2 / hypot, which the Newton iteration does not amplify."""
import sys

import numpy as np
import pytest

from mrtrix3_amd import peaks, phantom
from oracle import orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def fods():
    """FODs of 4,096 voxels of the small single-shell phantom, from the CSD oracle (lmax 8)."""
    ph = phantom.single_shell_small()
    o = orc.CsdOracle(ph.grad, ph.response)
    dw = np.ascontiguousarray(ph.dw[:4096])
    fod, _ = o.run_volume(dw, (64, 64, 1), nthreads=8)
    return fod


def compare(got, ref, npeaks, sh):
    """Fraction of voxels whose reported peaks agree (count, direction, amplitude).  Amplitudes are
    compared relative to the voxel's FOD scale (0.28 |sh|, the RMS of the series over the sphere
    times sqrt(4 pi) ~ 1/Y00): a side lobe ringing at 1e-4 of the scale has no meaningful relative
    error of its own (a one-ulp change of the input moves it by more than 1e-4 of itself)."""
    g = got.reshape(len(got), npeaks, 3).astype(np.float64)
    r = ref.reshape(len(ref), npeaks, 3).astype(np.float64)
    scale = 0.28 * np.linalg.norm(np.nan_to_num(np.asarray(sh, dtype=np.float64)), axis=1)
    ok = np.ones(len(g), dtype=bool)
    worst_angle = 0.0
    for v in range(len(g)):
        fg, fr = np.isfinite(g[v, :, 0]), np.isfinite(r[v, :, 0])
        if not np.array_equal(fg, fr):
            ok[v] = False
            continue
        for i in np.nonzero(fr)[0]:
            ag, ar = np.linalg.norm(g[v, i]), np.linalg.norm(r[v, i])
            c = abs(g[v, i] @ r[v, i]) / (ag * ar)
            ang = np.arccos(min(1.0, c))
            worst_angle = max(worst_angle, ang)
            if ang > 1e-3 or abs(ag - ar) > 1e-4 * max(ar, scale[v]):
                ok[v] = False
    return ok.mean(), worst_angle


def test_fods_match_oracle(torch_cuda, fods):
    torch = torch_cuda
    finder = peaks.SH2Peaks(45)
    say("[gpu peaks] %d voxels, lmax 8, 60 seeds, 3 peaks" % len(fods))
    stats = torch.zeros(4, dtype=torch.int64, device="cuda")
    out = finder.run(torch.from_numpy(fods).cuda(), stats=stats)
    torch.cuda.synchronize()
    ref, it = orc.sh2peaks(fods, nthreads=8)
    frac, worst = compare(out.cpu().numpy(), ref, 3, fods)
    st = stats.cpu().numpy()
    say("[gpu peaks] agree %.5f worst angle %.2e rad; stats voxels %d iterations %d (oracle %d) failed %d"
        % (frac, worst, st[0], st[1], it, st[2]))
    assert frac >= 0.999
    assert st[0] == len(fods) and st[3] == 0
    assert abs(int(st[1]) - it) <= 1e-3 * it


@pytest.mark.parametrize("lmax", [2, 4, 6, 10, 12])
def test_other_lmax(torch_cuda, lmax):
    torch = torch_cuda
    rng = np.random.default_rng(lmax)
    nsh = (lmax + 1) * (lmax + 2) // 2
    # random sums of three lobes with a little noise
    dirs = rng.standard_normal((300, 3, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    ang = np.stack([np.arctan2(dirs[..., 1], dirs[..., 0]), np.arccos(np.clip(dirs[..., 2], -1, 1))], axis=-1)
    Y = orc.sh_transform(ang.reshape(-1, 2), lmax).reshape(300, 3, nsh)
    w = rng.uniform(0.3, 1.0, (300, 3, 1))
    sh = ((w * Y).sum(axis=1) + 0.01 * rng.standard_normal((300, nsh))).astype(np.float32)
    say("[gpu peaks] lmax %d, 300 voxels" % lmax)
    out = peaks.SH2Peaks(nsh).run(torch.from_numpy(sh).cuda())
    torch.cuda.synchronize()
    ref, _ = orc.sh2peaks(sh, lmax=lmax, nthreads=8)
    frac, worst = compare(out.cpu().numpy(), ref, 3, sh)
    say("[gpu peaks] lmax %d agree %.5f worst angle %.2e" % (lmax, frac, worst))
    assert frac >= 0.99


@pytest.mark.parametrize("lmax", [10, 12])
def test_fast_table_other_lmax_and_infinite_coefficients(torch_cuda, lmax):
    """-fast at lmax 10 and 12 against the oracle's -fast run, with voxels whose series holds an
    infinite coefficient: their search directions become NaN, whose elevation takes the table's
    first row as the reference's conversion does on x86, and the outputs agree with the oracle's."""
    torch = torch_cuda
    rng = np.random.default_rng(40 + lmax)
    nsh = (lmax + 1) * (lmax + 2) // 2
    dirs = rng.standard_normal((400, 3, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    ang = np.stack([np.arctan2(dirs[..., 1], dirs[..., 0]), np.arccos(np.clip(dirs[..., 2], -1, 1))], axis=-1)
    Y = orc.sh_transform(ang.reshape(-1, 2), lmax).reshape(400, 3, nsh)
    w = rng.uniform(0.3, 1.0, (400, 3, 1))
    sh = ((w * Y).sum(axis=1) + 0.01 * rng.standard_normal((400, nsh))).astype(np.float32)
    sh[7, 5] = np.inf
    sh[8, 0] = -np.inf
    sh[9, nsh - 1] = np.inf
    say("[gpu peaks] -fast lmax %d, 400 voxels (3 with an infinite coefficient)" % lmax)
    out = peaks.SH2Peaks(nsh, fast=True).run(torch.from_numpy(sh).cuda())
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    ref, _ = orc.sh2peaks(sh, lmax=lmax, nthreads=8, fast=True)
    keep = np.ones(400, dtype=bool)
    keep[[7, 8, 9]] = False
    frac, worst = compare(out[keep], ref[keep], 3, sh[keep])
    say("[gpu peaks] -fast lmax %d agree %.5f worst angle %.2e" % (lmax, frac, worst))
    assert frac >= 0.99
    for v in (7, 8, 9):
        assert np.array_equal(np.isnan(out[v]), np.isnan(ref[v])), v
        assert np.allclose(out[v], ref[v], equal_nan=True), v


def test_options_and_input_rules(torch_cuda, fods):
    torch = torch_cuda
    sub = fods[:512].copy()
    sub[3, 10] = np.nan                 # NaN coefficient -> all NaN
    sub[4, 1:] = 0.0                    # isotropic -> all NaN
    mask = np.ones(len(sub), dtype=np.uint8)
    mask[5] = 0
    d_sub = torch.from_numpy(sub).cuda()
    d_mask = torch.from_numpy(mask).cuda()
    rng = np.random.default_rng(8)

    # -num 5 -threshold, mask
    thr = float(np.quantile(np.linalg.norm(orc.sh2peaks(sub, npeaks=1)[0][np.isfinite(orc.sh2peaks(sub, npeaks=1)[0][:, 0])], axis=1), 0.2))
    say("[gpu peaks] -num 5 -threshold %.3f with a mask" % thr)
    g = peaks.SH2Peaks(45, npeaks=5, threshold=thr).run(d_sub, mask=d_mask).cpu().numpy()
    r, _ = orc.sh2peaks(sub, npeaks=5, threshold=thr, mask=mask)
    assert np.all(np.isnan(g[3])) and np.all(np.isnan(g[4])) and np.all(np.isnan(g[5]))
    assert compare(g, r, 5, sub)[0] >= 0.998

    # custom seeds: 129 random directions (three batches of 64)
    seeds = np.stack([rng.uniform(-np.pi, np.pi, 129), np.arccos(rng.uniform(-1, 1, 129))], axis=1).astype(np.float32)
    say("[gpu peaks] -seeds (129 directions)")
    g = peaks.SH2Peaks(45, seeds=seeds).run(d_sub).cpu().numpy()
    r, _ = orc.sh2peaks(sub, seeds=seeds)
    assert compare(g, r, 3, sub)[0] >= 0.998

    # -direction: two fixed directions for every voxel
    dirs = np.array([[0.3, 1.1], [1.9, 0.7]], dtype=np.float32)
    unit = np.stack([np.cos(dirs[:, 0]) * np.sin(dirs[:, 1]), np.sin(dirs[:, 0]) * np.sin(dirs[:, 1]),
                     np.cos(dirs[:, 1])], axis=1).astype(np.float32)
    say("[gpu peaks] -direction x2")
    f = peaks.SH2Peaks(45, directions=dirs)
    assert f.npeaks == 2
    g = f.run(d_sub).cpu().numpy()
    r, _ = orc.sh2peaks(sub, directions=unit)
    assert compare(g, r, 2, sub)[0] >= 0.998

    # -peaks: per-voxel directions (the oracle's own peaks, perturbed); a NaN row -> all NaN
    base, _ = orc.sh2peaks(sub, npeaks=2)
    pin = (base + 0.05 * rng.standard_normal(base.shape)).astype(np.float32)
    pin[7, 0] = np.nan
    say("[gpu peaks] -peaks image")
    g = peaks.SH2Peaks(45, npeaks=2).run(d_sub, peaks=torch.from_numpy(pin).cuda()).cpu().numpy()
    r, _ = orc.sh2peaks(sub, peaks_image=pin)
    assert np.all(np.isnan(g[7]))
    assert compare(g, r, 2, sub)[0] >= 0.998


def test_host_pointer_entry_and_empty_batch(torch_cuda, fods):
    f = peaks.SH2Peaks(45)
    say("[gpu peaks] host-pointer entry point")
    g = f.run(fods[:300])
    r, _ = orc.sh2peaks(fods[:300])
    assert compare(g, r, 3, fods[:300])[0] >= 0.99
    assert f.run(np.zeros((0, 45), dtype=np.float32)).shape == (0, 9)


def test_brain_volume_properties(torch_cuda):
    """646,128 voxels of whole-brain FODs from the GPU CSD: every voxel gets at least one peak, no
    search exhausts its capacity, and a 2,000-voxel sample matches the oracle."""
    torch = torch_cuda
    from mrtrix3_amd import csd
    ph = phantom.single_shell_brain()
    solver = csd.CSD(csd.CSDShared(ph.grad, ph.response))
    say("[gpu peaks] brain: CSD on %d voxels" % ph.nvox)
    fod, _, _ = solver.run(torch.from_numpy(ph.dw).cuda())
    stats = torch.zeros(4, dtype=torch.int64, device="cuda")
    say("[gpu peaks] brain: sh2peaks")
    out = peaks.SH2Peaks(45).run(fod, stats=stats)
    torch.cuda.synchronize()
    st = stats.cpu().numpy()
    o = out.cpu().numpy()
    say("[gpu peaks] brain stats", st.tolist())
    assert st[0] == ph.nvox and st[3] == 0
    assert np.isfinite(o[:, 0]).mean() >= 0.999
    idx = np.sort(np.random.default_rng(9).choice(ph.nvox, 2000, replace=False))
    r, _ = orc.sh2peaks(fod.cpu().numpy()[idx], nthreads=8)
    frac, worst = compare(o[idx], r, 3, fod.cpu().numpy()[idx])
    say("[gpu peaks] brain sample agree %.5f worst angle %.2e" % (frac, worst))
    assert frac >= 0.999


def test_two_passes_many_seeds(torch_cuda, fods):
    """512 search directions on 163,840 voxels (the 4,096 FODs tiled 40 times): 84M searches, more
    than one pass of the work buffer holds, so the volume runs in two passes.  Every tile gives
    the same output bit for bit (a voxel's searches do not depend on where it sits or which pass
    runs it), and a sample of the first tile matches the oracle."""
    torch = torch_cuda
    rng = np.random.default_rng(11)
    seeds = np.stack([rng.uniform(-np.pi, np.pi, 512), np.arccos(rng.uniform(-1, 1, 512))], axis=1).astype(np.float32)
    tiled = np.tile(fods, (40, 1))
    say("[gpu peaks] 512 seeds on %d voxels (two passes)" % len(tiled))
    stats = torch.zeros(4, dtype=torch.int64, device="cuda")
    out = peaks.SH2Peaks(45, seeds=seeds).run(torch.from_numpy(tiled).cuda(), stats=stats).cpu().numpy()
    st = stats.cpu().numpy()
    assert st[0] == len(tiled) and st[3] == 0
    ref_tile = out[:len(fods)]
    for k in range(1, 40):
        assert np.array_equal(out[k * len(fods):(k + 1) * len(fods)], ref_tile, equal_nan=True), k
    r, _ = orc.sh2peaks(fods[:256], seeds=seeds, nthreads=8)
    frac, worst = compare(ref_tile[:256], r, 3, fods[:256])
    say("[gpu peaks] 512 seeds: tiles identical; oracle agree %.5f worst angle %.2e" % (frac, worst))
    assert frac >= 0.996


def test_fast_table_matches_oracle(torch_cuda, fods):
    """-fast (Legendre values from the 512-elevation table, SH.h:402-470) against the oracle's run
    with the same table, at the exact path's tolerance."""
    torch = torch_cuda
    finder = peaks.SH2Peaks(45, fast=True)
    say("[gpu peaks] -fast: %d voxels, lmax 8" % len(fods))
    out = finder.run(torch.from_numpy(fods).cuda())
    torch.cuda.synchronize()
    ref, _ = orc.sh2peaks(fods, nthreads=8, fast=True)
    exact, _ = orc.sh2peaks(fods[:2000], nthreads=8)
    frac, worst = compare(out.cpu().numpy(), ref, 3, fods)
    say("[gpu peaks] -fast agree %.5f worst angle %.2e rad" % (frac, worst))
    assert frac >= 0.999
    assert not np.array_equal(ref[:2000], exact)    # the table path, not the exact one
