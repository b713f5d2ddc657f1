"""The dwi2fod drop-in end to end on the GPU: .mif image and response text files in, the
reference's command line, .mif FOD images out, compared voxel by voxel with the CPU oracle
(cmd/dwi2fod.cpp: CSD_Processor :120-176 and MSMT_Processor :181-231 semantics, including the
mask and the header keys written)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from mrtrix3_amd import icls, mif, phantom
from oracle import orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mrtrix3_amd", "bin", "dwi2fod")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(args):
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return r


def write_response(path, rows):
    with open(path, "w") as f:
        for row in np.atleast_2d(rows):
            f.write(" ".join(repr(float(v)) for v in row) + "\n")


def test_cli_csd_matches_oracle(tmp_path):
    assert os.path.exists(EXE), "dwi2fod is not built"
    ph = phantom.single_shell_small()
    shape = (16, 16, 8)
    n = int(np.prod(shape))
    img = ph.dw[:n].reshape(shape + (ph.dw.shape[1],))
    rng = np.random.default_rng(5)
    mask = (rng.random(shape) < 0.8).astype(np.float32)
    dwi, msk, resp, out = (str(tmp_path / f) for f in ("dwi.mif", "mask.mif", "resp.txt", "fod.mif"))
    mif.write(dwi, img, vox=[2.0, 2.0, 2.0], keyval=[("dw_scheme", mif.scheme_text(ph.grad))])
    mif.write(msk, mask, vox=[2.0, 2.0, 2.0])
    write_response(resp, ph.response)
    say("[gpu cli] dwi2fod csd on %s voxels (%d in mask)" % (n, int(mask.sum())))
    run(["csd", dwi, resp, out, "-mask", msk, "-force"])
    fod, h = mif.read(out)
    assert fod.shape == shape + (45,)
    keys = [k for k, _ in h["keyval"]]
    assert "prior_dw_scheme" in keys and "dw_scheme" not in keys
    o = orc.CsdOracle(ph.grad, ph.response)
    rel = []
    for idx in np.ndindex(shape):
        if mask[idx] == 0:
            assert np.all(fod[idx] == 0)
            continue
        b = np.maximum(img[idx][o.vols].astype(np.float64), 0.0)
        F, _, _, _ = o.solve(b)
        rel.append(np.linalg.norm(fod[idx] - F) / max(np.linalg.norm(F), 1e-30))
    rel = np.array(rel)
    say("[gpu cli] csd: max rel %.3e over %d voxels" % (rel.max(), rel.size))
    assert np.mean(rel <= 1e-4) >= 0.999


def test_cli_msmt_matches_oracle(tmp_path):
    assert os.path.exists(EXE), "dwi2fod is not built"
    shape = (8, 8, 4)
    ph = phantom.multi_shell(seed=4, shape=shape, semi=None)
    img = ph.dw.reshape(shape + (ph.dw.shape[1],))
    dwi = str(tmp_path / "dwi.mif")
    mif.write(dwi, img, vox=[2.0, 2.0, 2.0], keyval=[("dw_scheme", mif.scheme_text(ph.grad))])
    args = ["msmt_csd", dwi]
    outs = []
    for name, r in zip(("wm", "gm", "csf"), ph.response):
        rp, op = str(tmp_path / (name + ".txt")), str(tmp_path / (name + ".mif"))
        write_response(rp, r)
        args += [rp, op]
        outs.append(op)
    say("[gpu cli] dwi2fod msmt_csd on %d voxels" % ph.nvox)
    pred_path = str(tmp_path / "pred.mif")
    run(args + ["-force", "-predicted_signal", pred_path])
    got = [mif.read(p)[0] for p in outs]
    assert got[0].shape == shape + (45,) and got[1].shape[:3] == shape and got[2].shape[:3] == shape
    s = icls.MSMTShared(ph.grad, ph.response)
    o = orc.MsmtOracle(ph.grad, ph.response)
    rel = []
    for idx in np.ndindex(shape):
        x, _, _ = o.solve(img[idx].astype(np.float64))
        g = np.concatenate([np.atleast_1d(got[0][idx]), np.atleast_1d(got[1][idx]).ravel(),
                            np.atleast_1d(got[2][idx]).ravel()])
        rel.append(np.linalg.norm(g - x) / max(np.linalg.norm(x), 1e-30))
    rel = np.array(rel)
    say("[gpu cli] msmt_csd: max rel %.3e over %d voxels" % (rel.max(), rel.size))
    assert np.mean(rel <= 1e-4) >= 0.999
    # -predicted_signal (dwi2fod.cpp:216-220): H x per voxel, H the problem matrix (here the oracle's
    # own), against the command's written solution and against the oracle's solution
    pred, hp = mif.read(pred_path)
    assert pred.shape == img.shape
    xs = np.concatenate([got[0].reshape(-1, 45), got[1].reshape(-1, 1), got[2].reshape(-1, 1)], axis=1)
    want = xs.astype(np.float64) @ o.C.T
    p2 = pred.reshape(-1, img.shape[3]).astype(np.float64)
    scale = np.abs(want).max(axis=1, keepdims=True)
    err = np.abs(p2 - want) / scale
    worst_o = 0.0
    for k, idx in enumerate(np.ndindex(shape)):
        x, _, _ = o.solve(img[idx].astype(np.float64))
        worst_o = max(worst_o, float(np.abs(p2[k] - o.C @ x).max() / scale[k, 0]))
    say("[gpu cli] -predicted_signal: %.2e of the signal scale from H x of the written solution, %.2e from the "
        "oracle's" % (err.max(), worst_o))
    assert err.max() <= 1e-6 and worst_o <= 1e-4


WARM_NET = os.path.join(ROOT, "tests", "golden", "warm_net_single_shell_brain.npz")


def test_cli_warm_net(tmp_path):
    """dwi2fod csd -warm_net: 16,384 voxels of the whole-brain protocol (8 b=0 + 64 b=3000) from the
    network trained for it; the same sets and coefficients within 1e-4 as the plain command in every
    voxel, fewer solves, and the plain path (with a warning) where the certificate does not hold."""
    ph = phantom.single_shell_brain()
    shape = (32, 32, 16)
    n = int(np.prod(shape))
    img = ph.dw[:n].reshape(shape + (ph.dw.shape[1],))
    dwi, resp = str(tmp_path / "dwi.mif"), str(tmp_path / "resp.txt")
    mif.write(dwi, img, vox=[2.0, 2.0, 2.0], keyval=[("dw_scheme", mif.scheme_text(ph.grad))])
    write_response(resp, ph.response)
    plain, warm = str(tmp_path / "plain.mif"), str(tmp_path / "warm.mif")
    say("[gpu cli] dwi2fod csd -warm_net on %d voxels" % n)
    run(["csd", dwi, resp, plain, "-force"])
    r = run(["csd", dwi, resp, warm, "-force", "-warm_net", WARM_NET, "-info"])
    info = [l for l in r.stderr.splitlines() if "warm start:" in l]
    say("[gpu cli] " + info[0])
    fp, fw = mif.read(plain)[0].reshape(n, 45).astype(np.float64), mif.read(warm)[0].reshape(n, 45).astype(np.float64)
    rel = np.linalg.norm(fw - fp, axis=1) / np.maximum(np.linalg.norm(fp, axis=1), 1e-30)
    say("[gpu cli] warm vs plain: max rel %.2e" % rel.max())
    assert rel.max() <= 1e-4
    solves, voxels = int(info[0].split(" solves in all")[0].split(", ")[-1]), n
    assert solves <= 2.7 * voxels
    # a threshold voids the certificate: plain path, same output as the plain command
    thr_p, thr_w = str(tmp_path / "thr_p.mif"), str(tmp_path / "thr_w.mif")
    run(["csd", dwi, resp, thr_p, "-force", "-threshold", "0.1"])
    r = run(["csd", dwi, resp, thr_w, "-force", "-threshold", "0.1", "-warm_net", WARM_NET])
    assert "-warm_net ignored" in r.stderr
    assert np.array_equal(mif.read(thr_p)[0], mif.read(thr_w)[0])
    # a network trained for another protocol is refused
    small = phantom.single_shell_small()
    dwi2 = str(tmp_path / "dwi60.mif")
    mif.write(dwi2, small.dw[:512].reshape(8, 8, 8, -1), vox=[2.0, 2.0, 2.0],
              keyval=[("dw_scheme", mif.scheme_text(small.grad))])
    resp2 = str(tmp_path / "resp60.txt")
    write_response(resp2, small.response)
    r = subprocess.run([EXE, "csd", dwi2, resp2, str(tmp_path / "x.mif"), "-warm_net", WARM_NET],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode != 0 and "trained for 64 DW volumes" in r.stderr


def count_peak_mismatches(g, ref, fod):
    """Voxels whose reported peaks differ from the oracle's: a different set of reported peaks, or a
    peak more than 1e-3 rad off or with an amplitude more than 1e-4 of the voxel's FOD scale off."""
    bad = 0
    for v in range(g.shape[0]):
        gv, rv = g[v].reshape(-1, 3), ref[v].reshape(-1, 3)
        if not np.array_equal(np.isfinite(gv[:, 0]), np.isfinite(rv[:, 0])):
            bad += 1
            continue
        for i in np.nonzero(np.isfinite(rv[:, 0]))[0]:
            c = abs(gv[i] @ rv[i]) / (np.linalg.norm(gv[i]) * np.linalg.norm(rv[i]))
            scale = max(np.linalg.norm(rv[i]), 0.28 * np.linalg.norm(fod[v]))   # the voxel's FOD scale
            if np.arccos(min(1.0, c)) > 1e-3 or abs(np.linalg.norm(gv[i]) - np.linalg.norm(rv[i])) > 1e-4 * scale:
                bad += 1
                break
    return bad


def test_cli_sh2peaks_matches_oracle(tmp_path):
    """sh2peaks SH peaks -num 2 -mask: .mif in and out (cmd/sh2peaks.cpp), against the oracle."""
    exe = os.path.join(ROOT, "mrtrix3_amd", "bin", "sh2peaks")
    assert os.path.exists(exe), "sh2peaks is not built"
    ph = phantom.single_shell_small()
    shape = (16, 16, 8)
    n = int(np.prod(shape))
    o = orc.CsdOracle(ph.grad, ph.response)
    fod, _ = o.run_volume(np.ascontiguousarray(ph.dw[:n]), (16, 16, 8), nthreads=8)
    rng = np.random.default_rng(6)
    mask = (rng.random(shape[::-1]) < 0.9).astype(np.float32)          # [z, y, x]
    img = fod.reshape(shape[2], shape[1], shape[0], 45).transpose(2, 1, 0, 3)   # [x, y, z, 45]
    shp, msk, out = (str(tmp_path / f) for f in ("fod.mif", "mask.mif", "peaks.mif"))
    mif.write(shp, img, vox=[2.0, 2.0, 2.0])
    mif.write(msk, mask.transpose(2, 1, 0), vox=[2.0, 2.0, 2.0])
    say("[gpu cli] sh2peaks on %d voxels" % n)
    r = subprocess.run([exe, shp, out, "-num", "2", "-mask", msk, "-force"], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    got, h = mif.read(out)
    assert got.shape == shape + (6,)
    ref, _ = orc.sh2peaks(fod, npeaks=2, mask=mask.reshape(-1).astype(np.uint8), nthreads=8)
    g = got.transpose(2, 1, 0, 3).reshape(n, 6)
    bad = count_peak_mismatches(g, ref, fod)
    say("[gpu cli] sh2peaks: %d of %d voxels differ" % (bad, n))
    assert bad <= 0.001 * n + 1
    # -peaks: the first run's output as the directions to match (cmd/sh2peaks.cpp:160-170)
    out2 = str(tmp_path / "peaks2.mif")
    r = subprocess.run([exe, shp, out2, "-peaks", out, "-mask", msk, "-force"], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    got2, _ = mif.read(out2)
    assert got2.shape == shape + (6,)
    g2 = got2.transpose(2, 1, 0, 3).reshape(n, 6)
    ref2, _ = orc.sh2peaks(fod, npeaks=2, peaks_image=np.ascontiguousarray(g, dtype=np.float32),
                           mask=mask.reshape(-1).astype(np.uint8), nthreads=8)
    bad2 = count_peak_mismatches(g2, ref2, fod)
    say("[gpu cli] sh2peaks -peaks: %d of %d voxels differ" % (bad2, n))
    assert bad2 <= 0.001 * n + 1
    assert np.all(np.isnan(g2[mask.reshape(-1) == 0]))
    assert np.all(np.isnan(g[mask.reshape(-1) == 0]))
    # -fast: the Legendre lookup table (SH.h:402-470)
    out3 = str(tmp_path / "peaks3.mif")
    r = subprocess.run([exe, shp, out3, "-num", "2", "-mask", msk, "-fast", "-force"], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    got3, _ = mif.read(out3)
    g3 = got3.transpose(2, 1, 0, 3).reshape(n, 6)
    ref3, _ = orc.sh2peaks(fod, npeaks=2, mask=mask.reshape(-1).astype(np.uint8), nthreads=8, fast=True)
    bad3 = count_peak_mismatches(g3, ref3, fod)
    say("[gpu cli] sh2peaks -fast: %d of %d voxels differ" % (bad3, n))
    assert bad3 <= 0.001 * n + 1
    assert not np.array_equal(g3, g, equal_nan=True)


def test_cli_fod2fixel_matches_oracle(tmp_path):
    """fod2fixel FOD fixels -afd -peak -disp -mask -fmls_no_thresholds (the call of dwi2response
    tournier, -peak standing for -peak_amp): the fixel directory's index, directions and data
    files against the oracle's segmentation of the same float32 FODs, in voxel order; then -maxnum
    and -dirpeak."""
    exe = os.path.join(ROOT, "mrtrix3_amd", "bin", "fod2fixel")
    assert os.path.exists(exe), "fod2fixel is not built"
    ph = phantom.single_shell_small()
    shape = (16, 16, 8)
    n = int(np.prod(shape))
    o = orc.CsdOracle(ph.grad, ph.response)
    fod, _ = o.run_volume(np.ascontiguousarray(ph.dw[:n]), shape, nthreads=8)
    fod = np.ascontiguousarray(fod, dtype=np.float32)
    rng = np.random.default_rng(8)
    mask = (rng.random(shape[::-1]) < 0.85).astype(np.float32)          # [z, y, x]
    img = fod.reshape(shape[2], shape[1], shape[0], 45).transpose(2, 1, 0, 3)
    fp, mp = str(tmp_path / "fod.mif"), str(tmp_path / "mask.mif")
    mif.write(fp, img, vox=[2.0, 2.0, 2.0])
    mif.write(mp, mask.transpose(2, 1, 0), vox=[2.0, 2.0, 2.0])
    m = mask.reshape(-1)

    def segment(args, out, oracle, maxnum=0, dirpeak=False):
        say("[gpu cli] fod2fixel %s on %d voxels" % (" ".join(args), n))
        r = subprocess.run([exe, fp, out] + args, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr
        idx, ih = mif.read(os.path.join(out, "index.mif"))
        assert ih["datatype"].lower().startswith("uint32") and idx.shape == shape + (2,)
        count = idx[..., 0].transpose(2, 1, 0).reshape(-1).astype(np.int64)
        offset = idx[..., 1].transpose(2, 1, 0).reshape(-1).astype(np.int64)
        nf = int(dict(ih["keyval"])["nfixels"])
        dirs, dh = mif.read(os.path.join(out, "directions.mif"))
        assert dirs.shape == (nf, 3, 1) and nf == count.sum()
        run = 0
        for v in range(n):
            lobes = np.zeros((0, 8))
            if m[v]:
                lobes, _ = oracle.segment(fod[v].astype(np.float64), max_out=2048)
                if maxnum:
                    lobes = lobes[:maxnum]
            assert count[v] == len(lobes), "voxel %d" % v
            assert offset[v] == (run if count[v] else 0)
            want = (lobes[:, 3:6] if dirpeak else lobes[:, 0:3]).astype(np.float32)
            np.testing.assert_allclose(dirs[run:run + count[v], :, 0], want, rtol=2e-6, atol=2e-7)
            run += count[v]
        return nf

    out = str(tmp_path / "fixels")
    ref = orc.FmlsOracle(8, integral=0.0, peak_value=0.0)
    nf = segment(["-afd", "afd.mif", "-peak", "peak.mif", "-disp", "disp.mif", "-mask", mp, "-fmls_no_thresholds"],
                 out, ref)
    lob = [ref.segment(fod[v].astype(np.float64), max_out=2048)[0] for v in range(n) if m[v]]
    lob = np.concatenate([x for x in lob if len(x)]) if lob else np.zeros((0, 8))
    afd, _ = mif.read(os.path.join(out, "afd.mif"))
    peak, _ = mif.read(os.path.join(out, "peak.mif"))
    disp, _ = mif.read(os.path.join(out, "disp.mif"))
    assert afd.shape == (nf, 1, 1) and peak.shape == (nf, 1, 1) and disp.shape == (nf, 1, 1)
    np.testing.assert_allclose(afd[:, 0, 0], lob[:, 6].astype(np.float32), rtol=2e-6, atol=2e-7)
    np.testing.assert_allclose(peak[:, 0, 0], lob[:, 7].astype(np.float32), rtol=2e-6, atol=2e-7)
    np.testing.assert_allclose(disp[:, 0, 0], lob[:, 6].astype(np.float32) / lob[:, 7].astype(np.float32),
                               rtol=4e-6, atol=2e-7)
    say("[gpu cli] fod2fixel: %d fixels in %d masked voxels" % (nf, int(m.sum())))
    # defaults with -maxnum 2 -dirpeak, no mask
    m = np.ones(n)
    segment(["-maxnum", "2", "-dirpeak"], str(tmp_path / "fixels2"), orc.FmlsOracle(8), maxnum=2, dirpeak=True)


def test_cli_fslgrad_matches_dw_scheme(tmp_path):
    """-fslgrad bvecs bvals (core/dwi/gradient.cpp:127-200) on an image whose transform is a
    rotation: the bvecs are the scheme's directions in image axes with the first component negated
    (positive determinant), one b=0 row has NaN directions.  The FODs equal those of the same
    image with the scheme in its header."""
    assert os.path.exists(EXE), "dwi2fod is not built"
    ph = phantom.single_shell_small()
    shape = (8, 8, 4)
    n = int(np.prod(shape))
    img = ph.dw[:n].reshape(shape + (ph.dw.shape[1],))
    a, b = 0.4, -0.3
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]) @ \
        np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
    T = np.concatenate([R, [[10.0], [-5.0], [3.0]]], axis=1)
    g = ph.grad
    bvecs = R.T @ g[:, :3].T
    bvecs[0] *= -1.0
    b0 = np.nonzero(g[:, 3] == 0)[0]
    assert len(b0) > 0
    bvecs[:, b0[0]] = np.nan
    d1, d2, resp = (str(tmp_path / f) for f in ("hdr.mif", "plain.mif", "resp.txt"))
    fv, fb = str(tmp_path / "bvecs"), str(tmp_path / "bvals")
    mif.write(d1, img, vox=[2.0, 2.0, 2.0], transform=T.tolist(), keyval=[("dw_scheme", mif.scheme_text(g))])
    mif.write(d2, img, vox=[2.0, 2.0, 2.0], transform=T.tolist())
    with open(fv, "w") as f:
        f.write("\n".join(" ".join(repr(float(v)) for v in row) for row in bvecs) + "\n")
    with open(fb, "w") as f:
        f.write("\n".join(repr(float(v)) for v in g[:, 3]) + "\n")       # one column
    write_response(resp, ph.response)
    say("[gpu cli] dwi2fod csd -fslgrad vs dw_scheme on %d voxels" % n)
    run(["csd", d1, resp, str(tmp_path / "a.mif"), "-force"])
    r = run(["csd", d2, resp, str(tmp_path / "b.mif"), "-fslgrad", fv, fb, "-force"])
    assert "1 row with NaN values detected in bvecs file" in r.stderr
    fa, _ = mif.read(str(tmp_path / "a.mif"))
    fb_, _ = mif.read(str(tmp_path / "b.mif"))
    rel = np.linalg.norm((fa - fb_).reshape(n, -1), axis=1) / np.maximum(np.linalg.norm(fa.reshape(n, -1), axis=1), 1e-30)
    say("[gpu cli] -fslgrad: max rel %.3e" % rel.max())
    assert np.mean(rel <= 1e-5) >= 0.999


def test_cli_amp2response_matches_oracle(tmp_path):
    """amp2response amps mask directions response (cmd/amp2response.cpp) on the small phantom's
    two shells (b=0 at lmax 0, b=3000 at lmax 10 by default), against the oracle's sums and solves;
    then -noconstraint."""
    from oracle import resp_oracle as R
    exe = os.path.join(ROOT, "mrtrix3_amd", "bin", "amp2response")
    assert os.path.exists(exe), "amp2response is not built"
    ph = phantom.single_shell_small()
    shape = (16, 16, 8)
    n = int(np.prod(shape))
    rng = np.random.default_rng(12)
    amp = ph.dw[:n].astype(np.float32)
    fib = rng.standard_normal((n, 3)).astype(np.float32)
    mask = (rng.random(n) < 0.3).astype(np.float32)
    paths = [str(tmp_path / f) for f in ("amps.mif", "mask.mif", "dirs.mif", "resp.txt", "resp_ols.txt")]
    mif.write(paths[0], amp.reshape(shape + (amp.shape[1],)), keyval=[("dw_scheme", mif.scheme_text(ph.grad))])
    mif.write(paths[1], mask.reshape(shape))
    mif.write(paths[2], fib.reshape(shape + (3,)))
    say("[gpu cli] amp2response on %d voxels (%d in mask)" % (n, int(mask.sum())))
    r = subprocess.run([exe] + paths[:4] + ["-force"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    r2 = subprocess.run([exe] + paths[:3] + [paths[4], "-noconstraint", "-force"], capture_output=True, text=True, timeout=240)
    assert r2.returncode == 0, r2.stderr
    got = np.loadtxt(paths[3], ndmin=2)
    got_ols = np.loadtxt(paths[4], ndmin=2)
    assert got.shape == (2, 6)
    b = ph.grad[:, 3]
    m = mask.astype(np.uint8)
    for row, (sel, lmax) in enumerate([(b <= 10, 0), (b > 10, 10)]):
        vols = np.nonzero(sel)[0]
        g = ph.grad[vols, :3]
        nrm = np.linalg.norm(g, axis=1, keepdims=True)
        dirs = np.where(nrm > 0, g / np.where(nrm > 0, nrm, 1), np.array([0.0, 0.0, 1.0]))
        M, bb, _ = R.accumulate(amp, fib, m, dirs, vols, lmax)
        for out, constrained in ((got, True), (got_ols, False)):
            ref, _ = R.solve(lmax, M, bb, constrained)
            k = lmax // 2 + 1
            err = np.abs(out[row, :k] - ref).max() / np.abs(ref).max()
            say("[gpu cli] amp2response shell %d constrained %s: rel %.2e" % (row, constrained, err))
            assert err <= 1e-5 and np.all(out[row, k:] == 0)


def test_cli_shells_option(tmp_path):
    """-shells (DWI::ShellsOption): csd on the b=3000 shell of a four-shell image equals the oracle
    on the image reduced to that shell; msmt_csd -shells 0,1000,3000 with three-row responses
    equals the oracle's ICLS on the problem whose b=2000 rows are zero (three tissues need three
    shells: with two, C^T C is singular but for the regularisation)."""
    shape = (8, 8, 4)
    ph = phantom.multi_shell(seed=4, shape=shape, semi=None)
    img = ph.dw.reshape(shape + (ph.dw.shape[1],))
    dwi = str(tmp_path / "dwi.mif")
    mif.write(dwi, img, vox=[2.0, 2.0, 2.0], keyval=[("dw_scheme", mif.scheme_text(ph.grad))])
    b = ph.grad[:, 3]
    # csd on b=3000
    resp = str(tmp_path / "wm3000.txt")
    write_response(resp, ph.response[0][3])
    out = str(tmp_path / "fod.mif")
    say("[gpu cli] dwi2fod csd -shells 3000 on %d voxels" % ph.nvox)
    run(["csd", dwi, resp, out, "-shells", "3000", "-force"])
    fod, _ = mif.read(out)
    sel = np.nonzero((b < 10) | (np.abs(b - 3000) < 100))[0]
    o = orc.CsdOracle(ph.grad[sel], ph.response[0][3])
    rel = []
    for idx in np.ndindex(shape):
        sig = np.maximum(img[idx][sel][o.vols].astype(np.float64), 0.0)
        F, _, _, _ = o.solve(sig)
        rel.append(np.linalg.norm(fod[idx] - F) / max(np.linalg.norm(F), 1e-30))
    rel = np.array(rel)
    say("[gpu cli] csd -shells: max rel %.3e" % rel.max())
    assert np.mean(rel <= 1e-4) >= 0.999
    # msmt_csd on b=0, 1000 and 3000
    args = ["msmt_csd", dwi]
    two = [r[[0, 1, 3]] for r in ph.response]
    outs = []
    for name, r in zip(("wm", "gm", "csf"), two):
        rp, op = str(tmp_path / (name + "2.txt")), str(tmp_path / (name + "2.mif"))
        write_response(rp, r)
        args += [rp, op]
        outs.append(op)
    say("[gpu cli] dwi2fod msmt_csd -shells 0,1000,3000")
    run(args + ["-shells", "0,1000,3000", "-force"])
    got = [mif.read(p)[0] for p in outs]
    s = icls.MSMTShared(ph.grad, two, shells=[0, 1000, 3000])
    assert s.nshell == 3
    o2 = orc.MsmtOracle(ph.grad, two, shells=[0, 1000, 3000])
    rel = []
    for idx in np.ndindex(shape):
        x, _, _ = o2.solve(img[idx].astype(np.float64))
        g = np.concatenate([np.atleast_1d(got[0][idx]), np.atleast_1d(got[1][idx]).ravel(),
                            np.atleast_1d(got[2][idx]).ravel()])
        rel.append(np.linalg.norm(g - x) / max(np.linalg.norm(x), 1e-30))
    rel = np.array(rel)
    say("[gpu cli] msmt_csd -shells: max rel %.3e" % rel.max())
    assert np.mean(rel <= 1e-4) >= 0.999


def test_cli_nifti_input_matches_mif(tmp_path):
    """NIfTI-1 input (core/formats/nifti1*.cpp, core/file/nifti_utils.cpp): the phantom as int16
    with scl_slope / scl_inter, in a .nii.gz with an s-form and a .nii with only a q-form, its
    gradients as FSL bvecs / bvals in image axes, and a uint8 NIfTI mask, give the FODs of the same
    values in a .mif with the scheme in its header."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import nifti_util
    ph = phantom.single_shell_small()
    shape = (16, 16, 8)
    n = int(np.prod(shape))
    rng = np.random.default_rng(13)
    dw = ph.dw[:n].reshape(shape + (ph.dw.shape[1],))
    slope, inter = np.float32(np.abs(dw).max() / 30000.0), np.float32(0.5)
    raw = np.round((dw - inter) / slope).astype(np.int16)
    deq = (np.float64(slope) * raw.astype(np.float64) + np.float64(inter)).astype(np.float32)
    mask = (rng.random(shape) < 0.7).astype(np.uint8)
    a, c = 0.3, -0.2
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]) @ \
        np.array([[np.cos(c), 0, np.sin(c)], [0, 1, 0], [-np.sin(c), 0, np.cos(c)]])
    vox = np.array([2.0, 2.0, 2.5])
    t = np.array([-30.0, 12.0, 5.0])
    g = ph.grad
    bvecs = R.T @ g[:, :3].T
    bvecs[0] *= -1.0                                     # positive determinant: FSL's flipped x
    fv, fb, resp = (str(tmp_path / f) for f in ("bvecs", "bvals", "resp.txt"))
    np.savetxt(fv, bvecs, fmt="%.17g")
    np.savetxt(fb, g[:, 3][None, :], fmt="%.17g")
    write_response(resp, ph.response)
    # reference run: .mif with the scheme in its header
    T = np.concatenate([R, t[:, None]], axis=1)
    dmif, mmif = str(tmp_path / "dwi.mif"), str(tmp_path / "mask.mif")
    mif.write(dmif, deq, vox=list(vox), transform=T.tolist(), keyval=[("dw_scheme", mif.scheme_text(g))])
    mif.write(mmif, mask.astype(np.float32), vox=list(vox), transform=T.tolist())
    run(["csd", dmif, resp, str(tmp_path / "a.mif"), "-mask", mmif, "-force"])
    ref, _ = mif.read(str(tmp_path / "a.mif"))
    # s-form, gzip
    S = np.concatenate([R * vox[None, :], t[:, None]], axis=1)
    d1, m1 = str(tmp_path / "dwi.nii.gz"), str(tmp_path / "mask.nii")
    nifti_util.write(d1, raw, vox=tuple(vox), sform=S, slope=float(slope), inter=float(inter))
    nifti_util.write(m1, mask, vox=tuple(vox), sform=S)
    # q-form only: the quaternion of R (w >= 0)
    w = 0.5 * np.sqrt(1.0 + np.trace(R))
    q = (np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]]) / (4 * w))
    d2 = str(tmp_path / "dwi_q.nii")
    nifti_util.write(d2, raw, vox=tuple(vox), quatern=tuple(q), qoffset=tuple(t), slope=float(slope), inter=float(inter))
    say("[gpu cli] dwi2fod csd from NIfTI (s-form .nii.gz, q-form .nii) on %d voxels" % n)
    run(["csd", d1, resp, str(tmp_path / "b.mif"), "-fslgrad", fv, fb, "-mask", m1, "-force"])
    run(["csd", d2, resp, str(tmp_path / "c.mif"), "-fslgrad", fv, fb, "-mask", m1, "-force"])
    for name in ("b.mif", "c.mif"):
        got, h = mif.read(str(tmp_path / name))
        assert got.shape == ref.shape
        assert np.allclose(np.array(h["transform"])[:, :3], R, atol=1e-6)
        diff = np.linalg.norm((got - ref).reshape(n, -1), axis=1)
        scale = np.maximum(np.linalg.norm(ref.reshape(n, -1), axis=1), 1e-30)
        rel = diff / scale
        say("[gpu cli] %s: max rel %.3e" % (name, rel.max()))
        assert np.mean(rel <= 1e-5) >= 0.999
        assert np.all(got.reshape(n, -1)[mask.reshape(-1) == 0] == 0)
