"""dwi2response tournier and tax on the GPU (mrtrix3_amd/bin/dwi2response) against the oracle's
restatement of the scripts' loops (oracle/dwi2response_oracle.py), iteration by iteration: each
oracle step starts from the command's own mask and response of the previous iteration (its
-scratch -nocleanup files), so every iteration is checked on its own inputs.

Bars: the cost image equal to the oracle's where either is NaN and within 1e-5 relative elsewhere
(the FODs of the GPU and of the oracle agree to about 1e-7, the cost is float32 arithmetic on
them); the selection and the next iteration's mask identical; the response within 1e-6 of its
first coefficient (FP64 sums over float32 signals and float32 fibre directions); the output files
equal to the last iteration's."""
import os
import subprocess
import sys

import numpy as np
import pytest

from mrtrix3_amd import mif, phantom
from oracle import dwi2response_oracle as T

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mrtrix3_amd", "bin", "dwi2response")
SHAPE = (16, 16, 8)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def flat(img):
    """[x, y, z] image -> flat x-fastest."""
    return np.asarray(img).transpose(2, 1, 0).reshape(-1)


def load_rf(path):
    return np.array([float(x) for x in open(path).read().split()])


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    d = tmp_path_factory.mktemp("tournier")
    ph = phantom.single_shell_small()
    n = int(np.prod(SHAPE))
    img = ph.dw[:n].reshape(SHAPE[2], SHAPE[1], SHAPE[0], -1).transpose(2, 1, 0, 3)
    scheme = "\n".join(",".join(repr(float(x)) for x in row) for row in ph.grad)
    mif.write(str(d / "dwi.mif"), img, vox=[2.0, 2.0, 2.0], keyval=[("dw_scheme", scheme)])
    rng = np.random.default_rng(12)
    mask = rng.random(n) < 0.9
    mif.write(str(d / "mask.mif"), mask.reshape(SHAPE[2], SHAPE[1], SHAPE[0]).transpose(2, 1, 0).astype(np.float32),
              vox=[2.0, 2.0, 2.0])
    keep = ph.grad[:, 3] > 10
    return dict(dir=d, ph=ph, mask=mask, dw=np.ascontiguousarray(ph.dw[:n][:, keep]), grad=ph.grad[keep])


def test_tournier_iterations_match_oracle(case):
    assert os.path.exists(EXE), "dwi2response is not built"
    d = case["dir"]
    args = [EXE, "tournier", str(d / "dwi.mif"), str(d / "resp.txt"), "-mask", str(d / "mask.mif"), "-number",
            "100", "-iter_voxels", "600", "-scratch", str(d / "scratch"), "-nocleanup", "-voxels",
            str(d / "voxels.mif"), "-info"]
    say("[gpu] dwi2response tournier on %d voxels" % int(case["mask"].sum()))
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    say(r.stderr.strip().replace("dwi2response: ", "  "))
    sc = d / "scratch"
    K = 0
    while (sc / ("iter%d_RF.txt" % K)).exists():
        K += 1
    assert K >= 2 and "Convergence of SF voxel selection detected at iteration %d" % K in r.stderr
    mask = case["mask"]
    for k in range(K):
        mask_in = mask if k == 0 else flat(mif.read(str(sc / ("iter%d_SF_dilated.mif" % (k - 1))))[0]) != 0
        rf_in = [1.0, -1.0, 1.0] if k == 0 else load_rf(str(sc / ("iter%d_RF.txt" % (k - 1))))
        s = T.step(case["dw"], case["grad"], SHAPE, mask_in, rf_in, 4 if k == 0 else None, 100)
        cf = flat(mif.read(str(sc / ("iter%d_CF.mif" % k)))[0])
        nan_g, nan_o = np.isnan(cf), np.isnan(s["cf"])
        assert np.array_equal(nan_g, nan_o), "iteration %d: NaN costs differ in %d voxels" % (k, int((nan_g ^ nan_o).sum()))
        np.testing.assert_allclose(cf[~nan_g], s["cf"][~nan_o], rtol=1e-5, atol=1e-6, err_msg="iteration %d" % k)
        sf = flat(mif.read(str(sc / ("iter%d_SF.mif" % k)))[0]) != 0
        assert np.array_equal(sf, s["sf"]), "iteration %d: selections differ in %d voxels" % (k, int((sf ^ s["sf"]).sum()))
        rf = load_rf(str(sc / ("iter%d_RF.txt" % k)))
        assert rf.shape == s["rf"].shape
        np.testing.assert_allclose(rf, s["rf"], rtol=0, atol=1e-6 * abs(s["rf"][0]), err_msg="iteration %d" % k)
        if k < K - 1:
            nxt = flat(mif.read(str(sc / ("iter%d_SF_dilated.mif" % k)))[0]) != 0
            assert np.array_equal(nxt, T.next_mask(s["cf"], s["count"], 600, 1, mask, SHAPE)), "iteration %d" % k
        say("[gpu] iteration %d: %d voxels tested, %d selected, rf[0] %.4f (oracle %.4f)" % (
            k, int(mask_in.sum()), int(sf.sum()), rf[0], s["rf"][0]))
    # outputs: the last iteration's response and selection
    np.testing.assert_array_equal(load_rf(str(d / "resp.txt")), load_rf(str(sc / ("iter%d_RF.txt" % (K - 1)))))
    assert np.array_equal(flat(mif.read(str(d / "voxels.mif"))[0]) != 0,
                          flat(mif.read(str(sc / ("iter%d_SF.mif" % (K - 1))))[0]) != 0)


def test_tournier_iteration_cap_and_lmax(case):
    """-max_iters 2 -lmax 6: the response of the second iteration (lmax 6, four coefficients) and
    the 'Exiting after maximum' path when the selection still changes."""
    d = case["dir"]
    out = d / "resp2.txt"
    r = subprocess.run([EXE, "tournier", str(d / "dwi.mif"), str(out), "-mask", str(d / "mask.mif"), "-number", "100",
                        "-iter_voxels", "600", "-max_iters", "2", "-lmax", "6"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    rf = load_rf(str(out))
    rf_o, sf_o, its, conv, hist = T.run(case["dw"], case["grad"], SHAPE, case["mask"], number=100, iter_voxels=600,
                                        max_iters=2, lmax=6)
    assert rf.shape == (4,) and rf_o.shape == (4,)
    assert ("Exiting after maximum 2 iterations" in r.stderr) == (not conv)
    np.testing.assert_allclose(rf, rf_o, rtol=0, atol=1e-5 * abs(rf_o[0]))
    say("[gpu] -max_iters 2 -lmax 6: %s (oracle %s)" % (np.round(rf, 3), np.round(rf_o, 3)))


def test_tax_iterations_match_oracle(case):
    """dwi2response tax: the initial response from the masked signal's statistics as mrstats prints
    them, then each iteration replayed by the oracle from the command's previous response and
    selection: peak-ratio image (NaN pattern, 1e-5), selection identical, response within 1e-6,
    termination by the convergence test."""
    d = case["dir"]
    sc = d / "scratch_tax"
    r = subprocess.run([EXE, "tax", str(d / "dwi.mif"), str(d / "resp_tax.txt"), "-mask", str(d / "mask.mif"),
                        "-scratch", str(sc), "-nocleanup", "-voxels", str(d / "voxels_tax.mif"), "-info"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    say(r.stderr.strip().replace("dwi2response: ", "  "))
    mask = case["mask"]
    init = load_rf(str(sc / "init_RF.txt"))
    np.testing.assert_allclose(init, T.tax_init_rf(case["dw"], mask), rtol=1e-12)
    K = 0
    while (sc / ("iter%d_RF.txt" % K)).exists():
        K += 1
    assert K >= 2
    prev = init
    for k in range(K):
        mask_in = mask if k == 0 else flat(mif.read(str(sc / ("iter%d_SF.mif" % (k - 1))))[0]) != 0
        s = T.tax_step(case["dw"], case["grad"], SHAPE, mask_in, prev, None, 0.1)
        ratio = flat(mif.read(str(sc / ("iter%d_peak_ratio.mif" % k)))[0])
        nan_g, nan_o = np.isnan(ratio), np.isnan(s["ratio"])
        assert np.array_equal(nan_g, nan_o), "iteration %d" % k
        np.testing.assert_allclose(ratio[~nan_g], s["ratio"][~nan_o], rtol=1e-5, atol=1e-7, err_msg="iteration %d" % k)
        sf = flat(mif.read(str(sc / ("iter%d_SF.mif" % k)))[0]) != 0
        assert np.array_equal(sf, s["sf"]), "iteration %d: selections differ in %d voxels" % (k, int((sf ^ s["sf"]).sum()))
        rf = load_rf(str(sc / ("iter%d_RF.txt" % k)))
        np.testing.assert_allclose(rf, s["rf"], rtol=0, atol=1e-6 * abs(s["rf"][0]), err_msg="iteration %d" % k)
        if k > 0:
            assert T.tax_converged(prev, rf, 0.5) == (k == K - 1), "iteration %d" % k
        say("[gpu] tax iteration %d: %d voxels tested, %d selected, rf[0] %.4f (oracle %.4f)" % (
            k, int(mask_in.sum()), int(sf.sum()), rf[0], s["rf"][0]))
        prev = rf
    assert "Exited at iteration %d with %d SF voxels due to unchanged RF coefficients" % (K, int(sf.sum())) in r.stderr
    np.testing.assert_array_equal(load_rf(str(d / "resp_tax.txt")), prev)
    assert np.array_equal(flat(mif.read(str(d / "voxels_tax.mif"))[0]) != 0, sf)
