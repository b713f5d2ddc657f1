"""The .mif format helpers and the dwi2fod drop-in's argument handling (CPU only: these paths
fail before any GPU work)."""
import os
import subprocess

import numpy as np
import pytest

from mrtrix3_amd import mif

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mrtrix3_amd", "bin", "dwi2fod")


def test_mif_roundtrip_layouts(tmp_path):
    rng = np.random.default_rng(0)
    data = rng.standard_normal((5, 4, 3, 7)).astype(np.float32)
    for vc in (True, False):
        p = str(tmp_path / ("a%d.mif" % vc))
        mif.write(p, data, vox=[2.0, 2.0, 2.5], keyval=[("dw_scheme", "0,0,1,0\n1,0,0,1000")],
                  volume_contiguous=vc)
        back, h = mif.read(p)
        assert np.array_equal(back, data)
        assert h["dim"] == [5, 4, 3, 7]
        assert [k for k, _ in h["keyval"]].count("dw_scheme") == 2
        raw = open(p, "rb").read()
        off = int(h["file"].split()[1])
        assert off % 4 == 0 and raw[:13] == b"mrtrix image\n"


def test_mif_reads_flipped_and_integer_layout(tmp_path):
    # hand-written header: int16 big-endian, axis 0 stored reversed, axis 1 fastest
    data = np.arange(2 * 3, dtype=np.int16).reshape(2, 3)
    p = str(tmp_path / "b.mif")
    head = b"mrtrix image\ndim: 2,3\nvox: 1,1\nlayout: -1,+0\ndatatype: Int16BE\nscaling: 1,2\nfile: . 96\nEND\n"
    filed = np.ascontiguousarray(np.flip(data, 0)).astype(">i2")   # C order (x slowest) = y fastest
    with open(p, "wb") as f:
        f.write(head + b"\0" * (96 - len(head)) + filed.tobytes())
    back, h = mif.read(p)
    assert np.array_equal(back, 1 + 2 * data.astype(np.float32))


@pytest.mark.skipif(not os.path.exists(EXE), reason="dwi2fod not built")
@pytest.mark.parametrize("args,msg", [
    (["csd"], "usage"),
    (["foo", "x.mif", "r.txt", "o.mif"], "algorithm must be one of"),
    (["csd", "/nonexistent.mif", "r.txt", "o.mif"], "failed to open image"),
])
def test_cli_argument_errors(args, msg):
    r = subprocess.run([EXE] + args, capture_output=True, text=True)
    assert r.returncode == 1 and msg in r.stderr


@pytest.mark.skipif(not os.path.exists(EXE), reason="dwi2fod not built")
def test_cli_csd_wrong_pair_count(tmp_path):
    p = str(tmp_path / "d.mif")
    mif.write(p, np.ones((2, 2, 2, 7), dtype=np.float32), keyval=[("dw_scheme", "\n".join(["0,0,1,1000"] * 7))])
    r = subprocess.run([EXE, "csd", p, "r.txt", "o.mif", "x.mif"], capture_output=True, text=True)
    assert r.returncode == 1 and "expects a single input response function" in r.stderr
    r = subprocess.run([EXE, "msmt_csd", p, "r.txt", "o.mif", "x.mif"], capture_output=True, text=True)
    assert r.returncode == 1 and "expects pairs" in r.stderr


@pytest.mark.skipif(not os.path.exists(EXE), reason="dwi2fod not built")
def test_cli_fslgrad_errors(tmp_path):
    """core/dwi/gradient.cpp:127-200 and :261-265: the bvecs / bvals checks, all before any GPU work."""
    p = str(tmp_path / "d.mif")
    mif.write(p, np.ones((2, 2, 2, 4), dtype=np.float32))
    bvecs, bvals, grad = (str(tmp_path / f) for f in ("bvecs", "bvals", "grad.txt"))

    def put(path, rows):
        with open(path, "w") as f:
            f.write("\n".join(" ".join(str(v) for v in r) for r in rows) + "\n")

    def err(*extra):
        r = subprocess.run([EXE, "csd", p, "r.txt", str(tmp_path / "o.mif")] + list(extra),
                           capture_output=True, text=True)
        assert r.returncode == 1
        return r.stderr

    put(grad, [[0, 0, 1, 1000]] * 4)
    put(bvecs, [[1, 0, 0, 1], [0, 1, 0, 0], [0, 0, 1, 0]])
    put(bvals, [[0, 1000, 1000, 1000]])
    assert "either -grad or -fslgrad option, but NOT both" in err("-grad", grad, "-fslgrad", bvecs, bvals)
    put(bvals, [[0, 1000], [1000, 1000]])
    assert "bvals file must contain 1 row or column only" in err("-fslgrad", bvecs, bvals)
    put(bvals, [[0, 1000, 1000]])
    assert "must have same number of diffusion directions (file" in err("-fslgrad", bvecs, bvals)
    put(bvecs, [[1, 0], [0, 1]])
    assert "bvecs file must contain exactly 3 rows or columns" in err("-fslgrad", bvecs, bvals)
    put(bvecs, [[1, 0, 0]] * 3)                       # three columns: read as one vector per row
    assert "as DW-image (gradients: 3, image: 4)" in err("-fslgrad", bvecs, bvals)
    put(bvecs, [[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 1]])
    put(bvals, [[1000], [float("nan")], [1000], [1000]])
    assert "NaN present in bval but valid direction in bvec" in err("-fslgrad", bvecs, bvals)


AMP2RESP = os.path.join(ROOT, "mrtrix3_amd", "bin", "amp2response")


@pytest.mark.skipif(not os.path.exists(AMP2RESP), reason="amp2response not built")
def test_amp2response_argument_errors(tmp_path):
    """cmd/amp2response.cpp:241-300: checks made before any GPU work."""
    amps, mask, dirs = (str(tmp_path / f) for f in ("amps.mif", "mask.mif", "dirs.mif"))
    scheme = "\n".join(["0,0,1,0"] * 2 + ["1,0,0,3000", "0,1,0,3000", "0,0,1,3000", "0.6,0.8,0,3000"])
    mif.write(amps, np.ones((2, 2, 2, 6), dtype=np.float32), keyval=[("dw_scheme", scheme)])
    mif.write(mask, np.zeros((2, 2, 2), dtype=np.float32))
    mif.write(dirs, np.ones((2, 2, 2, 3), dtype=np.float32))

    def err(*args):
        r = subprocess.run([AMP2RESP] + list(args), capture_output=True, text=True)
        assert r.returncode == 1, r.stderr
        return r.stderr

    assert "usage" in err(amps, mask, dirs)
    assert "input mask does not contain any voxels" in err(amps, mask, dirs, str(tmp_path / "r.txt"))
    mif.write(mask, np.ones((2, 2, 2), dtype=np.float32))
    assert "does not match number of b-value shells" in err(amps, mask, dirs, str(tmp_path / "r.txt"), "-lmax", "0,4,8")
    assert "must be even" in err(amps, mask, dirs, str(tmp_path / "r.txt"), "-lmax", "0,5")
    mif.write(dirs, np.ones((2, 2, 2, 2), dtype=np.float32))
    assert "does not have expected dimensions" in err(amps, mask, dirs, str(tmp_path / "r.txt"))


@pytest.mark.skipif(not os.path.exists(EXE), reason="dwi2fod not built")
def test_nifti_input_checks(tmp_path):
    """NIfTI-1 input (core/file/nifti_utils.cpp): what the reader rejects, before any GPU work."""
    import nifti_util
    dwi = str(tmp_path / "dwi.nii")
    nifti_util.write(dwi, np.ones((2, 2, 2, 7), dtype=np.float32))
    r = subprocess.run([EXE, "csd", dwi, "r.txt", str(tmp_path / "o.mif")], capture_output=True, text=True)
    assert r.returncode == 1 and "gradient" in r.stderr.lower()     # NIfTI holds no scheme
    bad = str(tmp_path / "two.nii")
    nifti_util.write(bad, np.ones((2, 2, 2), dtype=np.float32), sizeof_hdr=540)
    r = subprocess.run([EXE, "csd", bad, "r.txt", str(tmp_path / "o.mif")], capture_output=True, text=True)
    assert r.returncode == 1 and "NIfTI-2" in r.stderr
    short = str(tmp_path / "short.nii.gz")
    import gzip
    with gzip.open(short, "wb") as f:
        f.write(b"\0" * 100)
    r = subprocess.run([EXE, "csd", short, "r.txt", str(tmp_path / "o.mif")], capture_output=True, text=True)
    assert r.returncode == 1 and "too small" in r.stderr


F2F = os.path.join(ROOT, "mrtrix3_amd", "bin", "fod2fixel")


@pytest.mark.skipif(not os.path.exists(F2F), reason="fod2fixel not built")
def test_fod2fixel_argument_errors(tmp_path):
    """fod2fixel's checks before any GPU work (cmd/fod2fixel.cpp:285-320, fixel/helpers.h:190-211,
    Math::SH::check): usage, SH image shape, mask dimensions, the output directory, option names
    (unambiguous prefixes accepted, as the reference's parser does) and option ranges."""
    fod = str(tmp_path / "fod.mif")
    mif.write(fod, np.ones((2, 2, 2, 15), dtype=np.float32))
    bad = str(tmp_path / "bad.mif")
    mif.write(bad, np.ones((2, 2, 2, 14), dtype=np.float32))
    three = str(tmp_path / "three.mif")
    mif.write(three, np.ones((2, 2, 2), dtype=np.float32))
    msk = str(tmp_path / "mask.mif")
    mif.write(msk, np.ones((2, 3, 2), dtype=np.float32))
    full = tmp_path / "full"
    full.mkdir()
    (full / "x.txt").write_text("x")
    notdir = str(tmp_path / "file")
    open(notdir, "w").close()

    def fails(args, msg):
        r = subprocess.run([F2F] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)

    out = str(tmp_path / "fixels")
    fails([fod], "usage")
    fails([fod, out, "-nii"], "-nii")
    fails([bad, out], "unexpected number of coefficients")
    fails([three, out], "not 4D")
    fails([fod, out, "-mask", msk], "dimensions do not match FOD image")
    fails([fod, str(full)], "is not empty")
    fails([fod, notdir], "is not a directory")
    fails([fod, out, "-bogus"], "unknown option")
    fails([fod, out, "-fmls", "1"], "several matches possible")
    fails([fod, out, "-fmls_lobe_merge_ratio", "1.5"], "out of bounds")
    fails([fod, out, "-maxnum", "0"], "out of bounds")
    fails([fod, out, "-maxnum", "1.5"], "not an integer")
    fails([fod, out, "-fmls_integral"], "no value supplied")
    assert not os.path.exists(out)


D2R = os.path.join(ROOT, "mrtrix3_amd", "bin", "dwi2response")


@pytest.mark.skipif(not os.path.exists(D2R), reason="dwi2response not built")
def test_dwi2response_argument_errors(tmp_path):
    """dwi2response tournier's checks before any GPU work (bin/dwi2response execute(),
    tournier.py execute(), dwiextract's shell selection)."""
    scheme = "\n".join(["0,0,1,0"] * 2 + ["%r,%r,0,3000" % (float(np.cos(a)), float(np.sin(a))) for a in np.linspace(0, 3, 30)])
    dwi = str(tmp_path / "dwi.mif")
    mif.write(dwi, np.ones((3, 3, 2, 32), dtype=np.float32), keyval=[("dw_scheme", scheme)])
    nograd = str(tmp_path / "nograd.mif")
    mif.write(nograd, np.ones((3, 3, 2, 32), dtype=np.float32))
    m_ok = str(tmp_path / "m.mif")
    mif.write(m_ok, np.ones((3, 3, 2), dtype=np.float32))
    m_bad = str(tmp_path / "mbad.mif")
    mif.write(m_bad, np.ones((3, 2, 2), dtype=np.float32))
    m_empty = str(tmp_path / "m0.mif")
    mif.write(m_empty, np.zeros((3, 3, 2), dtype=np.float32))
    exists = str(tmp_path / "exists.txt")
    open(exists, "w").close()
    out = str(tmp_path / "r.txt")

    def fails(args, msg):
        r = subprocess.run([D2R] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)

    fails(["tournier", dwi], "usage")
    fails(["dhollander", dwi, out], "not part of this build")
    fails(["tax", dwi, out, "-mask", m_ok, "-number", "10"], "unrecognized arguments: -number")
    fails(["tournier", dwi, out, "-mask", m_ok, "-peak_ratio", "0.2"], "unrecognized arguments: -peak_ratio")
    fails(["tax", dwi, out, "-mask", m_ok, "-peak_ratio", "x"], "invalid float value")
    fails(["tournier", dwi, exists, "-mask", m_ok], "already exists")
    fails(["tournier", dwi, out, "-mask", m_ok, "-lmax", "7"], "Parameter lmax must be a number")
    fails(["tournier", dwi, out, "-mask", m_ok, "-lmax", "4,6"], "single lmax value")
    fails(["tournier", dwi, out, "-mask", m_ok, "-shells", "1000,3000"], "single b-value shell")
    fails(["tournier", dwi, out, "-mask", m_ok, "-shells", "x"], "comma-separated list of b-values")
    fails(["tournier", dwi, out, "-mask", m_ok, "-max_iters", "1"], "at least 2")
    fails(["tournier", dwi, out, "-mask", m_ok, "-number", "10", "-iter_voxels", "5"], "-iter_voxels")
    fails(["tournier", nograd, out, "-mask", m_ok], "requires diffusion gradient table")
    fails(["tournier", dwi, out], "-mask is required")
    fails(["tournier", dwi, out, "-mask", m_bad], "do not match DWI")
    fails(["tournier", dwi, out, "-mask", m_empty], "does not contain any voxels")
    fails(["tournier", dwi, out, "-mask", m_ok, "-shells", "0"], "User selected 0 non b=0 shells")
    assert not os.path.exists(out)
