"""The product's host-side CSD::Shared set-up (libmrcsd.so, no GPU needed) against the oracle."""
import numpy as np
import pytest

from mrtrix3_amd import csd, phantom
from mrtrix3_amd._lib import MrcsdError
from oracle import orc


def compare(shared, o, rconv_tol=1e-12):
    """rconv comes from normal equations of a (near-)square system, so two correct
    implementations differ by cond(fconv)^2 * eps; pass a looser rconv_tol for such tables."""
    assert (shared.n_dw, shared.n_sh, shared.n_init, shared.n_hr) == (o.n_dw, o.n_sh, o.n_init, o.n_hr)
    assert (shared.lmax, shared.lmax_response, shared.lmax_data) == (o.lmax, o.lmax_response, o.lmax_data)
    assert shared.threshold == pytest.approx(o.threshold, rel=1e-14, abs=1e-300)
    assert shared.norm_lambda == pytest.approx(o.norm_lambda, rel=1e-14, abs=1e-300)
    assert np.array_equal(shared.dwis, o.vols)
    for name, ref in (("rconv", o.rconv), ("M", o.M), ("Mt_M", o.MtM), ("HR_trans", o.HR_trans)):
        got = shared.matrix(name)
        scale = max(np.abs(ref).max(), 1e-300)
        tol = rconv_tol if name == "rconv" else 1e-12
        assert np.abs(got - ref).max() <= tol * scale, name


def test_default_setup(small_phantom, small_oracle):
    compare(csd.CSDShared(small_phantom.grad, small_phantom.response), small_oracle)


@pytest.mark.parametrize("kw", [dict(lmax=6), dict(lmax=4), dict(neg_lambda=0.5, norm_lambda=2.0),
                                dict(threshold=0.1), dict(niter=7), dict(filter=[1.0, 0.5, 0.2, 0.1, 0.05]),
                                dict(filter=[1.0, 1.0])])
def test_option_variants(small_phantom, kw):
    ph = small_phantom
    compare(csd.CSDShared(ph.grad, ph.response, **kw), orc.CsdOracle(ph.grad, ph.response, **kw))


def test_custom_directions(small_phantom):
    ph = small_phantom
    rng = np.random.default_rng(4)
    v = rng.standard_normal((120, 3))
    d = np.stack([np.arctan2(v[:, 1], v[:, 0]), np.arccos(v[:, 2] / np.linalg.norm(v, axis=1))], axis=1)
    compare(csd.CSDShared(ph.grad, ph.response, directions=d),
            orc.CsdOracle(ph.grad, ph.response, directions=d))


def test_multishell_table_selects_largest_shell():
    rng = np.random.default_rng(9)
    grad = phantom.gradient_table([(0, 6), (1000, 20), (2000, 30), (3000, 45)], rng)
    grad[:, 3] += rng.uniform(-30, 30, grad.shape[0]) * (grad[:, 3] > 0)
    resp = phantom.zonal_response(3000, 8)
    s = csd.CSDShared(grad, resp)
    o = orc.CsdOracle(grad, resp)
    assert s.n_dw == 45
    compare(s, o, rconv_tol=1e-8)


def test_bvalue_scaling_of_non_unit_vectors():
    rng = np.random.default_rng(2)
    grad = phantom.gradient_table([(0, 4), (2500, 40)], rng)
    scale = rng.uniform(0.8, 1.2, grad.shape[0])
    g2 = grad.copy()
    g2[:, :3] *= np.sqrt(scale)[:, None]
    got = csd.CSDShared(g2, phantom.zonal_response(2500, 8)).matrix("grad")
    want, scaled = orc.normalise_grad(g2)
    assert scaled
    assert np.allclose(got, want, rtol=1e-14, atol=1e-12)
    assert np.allclose(np.linalg.norm(got[4:, :3], axis=1), 1.0)


def test_shell_clustering_rules():
    b = np.array([0, 5, 1000, 1010, 990, 1030, 2000, 2040, 1990, 2010, 3000, 2950, 3020, 3010, 3500])
    n, cl = orc.shells_cluster(b.astype(float))
    # b=0 cluster, three shells, and the lone 3500 left unassigned (fewer than 3 neighbours)
    assert n == 4 and cl[-1] == 0 and cl[0] == cl[1] == 1
    assert len(set(cl[2:6])) == 1 and len(set(cl[6:10])) == 1 and len(set(cl[10:14])) == 1


@pytest.mark.parametrize("resp,kw,msg", [
    ([500.0], {}, "response function does not contain anisotropic terms"),
    ([500.0, -200.0, 50.0], dict(lmax=5), "lmax must be a positive even integer"),
])
def test_error_messages(small_phantom, resp, kw, msg):
    with pytest.raises(MrcsdError, match=msg):
        csd.CSDShared(small_phantom.grad, resp, **kw)


def test_too_few_directions():
    rng = np.random.default_rng(1)
    grad = phantom.gradient_table([(0, 2), (3000, 5)], rng)
    with pytest.raises(MrcsdError, match="too few directions even for lmax = 2"):
        csd.CSDShared(grad, phantom.zonal_response(3000, 8))


def test_gradient_table_validation():
    with pytest.raises(ValueError, match="gradient table must be"):
        csd.CSDShared(np.zeros((10, 3)), [1.0, 2.0])
    with pytest.raises(MrcsdError, match="could not be classified into b-value shells"):
        csd.CSDShared(np.array([[1.0, 0, 0, 1000.0], [0, 1.0, 0, 3000.0]]), [1.0, 2.0])


def test_shells_option_selects_the_requested_shell():
    """-shells (DWI::ShellsOption, shells.cpp:100-281): CSD on a four-shell table uses the
    requested shell; its set-up equals the oracle's on the table reduced to that shell."""
    rng = np.random.default_rng(9)
    grad = phantom.gradient_table([(0, 6), (1000, 20), (2000, 30), (3000, 45)], rng)
    resp = phantom.zonal_response(2000, 8)
    s = csd.CSDShared(grad, resp, shells=[2000])
    sel = np.nonzero((grad[:, 3] == 0) | (np.abs(grad[:, 3] - 2000) < 100))[0]
    o = orc.CsdOracle(grad[sel], resp)
    assert s.n_dw == 30 == o.n_dw
    assert np.array_equal(s.dwis, sel[o.vols])
    for name, ref in (("rconv", o.rconv), ("M", o.M), ("Mt_M", o.MtM), ("HR_trans", o.HR_trans)):
        assert np.abs(s.matrix(name) - ref).max() <= 1e-8 * np.abs(ref).max(), name
    with pytest.raises(MrcsdError):
        csd.CSDShared(grad, resp, shells=[1500])


def test_msmt_shells_option_zeroes_unselected_volumes():
    from mrtrix3_amd import icls
    rng = np.random.default_rng(10)
    grad = phantom.gradient_table([(0, 6), (1000, 20), (3000, 45)], rng)
    wm = np.stack([phantom.zonal_response(b, 8) if b else np.r_[phantom.zonal_response(0, 8)[:1], np.zeros(4)]
                   for b in (0, 3000)])
    gm = np.array([[2.0], [0.5]])
    s = icls.MSMTShared(grad, [wm, gm], shells=[0, 3000])
    assert s.nshell == 2
    C = s.matrix("C")
    b1000 = np.abs(grad[:, 3] - 1000) < 100
    assert np.all(C[b1000] == 0) and np.all(np.abs(C[~b1000]).sum(axis=1) > 0)
