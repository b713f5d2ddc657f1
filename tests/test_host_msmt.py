"""Host-side MSMT_CSD::Shared set-up (libmrcsd.so, no GPU) against the oracle's restatement of
msmt_csd.h:100-218."""
import ctypes as C

import numpy as np
import pytest

from mrtrix3_amd import icls, phantom
from mrtrix3_amd._lib import MrcsdError
from oracle import orc


def oracle_msmt(grad, responses, lmax):
    g, _ = orc.normalise_grad(grad)
    n, cl = orc.shells_cluster(g[:, 3])
    ids = [c for c in np.unique(cl) if c > 0]
    means = {c: g[cl == c, 3].mean() for c in ids}
    order = sorted(ids, key=lambda c: means[c])
    shell_of = np.array([order.index(c) if c > 0 else -1 for c in cl], dtype=np.int32)
    resp = np.ascontiguousarray(np.concatenate([np.asarray(r, float).ravel() for r in responses]))
    cols = np.array([np.atleast_2d(r).shape[1] for r in responses], dtype=np.int32)
    lm = np.array(lmax, dtype=np.int32)
    nparams = int(sum((l + 1) * (l + 2) // 2 for l in lmax))
    ncon = int(sum(300 if l > 0 else 1 for l in lmax))
    Cm = np.zeros((g.shape[0], nparams))
    A = np.zeros((ncon, nparams))
    nc, nr = C.c_int(), C.c_int()
    hr = orc.load_dirs300()
    orc.lib().orc_msmt_build(orc._d(g), g.shape[0], shell_of.ctypes.data_as(orc.I), len(order), orc._d(resp),
                             cols.ctypes.data_as(orc.I), len(responses), lm.ctypes.data_as(orc.I), orc._d(hr),
                             300, orc._d(Cm), orc._d(A), C.byref(nc), C.byref(nr))
    return Cm, A


@pytest.fixture(scope="module")
def ms_small():
    return phantom.multi_shell(seed=4, shape=(8, 8, 4), semi=None)


def test_msmt_matrices_match_oracle(ms_small):
    ph = ms_small
    s = icls.MSMTShared(ph.grad, ph.response)
    assert s.lmax == [8, 0, 0] and s.nshell == 4 and s.nparams == 47 and s.ncon == 302
    Cr, Ar = oracle_msmt(ph.grad, ph.response, s.lmax)
    assert np.allclose(s.matrix("C"), Cr, rtol=1e-13, atol=1e-12 * np.abs(Cr).max())
    assert np.allclose(s.matrix("A"), Ar, rtol=1e-13, atol=1e-13)


def test_msmt_lmax_option(ms_small):
    ph = ms_small
    s = icls.MSMTShared(ph.grad, ph.response, lmax=[6, 0, 0])
    assert s.lmax == [6, 0, 0] and s.nparams == 30
    Cr, Ar = oracle_msmt(ph.grad, ph.response, [6, 0, 0])
    assert np.allclose(s.matrix("C"), Cr, rtol=1e-13, atol=1e-12 * np.abs(Cr).max())


def test_msmt_errors(ms_small):
    ph = ms_small
    with pytest.raises(MrcsdError, match="does not match number of tissues"):
        icls.MSMTShared(ph.grad, ph.response, lmax=[8, 0])
    with pytest.raises(MrcsdError, match="non-negative even integer"):
        icls.MSMTShared(ph.grad, ph.response, lmax=[7, 0, 0])
    bad = [ph.response[0][:3], ph.response[1], ph.response[2]]
    with pytest.raises(MrcsdError, match="must match number of b-value shells"):
        icls.MSMTShared(ph.grad, bad)
