"""Host-side MSMT_CSD::Shared set-up (libmrcsd.so, no GPU) against the oracle's restatement of
msmt_csd.h:100-218."""
import numpy as np
import pytest

from mrtrix3_amd import icls, phantom
from mrtrix3_amd._lib import MrcsdError
from oracle import orc


def oracle_msmt(grad, responses, lmax):
    C, A, _ = orc.msmt_matrices(grad, responses, lmax=lmax)
    return C, A


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
