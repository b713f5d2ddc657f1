"""mricls_solve_noalloc (include/mricls.h) through the C ABI against mricls_solve: the same kernels on
the same inputs, so bit-identical solutions, iteration counts and active-set sizes, on two calls in
a row (the second smaller, with the first call's queue contents left in the fixed scratch), with a
16-slot first pass so that the voxel queues of the scratch carry every voxel to the 64-slot pass."""
import ctypes as C
import sys

import numpy as np
import pytest

from mrtrix3_amd import _lib, icls, phantom

pytestmark = pytest.mark.gpu

_P = C.c_void_p


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def noalloc_fn():
    return _lib.register("mricls_solve_noalloc", C.c_int,
                         [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P, _P, _P])


def solve_noalloc(torch, prob, b, valid=None):
    nvox, ld = b.shape
    x = torch.full((nvox, prob.n), float("nan"), dtype=torch.float32, device=b.device)
    it = torch.zeros(nvox, dtype=torch.int16, device=b.device)
    na = torch.zeros(nvox, dtype=torch.int16, device=b.device)
    st = torch.zeros(8, dtype=torch.int64, device=b.device)
    stream = torch.cuda.current_stream(b.device).cuda_stream
    _lib.check(noalloc_fn()(prob._h, b.data_ptr(), ld, valid.data_ptr() if valid is not None else None, nvox,
                            x.data_ptr(), it.data_ptr(), na.data_ptr(), st.data_ptr(), stream))
    return x, it, na, st


def test_noalloc_matches_solve_on_two_calls(torch_cuda):
    torch = torch_cuda
    ph = phantom.multi_shell(seed=4, shape=(16, 16, 8), semi=None)
    prob = icls.MSMTShared(ph.grad, ph.response).problem()
    prob.capacity = 16
    cases = [(ph.dw, np.arange(ph.nvox) % 7 != 3),
             (phantom.multi_shell(seed=9, shape=(8, 8, 8), semi=None).dw[:700], None)]
    for call, (dw, keep) in enumerate(cases):
        b = torch.from_numpy(np.ascontiguousarray(dw)).cuda()
        valid = torch.from_numpy(keep.astype(np.uint8)).cuda() if keep is not None else None
        say("[gpu noalloc] call %d: %d voxels, first pass %d slots" % (call, b.shape[0], prob.capacity))
        x0, it0, na0 = prob.solve(b, valid=valid)
        x1, it1, na1, st = solve_noalloc(torch, prob, b, valid)
        torch.cuda.synchronize()
        nvalid = int(keep.sum()) if keep is not None else b.shape[0]
        say("[gpu noalloc] call %d: stats %s" % (call, st.cpu().numpy().tolist()))
        assert torch.equal(x0, x1) and torch.equal(it0, it1) and torch.equal(na0, na1)
        assert int(st[0]) == nvalid and int(st[7]) == 0 and int(na1.max()) > 16


def test_noalloc_rejects_oversized_batches(torch_cuda):
    torch = torch_cuda
    ph = phantom.multi_shell(seed=4, shape=(4, 4, 4), semi=None)
    prob = icls.MSMTShared(ph.grad, ph.response).problem()
    b = torch.from_numpy(np.ascontiguousarray(ph.dw)).cuda()
    x = torch.empty((b.shape[0], prob.n), dtype=torch.float32, device="cuda")
    rc = noalloc_fn()(prob._h, b.data_ptr(), b.shape[1], None, (1 << 20) + 1, x.data_ptr(), None, None, None, None)
    assert rc == -1 and b"at most 1048576" in _lib.load().mrcsd_last_error()
