"""amp2response on the GPU (libmrcsd.so, include/mrresp.h): the response function's normal
equations accumulated over single-fibre voxels (cmd/amp2response.cpp:152-191) and the per-shell
solve (:336-361), constrained by the batched ICLS kernel in standard form.

    acc = Amp2Response(dirs, volumes, lmax=10)     # one shell: unit directions [nvol, 3], volume indices
    M, b, count = acc.accumulate(amp, fibre, mask)  # device tensors: [nvox, >=nvol], [nvox, >=3], [nvox]
    rf, iters = solve(10, M, b)                     # zonal coefficients l = 0, 2, ..., 10
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_registered = False


def _register():
    global _registered
    if _registered:
        return
    _lib.register("mrresp_create", C.c_int, [_D, _I, C.c_int, C.c_int, C.POINTER(_P)])
    _lib.register("mrresp_destroy", None, [_P])
    _lib.register("mrresp_ncoef", C.c_int, [_P])
    _lib.register("mrresp_accumulate", C.c_int, [_P, _P, C.c_int, _P, C.c_int, _P, C.c_size_t, _D, _D,
                                                 C.POINTER(C.c_uint64), _P])
    _lib.register("mrresp_solve", C.c_int, [C.c_int, _D, _D, C.c_int, _D, C.POINTER(C.c_int)])
    _registered = True


class Amp2Response:
    def __init__(self, dirs, volumes, lmax=10):
        _register()
        lib = _lib.load()
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(volumes, dtype=np.int32).reshape(-1)
        if len(v) != len(d):
            raise ValueError("one volume index per direction")
        h = C.c_void_p()
        check(lib.mrresp_create(dptr(d), v.ctypes.data_as(_I), len(v), int(lmax), C.byref(h)))
        self._h = h
        self.lmax = int(lmax)
        self.n = lib.mrresp_ncoef(h)

    def accumulate(self, amp, fibre, mask=None):
        """Device (torch CUDA) float32 amp [nvox, ld], fibre [nvox, >=3], optional uint8 mask [nvox].
        Returns (M [n, n], b [n], count) as numpy float64 / int."""
        import torch
        assert amp.is_cuda and amp.dtype == torch.float32 and amp.is_contiguous()
        assert fibre.is_cuda and fibre.dtype == torch.float32 and fibre.is_contiguous()
        assert fibre.shape[0] == amp.shape[0] and fibre.shape[1] >= 3
        if mask is not None:
            assert mask.is_cuda and mask.dtype == torch.uint8 and mask.numel() == amp.shape[0]
        M = np.zeros((self.n, self.n))
        b = np.zeros(self.n)
        cnt = C.c_uint64()
        stream = torch.cuda.current_stream(amp.device).cuda_stream
        check(_lib.load().mrresp_accumulate(self._h, amp.data_ptr(), amp.shape[1], fibre.data_ptr(), fibre.shape[1],
                                            mask.data_ptr() if mask is not None else None, amp.shape[0],
                                            dptr(M), dptr(b), C.byref(cnt), stream))
        return M, b, int(cnt.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrresp_destroy(h)
        self._h = None


def solve(lmax, M, b, constrained=True):
    """The shell's zonal response (amp2response.cpp:336-361); returns (rf, ICLS iterations)."""
    _register()
    n = int(lmax) // 2 + 1
    Mc = np.ascontiguousarray(M, dtype=np.float64).reshape(n, n)
    bc = np.ascontiguousarray(b, dtype=np.float64).reshape(n)
    rf = np.zeros(n)
    it = C.c_int()
    check(_lib.load().mrresp_solve(int(lmax), dptr(Mc), dptr(bc), int(bool(constrained)), dptr(rf), C.byref(it)))
    return rf, it.value
