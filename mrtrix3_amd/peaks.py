"""The peaks of an SH series in every voxel on the GPU (libmrcsd.so, include/mrpeaks.h): the
sh2peaks command's Processor (cmd/sh2peaks.cpp:145-300) over a batch of voxels.

    finder = SH2Peaks(n_sh=45)                       # -num 3, built-in 60 search directions
    peaks = finder.run(fod)                          # [nvox, 9]: amplitude x direction per peak

Options mirror the command's: npeaks (-num), threshold (-threshold), seeds (-seeds: [n, 2]
azimuth / elevation in radians), directions (-direction: [n, 2] phi / theta in radians), and per
run a mask (-mask) and per-voxel directions to match (peaks, -peaks).  Arithmetic is single
precision, as the reference's.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check

_P = C.c_void_p
_registered = False


class Options(C.Structure):
    _fields_ = [("n_sh", C.c_int), ("npeaks", C.c_int), ("threshold", C.c_float),
                ("seeds", C.POINTER(C.c_float)), ("n_seeds", C.c_int),
                ("directions", C.POINTER(C.c_float)), ("n_directions", C.c_int), ("fast", C.c_int)]


def _register():
    global _registered
    if _registered:
        return
    _lib.register("mrpeaks_options_default", None, [C.POINTER(Options)])
    _lib.register("mrpeaks_create", C.c_int, [C.POINTER(Options), C.POINTER(_P)])
    _lib.register("mrpeaks_destroy", None, [_P])
    _lib.register("mrpeaks_npeaks", C.c_int, [_P])
    _lib.register("mrpeaks_run", C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, C.c_size_t, _P, _P, _P])
    _lib.register("mrpeaks_run_host", C.c_int, [_P, _P, C.c_int, _P, _P, C.c_int, C.c_size_t, _P])
    _registered = True


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class SH2Peaks:
    def __init__(self, n_sh, npeaks=3, threshold=-math.inf, seeds=None, directions=None, fast=False):
        _register()
        lib = _lib.load()
        opt = Options()
        lib.mrpeaks_options_default(C.byref(opt))
        opt.n_sh = int(n_sh)
        opt.npeaks = int(npeaks)
        opt.threshold = float(threshold)
        opt.fast = 1 if fast else 0
        self._keep = []
        if seeds is not None:
            s = np.ascontiguousarray(seeds, dtype=np.float32).reshape(-1, 2)
            self._keep.append(s)
            opt.seeds = s.ctypes.data_as(C.POINTER(C.c_float))
            opt.n_seeds = s.shape[0]
        if directions is not None:
            d = np.ascontiguousarray(directions, dtype=np.float32).reshape(-1, 2)
            self._keep.append(d)
            opt.directions = d.ctypes.data_as(C.POINTER(C.c_float))
            opt.n_directions = d.shape[0]
        h = _P()
        check(lib.mrpeaks_create(C.byref(opt), C.byref(h)))
        self._h = h
        self.n_sh = int(n_sh)
        self.npeaks = int(lib.mrpeaks_npeaks(h))

    def run(self, sh, mask=None, peaks=None, stats=None, out=None):
        """sh: [nvox, >= n_sh] float32 (torch CUDA tensor or numpy).  Returns [nvox, 3 npeaks]."""
        lib = _lib.load()
        if _is_torch(sh):
            import torch
            assert sh.is_cuda and sh.dtype == torch.float32 and sh.is_contiguous() and sh.dim() == 2
            nvox, ld = sh.shape
            if out is None:
                out = torch.empty((nvox, 3 * self.npeaks), dtype=torch.float32, device=sh.device)
            if mask is not None:
                assert mask.is_cuda and mask.dtype == torch.uint8 and mask.numel() == nvox
            ldp = 0
            if peaks is not None:
                assert peaks.is_cuda and peaks.dtype == torch.float32 and peaks.shape[0] == nvox
                ldp = peaks.shape[1]
            stream = torch.cuda.current_stream(sh.device).cuda_stream
            check(lib.mrpeaks_run(self._h, sh.data_ptr(), ld, mask.data_ptr() if mask is not None else None,
                                  peaks.data_ptr() if peaks is not None else None, ldp, nvox, out.data_ptr(),
                                  stats.data_ptr() if stats is not None else None, stream))
            return out
        sh = np.ascontiguousarray(np.atleast_2d(sh), dtype=np.float32)
        nvox, ld = sh.shape
        res = np.empty((nvox, 3 * self.npeaks), dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        p = None if peaks is None else np.ascontiguousarray(peaks, dtype=np.float32)
        check(lib.mrpeaks_run_host(self._h, sh.ctypes.data, ld, m.ctypes.data if m is not None else None,
                                   p.ctypes.data if p is not None else None, p.shape[1] if p is not None else 0,
                                   nvox, res.ctypes.data))
        return res

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrpeaks_destroy(h)
        self._h = None
