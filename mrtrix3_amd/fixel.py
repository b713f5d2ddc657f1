"""FOD segmentation into fixels on the GPU (libmrcsd.so, include/mrfixel.h): DWI::FMLS::Segmenter
(src/dwi/fmls.cpp:144-377) with fod2fixel's per-voxel output (cmd/fod2fixel.cpp:108-283).

    seg = Segmenter(n_sh=45)                       # 1281 directions, peak threshold 0.1
    res = seg.segment(fod, mask)                   # fixels of every voxel, in voxel order
    res.count, res.offset                          # the index image's two volumes
    res.directions, res.afd, res.peak_amp, res.disp

Options mirror the command's: integral_threshold (-fmls_integral), peak_value_threshold
(-fmls_peak_value), no_thresholds (-fmls_no_thresholds), lobe_merge_ratio
(-fmls_lobe_merge_ratio), maxnum (-maxnum), dir_from_peak (-dirpeak).  Arithmetic is double
precision, as the reference's; outputs are float32 as fod2fixel writes them.
"""
import ctypes as C
import warnings

import numpy as np

from . import _lib
from ._lib import check

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_registered = False


class Options(C.Structure):
    _fields_ = [("n_sh", C.c_int), ("integral_threshold", C.c_double), ("peak_value_threshold", C.c_double),
                ("lobe_merge_ratio", C.c_double), ("maxnum", C.c_int), ("dir_from_peak", C.c_int),
                ("directions", _D), ("n_directions", C.c_int)]


class SetInfo(C.Structure):
    _fields_ = [("ndir", C.c_int), ("n_az", C.c_int), ("n_el", C.c_int), ("n_cell_dirs", C.c_int)]


def _register():
    global _registered
    if _registered:
        return
    _lib.register("mrfixel_options_default", None, [C.POINTER(Options)])
    _lib.register("mrfixel_create", C.c_int, [C.POINTER(Options), C.POINTER(_P)])
    _lib.register("mrfixel_destroy", None, [_P])
    _lib.register("mrfixel_segment", C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P, C.c_size_t, _P, _P, _P])
    _lib.register("mrfixel_segment_host", C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P, C.c_size_t, _P,
                                                    C.POINTER(C.c_uint64)])
    _lib.register("mrfixel_direction_set", C.c_int, [_D, C.c_int, C.POINTER(SetInfo), _D, _I, _D, _I, _I])
    _lib.register("mrfixel_legendre_table", C.c_int, [C.c_int, C.c_int, _D])
    _registered = True


class DirectionSet:
    """The tables a segmenter is built on (host only): unit vectors, hull adjacency [n, 8] (-1
    padded), integration weights and the FastLookupSet grid."""

    def __init__(self, az_el=None):
        _register()
        lib = _lib.load()
        info = SetInfo()
        a = None if az_el is None else np.ascontiguousarray(az_el, dtype=np.float64).reshape(-1, 2)
        ap = None if a is None else a.ctypes.data_as(_D)
        n = 0 if a is None else a.shape[0]
        check(lib.mrfixel_direction_set(ap, n, C.byref(info), None, None, None, None, None))
        self.n, self.n_az, self.n_el = info.ndir, info.n_az, info.n_el
        self.unit = np.zeros((info.ndir, 3))
        self.adj = np.zeros((info.ndir, 8), dtype=np.int32)
        self.weights = np.zeros(info.ndir)
        self.cell_start = np.zeros(info.n_az * info.n_el + 1, dtype=np.int32)
        self.cell_dirs = np.zeros(info.n_cell_dirs, dtype=np.int32)
        check(lib.mrfixel_direction_set(ap, n, C.byref(info), self.unit.ctypes.data_as(_D),
                                        self.adj.ctypes.data_as(_I), self.weights.ctypes.data_as(_D),
                                        self.cell_start.ctypes.data_as(_I), self.cell_dirs.ctypes.data_as(_I)))


def legendre_table(lmax, nrows):
    """PrecomputedAL<double> (core/math/SH.h:418-434): [nrows, (lmax/2+1)^2]."""
    _register()
    out = np.zeros((nrows, (lmax // 2 + 1) ** 2))
    check(_lib.load().mrfixel_legendre_table(int(lmax), int(nrows), out.ctypes.data_as(_D)))
    return out


class Fixels:
    """fod2fixel's output for a batch of voxels: count / offset per voxel (the index image) and the
    fixels in voxel order; summary [nvox, 5] = peak amplitude of fixels 0 and 1, direction of
    fixel 0 (when asked for)."""

    def __init__(self, count, offset, fixels, summary=None):
        self.count, self.offset, self.fixels, self.summary = count, offset, fixels, summary

    @property
    def directions(self):
        return self.fixels[:, 0:3]

    @property
    def afd(self):
        return self.fixels[:, 3]

    @property
    def peak_amp(self):
        return self.fixels[:, 4]

    @property
    def disp(self):
        """integral / maximal peak amplitude in single precision (fod2fixel.cpp:272-276)."""
        return self.fixels[:, 3] / self.fixels[:, 4]


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class Segmenter:
    def __init__(self, n_sh, integral_threshold=0.0, peak_value_threshold=0.1, lobe_merge_ratio=1.0,
                 no_thresholds=False, maxnum=0, dir_from_peak=False, directions=None):
        _register()
        lib = _lib.load()
        opt = Options()
        lib.mrfixel_options_default(C.byref(opt))
        opt.n_sh = int(n_sh)
        # load_fmls_thresholds (fmls.cpp:56-91): -fmls_no_thresholds overrides both thresholds
        if no_thresholds:
            if integral_threshold != 0.0:
                warnings.warn("Option -fmls_integral ignored: -fmls_no_thresholds overrides this")
            if peak_value_threshold != 0.1:
                warnings.warn("Option -fmls_peak_value ignored: -fmls_no_thresholds overrides this")
            integral_threshold = peak_value_threshold = 0.0
        opt.integral_threshold = float(integral_threshold)
        opt.peak_value_threshold = float(peak_value_threshold)
        opt.lobe_merge_ratio = float(lobe_merge_ratio)
        opt.maxnum = int(maxnum)
        opt.dir_from_peak = 1 if dir_from_peak else 0
        self._keep = None
        if directions is not None:
            d = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 2)
            self._keep = d
            opt.directions = d.ctypes.data_as(_D)
            opt.n_directions = d.shape[0]
        h = _P()
        check(lib.mrfixel_create(C.byref(opt), C.byref(h)))
        self._h = h
        self.n_sh = int(n_sh)

    def segment(self, fod, mask=None, summary=False, capacity=None):
        """fod: [nvox, >= n_sh] float32 (torch device tensor or numpy); mask: [nvox] uint8 or None.
        Returns Fixels (torch tensors on the device for a torch input, numpy arrays otherwise)."""
        lib = _lib.load()
        if _is_torch(fod):
            import torch
            assert fod.is_cuda and fod.dtype == torch.float32 and fod.is_contiguous() and fod.dim() == 2
            nvox, ld = fod.shape
            dev = fod.device
            if mask is not None:
                assert mask.is_cuda and mask.dtype == torch.uint8 and mask.numel() == nvox
            cap = int(capacity) if capacity is not None else max(64, 4 * nvox)
            count = torch.empty(nvox, dtype=torch.int32, device=dev)
            offset = torch.empty(nvox, dtype=torch.int32, device=dev)
            total = torch.zeros(1, dtype=torch.int64, device=dev)
            summ = torch.empty((nvox, 5), dtype=torch.float32, device=dev) if summary else None
            stream = torch.cuda.current_stream(dev).cuda_stream
            while True:
                fix = torch.empty((cap, 5), dtype=torch.float32, device=dev)
                check(lib.mrfixel_segment(self._h, fod.data_ptr(), ld, mask.data_ptr() if mask is not None else None,
                                          nvox, count.data_ptr(), offset.data_ptr(), fix.data_ptr(), cap,
                                          summ.data_ptr() if summ is not None else None, total.data_ptr(), stream))
                n = int(total.item())
                if n <= cap:
                    return Fixels(count, offset, fix[:n], summ)
                cap = n
        fod = np.ascontiguousarray(np.atleast_2d(fod), dtype=np.float32)
        nvox, ld = fod.shape
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        cap = int(capacity) if capacity is not None else max(64, 4 * nvox)
        count = np.zeros(nvox, dtype=np.uint32)
        offset = np.zeros(nvox, dtype=np.uint32)
        summ = np.zeros((nvox, 5), dtype=np.float32) if summary else None
        total = C.c_uint64()
        while True:
            fix = np.zeros((cap, 5), dtype=np.float32)
            check(lib.mrfixel_segment_host(self._h, fod.ctypes.data, ld, m.ctypes.data if m is not None else None,
                                           nvox, count.ctypes.data, offset.ctypes.data, fix.ctypes.data, cap,
                                           summ.ctypes.data if summ is not None else None, C.byref(total)))
            if total.value <= cap:
                return Fixels(count, offset, fix[:total.value], summ)
            cap = int(total.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrfixel_destroy(h)
        self._h = None
