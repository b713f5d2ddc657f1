"""Python mirror of the CSD call surface, backed by libmrcsd.so.

    shared = CSDShared(grad, response, lmax=0, ...)   # DWI::SDeconv::CSD::Shared (csd.h:49-212)
    solver = CSD(shared)                              # GPU solver context (device matrices)
    fod, iters, nneg = solver.run(dw, valid)          # dwi2fod's CSD_Processor over a batch

`run` accepts torch CUDA tensors (the solve stays on the device, asynchronous on the current
torch stream) or numpy arrays (copied in and out).  Errors raise MrcsdError with the message of
the C layer, which mirrors the reference's exception texts.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr

MATRIX_IDS = {"rconv": 0, "M": 1, "Mt_M": 2, "HR_trans": 3, "dwis": 4, "grad": 5}


class CSDShared:
    def __init__(self, grad, response, lmax=0, filter=None, directions=None, neg_lambda=1.0,
                 norm_lambda=1.0, threshold=0.0, niter=50, bvalue_scaling=0, shells=None):
        lib = _lib.load()
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        if grad.ndim != 2 or grad.shape[1] != 4:
            raise ValueError("gradient table must be [nvol, 4]")
        resp = np.ascontiguousarray(np.ravel(response), dtype=np.float64)
        opt = _lib.Options()
        lib.mrcsd_options_default(C.byref(opt))
        opt.lmax = int(lmax)
        self._keep = []
        if filter is not None:
            f = np.ascontiguousarray(filter, dtype=np.float64)
            self._keep.append(f)
            opt.filter = dptr(f)
            opt.n_filter = f.size
        if directions is not None:
            d = np.ascontiguousarray(directions, dtype=np.float64)
            self._keep.append(d)
            opt.directions = dptr(d)
            opt.n_directions = d.shape[0]
        if shells is not None:
            sh = np.ascontiguousarray(shells, dtype=np.float64).reshape(-1)
            self._keep.append(sh)
            opt.shells = dptr(sh)
            opt.n_shells = sh.size
        opt.neg_lambda = float(neg_lambda)
        opt.norm_lambda = float(norm_lambda)
        opt.threshold = float(threshold)
        opt.niter = int(niter)
        opt.bvalue_scaling = int(bvalue_scaling)
        h = C.c_void_p()
        check(lib.mrcsd_shared_create(dptr(grad), grad.shape[0], dptr(resp), resp.size, C.byref(opt),
                                      C.byref(h)))
        self._h = h
        self.nvol = grad.shape[0]
        info = _lib.SharedInfo()
        check(lib.mrcsd_shared_info_get(h, C.byref(info)))
        for name, _ in info._fields_:
            setattr(self, name, getattr(info, name))

    def nSH(self):
        return self.n_sh

    def matrix(self, name):
        shapes = {"rconv": (self.n_init, self.n_dw), "M": (self.n_dw, self.n_sh),
                  "Mt_M": (self.n_sh, self.n_sh), "HR_trans": (self.n_hr, self.n_sh),
                  "dwis": (self.n_dw,), "grad": (self.nvol, 4)}
        out = np.empty(shapes[name], dtype=np.float64)
        check(_lib.load().mrcsd_shared_matrix(self._h, MATRIX_IDS[name], dptr(out), out.size))
        return out

    @property
    def dwis(self):
        return self.matrix("dwis").astype(np.int64)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrcsd_shared_destroy(h)
        self._h = None


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class CSD:
    """GPU solver built from a CSDShared (or from raw matrices via CSD.from_params)."""

    def __init__(self, shared=None, _handle=None, n_sh=None):
        lib = _lib.load()
        if _handle is None:
            h = C.c_void_p()
            check(lib.mrcsd_create_from_shared(shared._h, C.byref(h)))
            self._h = h
            self.n_sh = shared.n_sh
            self.n_hr = shared.n_hr
            self.niter = shared.niter
        else:
            self._h = _handle
            self.n_sh = n_sh

    @classmethod
    def from_params(cls, rconv, M, MtM, HR_trans, dwis, threshold, niter):
        lib = _lib.load()
        rconv = np.ascontiguousarray(rconv, dtype=np.float64)
        M = np.ascontiguousarray(M, dtype=np.float64)
        MtM = np.ascontiguousarray(MtM, dtype=np.float64)
        HR = np.ascontiguousarray(HR_trans, dtype=np.float64)
        dw = np.ascontiguousarray(dwis, dtype=np.int32)
        p = _lib.Params(M.shape[0], M.shape[1], rconv.shape[0], HR.shape[0], dptr(rconv), dptr(M),
                        dptr(MtM), dptr(HR), dw.ctypes.data_as(C.POINTER(C.c_int)), float(threshold),
                        int(niter))
        h = C.c_void_p()
        check(lib.mrcsd_create(C.byref(p), C.byref(h)))
        obj = cls(_handle=h, n_sh=M.shape[1])
        obj.niter = int(niter)
        obj.n_hr = HR.shape[0]
        return obj

    def warm_supported(self):
        """Whether the solver accepts a warm-start proposal (threshold 0 and a positive definite
        normal matrix, where the certified fixed point is the plain solver's answer)."""
        return bool(_lib.load().mrcsd_warm_supported(self._h))

    def run(self, dw, valid=None, iters=True, nneg=True, stats=None, warm_F0=None, fod=None,
            negmask=False, warm_neg=None, warm_budget=6):
        """Solve every row of dw ([nvox, nvol] float32).  Returns (fod, iters, nneg), plus the
        neg-set bitmask [nvox, words] when negmask=True (device path only)."""
        lib = _lib.load()
        if _is_torch(dw):
            import torch
            assert dw.is_cuda and dw.dtype == torch.float32 and dw.is_contiguous()
            nvox, ld = dw.shape
            dev = dw.device
            if fod is None:
                fod = torch.empty((nvox, self.n_sh), dtype=torch.float32, device=dev)
            it = torch.empty(nvox, dtype=torch.int16, device=dev) if iters else None
            nn = torch.empty(nvox, dtype=torch.int16, device=dev) if nneg else None
            if valid is not None:
                assert valid.dtype == torch.uint8 and valid.is_cuda and valid.numel() == nvox
            warm = None
            if warm_F0 is not None or warm_neg is not None:
                if warm_F0 is not None:
                    assert warm_F0.is_cuda and warm_F0.dtype == torch.float32 and warm_F0.shape == fod.shape
                if warm_neg is not None:
                    assert warm_neg.is_cuda and warm_neg.dtype == torch.int64
                    assert warm_neg.shape == (nvox, (self.n_hr + 63) // 64)
                warm = _lib.Warm(warm_F0.data_ptr() if warm_F0 is not None else None,
                                 warm_neg.data_ptr() if warm_neg is not None else None, int(warm_budget))
            nm = None
            if negmask:
                nm = torch.empty((nvox, (self.n_hr + 63) // 64), dtype=torch.int64, device=dev)
            outs = _lib.Outputs(it.data_ptr() if it is not None else None,
                                nn.data_ptr() if nn is not None else None,
                                nm.data_ptr() if nm is not None else None,
                                stats.data_ptr() if stats is not None else None)
            stream = torch.cuda.current_stream(dev).cuda_stream
            check(lib.mrcsd_run(self._h, dw.data_ptr(), ld, valid.data_ptr() if valid is not None else None,
                                nvox, fod.data_ptr(), C.byref(outs),
                                C.byref(warm) if warm is not None else None, stream))
            if negmask:
                return fod, it, nn, nm
            return fod, it, nn
        dw = np.ascontiguousarray(dw, dtype=np.float32)
        nvox, ld = dw.shape
        out = np.empty((nvox, self.n_sh), dtype=np.float32)
        it = np.empty(nvox, dtype=np.uint16) if iters else None
        nn = np.empty(nvox, dtype=np.uint16) if nneg else None
        v = None
        if valid is not None:
            v = np.ascontiguousarray(valid, dtype=np.uint8)
        check(lib.mrcsd_run_host(self._h, dw.ctypes.data, ld, v.ctypes.data if v is not None else None,
                                 nvox, out.ctypes.data, it.ctypes.data if it is not None else None,
                                 nn.ctypes.data if nn is not None else None))
        return out, it, nn

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrcsd_destroy(h)
        self._h = None
