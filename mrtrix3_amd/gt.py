"""Global tractography's external energy on the GPU (libmrcsd.so, include/mrgt.h):
MR::DWI::Tractography::GT::ExternalEnergyComputer (src/dwi/tractography/GT/externalenergy.cpp).

    E = ExternalEnergy(grad, resp_wm, resp_iso=[gm, csf], lmax=8, weight=1.0, ppot=0.0)
    eext, fiso, total = E.energy(dwi, tod)       # calcEnergy over every row, and the sum

dwi [nvox, >= nvol] and tod [nvox, >= NforL(lmax)] are float32 torch CUDA tensors (the result stays
on the device) or numpy arrays.  Each voxel's isotropic fractions come from an NNLS fit on the GPU's
batched ICLS kernel in double precision, as the reference's Solver<double> does.
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
    _lib.register("mrgt_create", C.c_int, [_D, C.c_int, _D, C.c_int, C.c_int, _D, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.POINTER(_P)])
    _lib.register("mrgt_destroy", None, [_P])
    _lib.register("mrgt_kernels", C.c_int, [_D, C.c_int, _D, C.c_int, C.c_int, _D, C.c_int, C.c_int, C.c_double,
                                            _D, _D])
    _lib.register("mrgt_dims", C.c_int, [_P, _I, _I, _I])
    _lib.register("mrgt_matrix", C.c_int, [_P, C.c_int, _D, C.c_size_t])
    _lib.register("mrgt_energy", C.c_int, [_P, _P, C.c_int, _P, C.c_int, C.c_size_t, _P, _P, _P, _P])
    _lib.register("mrgt_energy_host", C.c_int, [_P, _P, C.c_int, _P, C.c_int, C.c_size_t, _P, _P, _P])
    _registered = True


def kernels(grad, resp_wm, resp_iso=(), lmax=8, weight=1.0):
    """ExternalEnergyComputer's K and Ak from the library's host set-up (no GPU needed)."""
    _register()
    g = np.ascontiguousarray(grad, dtype=np.float64)
    wm = np.ascontiguousarray(np.atleast_2d(resp_wm), dtype=np.float64)
    iso = np.ascontiguousarray(np.asarray(resp_iso, dtype=np.float64).reshape(len(resp_iso), -1)) \
        if len(resp_iso) else None
    K = np.empty((g.shape[0], (lmax + 1) * (lmax + 2) // 2))
    Ak = np.empty((g.shape[0], len(resp_iso) + 1))
    check(_lib.load().mrgt_kernels(dptr(g), g.shape[0], dptr(wm), wm.shape[0], wm.shape[1],
                                   dptr(iso) if iso is not None else None, len(resp_iso), int(lmax), float(weight),
                                   dptr(K), dptr(Ak)))
    return K, Ak


class ExternalEnergy:
    def __init__(self, grad, resp_wm, resp_iso=(), lmax=8, weight=1.0, ppot=0.0):
        _register()
        g = np.ascontiguousarray(grad, dtype=np.float64)
        wm = np.ascontiguousarray(np.atleast_2d(resp_wm), dtype=np.float64)
        iso = np.ascontiguousarray(np.asarray(resp_iso, dtype=np.float64).reshape(len(resp_iso), -1)) \
            if len(resp_iso) else None
        h = C.c_void_p()
        check(_lib.load().mrgt_create(dptr(g), g.shape[0], dptr(wm), wm.shape[0], wm.shape[1],
                                      dptr(iso) if iso is not None else None, len(resp_iso), int(lmax),
                                      float(weight), float(ppot), C.byref(h)))
        self._h = h
        d = [C.c_int() for _ in range(3)]
        check(_lib.load().mrgt_dims(h, *[C.byref(v) for v in d]))
        self.nrows, self.ncols, self.nf = [v.value for v in d]

    def matrix(self, name):
        shape = {"K": (self.nrows, self.ncols), "Ak": (self.nrows, self.nf + 1)}[name]
        out = np.empty(shape)
        check(_lib.load().mrgt_matrix(self._h, 0 if name == "K" else 1, dptr(out), out.size))
        return out

    def energy(self, dwi, tod):
        """Returns (eext [nvox] float32, fiso [nvox, nf] float32 or None, total energy)."""
        lib = _lib.load()
        if type(dwi).__module__.startswith("torch"):
            import torch
            assert dwi.is_cuda and tod.is_cuda and dwi.dtype == torch.float32 and tod.dtype == torch.float32
            assert dwi.is_contiguous() and tod.is_contiguous() and dwi.shape[0] == tod.shape[0]
            n = dwi.shape[0]
            e = torch.empty(n, dtype=torch.float32, device=dwi.device)
            f = torch.empty((n, self.nf), dtype=torch.float32, device=dwi.device) if self.nf else None
            total = torch.empty(1, dtype=torch.float64, device=dwi.device)
            stream = torch.cuda.current_stream(dwi.device).cuda_stream
            check(lib.mrgt_energy(self._h, dwi.data_ptr(), dwi.shape[1], tod.data_ptr(), tod.shape[1], n,
                                  e.data_ptr(), f.data_ptr() if f is not None else None, total.data_ptr(), stream))
            return e, f, total
        dwi = np.ascontiguousarray(np.atleast_2d(dwi), dtype=np.float32)
        tod = np.ascontiguousarray(np.atleast_2d(tod), dtype=np.float32)
        n = dwi.shape[0]
        e = np.empty(n, dtype=np.float32)
        f = np.empty((n, self.nf), dtype=np.float32) if self.nf else None
        total = C.c_double()
        check(lib.mrgt_energy_host(self._h, dwi.ctypes.data, dwi.shape[1], tod.ctypes.data, tod.shape[1], n,
                                   e.ctypes.data, f.ctypes.data if f is not None else None, C.byref(total)))
        return e, f, total.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrgt_destroy(h)
        self._h = None
