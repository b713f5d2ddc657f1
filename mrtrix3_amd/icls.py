"""Python mirror of Math::ICLS and MSMT_CSD::Shared, backed by libmrcsd.so (include/mricls.h).

    problem = IclsProblem(H, A, t=None, num_eq=0, ...)       # ICLS::Problem (:91-152) + GPU copy
    x, iters, nactive = problem.solve(b)                    # Solver::operator() per row of b

    shared = MSMTShared(grad, [wm, gm, csf], lmax=None)     # MSMT_CSD::Shared (msmt_csd.h:47-257)
    problem = shared.problem()                              # its ICLS problem on the GPU
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr

_P = C.c_void_p
_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)
_registered = False


def _register():
    global _registered
    if _registered:
        return
    _lib.register("mricls_create", C.c_int, [_D, C.c_int, C.c_int, _D, C.c_int, _D, C.c_int, C.c_double,
                                            C.c_double, C.c_int, C.c_double, C.c_int, C.POINTER(_P)])
    _lib.register("mricls_destroy", None, [_P])
    _lib.register("mricls_dims", C.c_int, [_P, _I, _I, _I, _I])
    _lib.register("mricls_solve", C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P, _P, _P])
    _lib.register("mricls_solve_host", C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P])
    _lib.register("mricls_solve_f64", C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P, _P, _P])
    _lib.register("mricls_solve_host_f64", C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P])
    _lib.register("mricls_set_capacity", C.c_int, [_P, C.c_int])
    _lib.register("mricls_get_capacity", C.c_int, [_P])
    _lib.register("mrmsmt_shared_create", C.c_int, [_D, C.c_int, _D, _I, _I, C.c_int, _I, C.c_int, _D,
                                                   C.c_int, C.c_double, C.c_double, _D, C.c_int, C.POINTER(_P)])
    _lib.register("mrmsmt_shared_destroy", None, [_P])
    _lib.register("mrmsmt_shared_info", C.c_int, [_P, _I, _I, _I, _I])
    _lib.register("mrmsmt_shared_matrix", C.c_int, [_P, C.c_int, _D, C.c_size_t])
    _lib.register("mrmsmt_icls_create", C.c_int, [_P, C.POINTER(_P)])
    _registered = True


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class IclsProblem:
    def __init__(self, H=None, A=None, t=None, num_eq=0, sol_reg=0.0, con_reg=0.0, max_iter=0, tol=0.0,
                 standard_form=False, _handle=None):
        _register()
        lib = _lib.load()
        if _handle is not None:
            self._h = _handle
        else:
            H = np.ascontiguousarray(H, dtype=np.float64)
            A = np.ascontiguousarray(A, dtype=np.float64)
            tt = None if t is None or np.size(t) == 0 else np.ascontiguousarray(t, dtype=np.float64)
            h = C.c_void_p()
            check(lib.mricls_create(dptr(H), H.shape[0], H.shape[1], dptr(A), A.shape[0],
                                    dptr(tt) if tt is not None else None, int(num_eq), float(sol_reg),
                                    float(con_reg), int(max_iter), float(tol), int(bool(standard_form)),
                                    C.byref(h)))
            self._h = h
        d = [C.c_int() for _ in range(4)]
        check(lib.mricls_dims(self._h, *[C.byref(v) for v in d]))
        self.nmeas, self.n, self.ncon, self.max_niter = [v.value for v in d]

    @property
    def capacity(self):
        """Active-set slots of the solver's first pass (voxels needing more are re-solved with 64,
        then 128 slots)."""
        return int(_lib.load().mricls_get_capacity(self._h))

    @capacity.setter
    def capacity(self, slots):
        check(_lib.load().mricls_set_capacity(self._h, int(slots)))

    def solve(self, b, valid=None, stats=None, x=None):
        """b: [nvox, >= nmeas] float32, or float64 for double-precision measurements and solutions
        (torch CUDA tensor or numpy).  Returns (x, iters, nactive), x in b's precision."""
        lib = _lib.load()
        if _is_torch(b):
            import torch
            assert b.is_cuda and b.dtype in (torch.float32, torch.float64) and b.is_contiguous()
            f64 = b.dtype == torch.float64
            nvox, ld = b.shape
            if x is None:
                x = torch.empty((nvox, self.n), dtype=b.dtype, device=b.device)
            assert x.dtype == b.dtype
            it = torch.empty(nvox, dtype=torch.int16, device=b.device)
            na = torch.empty(nvox, dtype=torch.int16, device=b.device)
            stream = torch.cuda.current_stream(b.device).cuda_stream
            fn = lib.mricls_solve_f64 if f64 else lib.mricls_solve
            check(fn(self._h, b.data_ptr(), ld, valid.data_ptr() if valid is not None else None, nvox,
                     x.data_ptr(), it.data_ptr(), na.data_ptr(), stats.data_ptr() if stats is not None else None,
                     stream))
            return x, it, na
        f64 = np.asarray(b).dtype == np.float64
        dt = np.float64 if f64 else np.float32
        b = np.ascontiguousarray(np.atleast_2d(b), dtype=dt)
        nvox, ld = b.shape
        out = np.empty((nvox, self.n), dtype=dt)
        it = np.empty(nvox, dtype=np.uint16)
        na = np.empty(nvox, dtype=np.uint16)
        v = None if valid is None else np.ascontiguousarray(valid, dtype=np.uint8)
        fn = lib.mricls_solve_host_f64 if f64 else lib.mricls_solve_host
        check(fn(self._h, b.ctypes.data, ld, v.ctypes.data if v is not None else None, nvox, out.ctypes.data,
                 it.ctypes.data, na.ctypes.data))
        return out, it, na

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mricls_destroy(h)
        self._h = None


class MSMTShared:
    def __init__(self, grad, responses, lmax=None, directions=None, norm_lambda=-1.0, neg_lambda=-1.0, shells=None):
        _register()
        lib = _lib.load()
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        resp = [np.atleast_2d(np.asarray(r, dtype=np.float64)) for r in responses]
        flat = np.ascontiguousarray(np.concatenate([r.ravel() for r in resp]))
        rows = np.array([r.shape[0] for r in resp], dtype=np.int32)
        cols = np.array([r.shape[1] for r in resp], dtype=np.int32)
        lm = None if lmax is None else np.ascontiguousarray(lmax, dtype=np.int32)
        d = None if directions is None else np.ascontiguousarray(directions, dtype=np.float64)
        sh = None if shells is None else np.ascontiguousarray(shells, dtype=np.float64).reshape(-1)
        h = C.c_void_p()
        check(lib.mrmsmt_shared_create(dptr(grad), grad.shape[0], dptr(flat), rows.ctypes.data_as(_I),
                                       cols.ctypes.data_as(_I), len(resp),
                                       lm.ctypes.data_as(_I) if lm is not None else None,
                                       0 if lm is None else lm.size, dptr(d) if d is not None else None, 0 if d is None else d.shape[0],
                                       float(norm_lambda), float(neg_lambda),
                                       dptr(sh) if sh is not None else None, 0 if sh is None else sh.size, C.byref(h)))
        self._h = h
        self.nvol = grad.shape[0]
        self.ntissue = len(resp)
        ns, npar, nc = C.c_int(), C.c_int(), C.c_int()
        lo = np.zeros(len(resp), dtype=np.int32)
        check(lib.mrmsmt_shared_info(h, C.byref(ns), C.byref(npar), C.byref(nc), lo.ctypes.data_as(_I)))
        self.nshell, self.nparams, self.ncon = ns.value, npar.value, nc.value
        self.lmax = [int(v) for v in lo]

    def matrix(self, name):
        which, shape = {"C": (0, (self.nvol, self.nparams)), "A": (1, (self.ncon, self.nparams))}[name]
        out = np.empty(shape, dtype=np.float64)
        check(_lib.load().mrmsmt_shared_matrix(self._h, which, dptr(out), out.size))
        return out

    def problem(self):
        h = C.c_void_p()
        check(_lib.load().mrmsmt_icls_create(self._h, C.byref(h)))
        return IclsProblem(_handle=h)

    def split(self, x):
        """Per-tissue views of solutions [nvox, nparams] (dwi2fod.cpp:209-214)."""
        out, j = [], 0
        for l in self.lmax:
            n = (l + 1) * (l + 2) // 2
            out.append(x[:, j:j + n])
            j += n
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrmsmt_shared_destroy(h)
        self._h = None
