"""Launch-shape autotuners and captured solve steps (libmrcsd.so, include/mrtune.h).

    report = autotune_csd(solver, dw)                          # persistent grid of solver.run
    report = autotune_csd(solver, dw, warm_neg=neg)            # ... of the warm-started runs
    report = autotune_icls(problem, b)                          # ICLS first-pass capacity
    step = SolveGraph(solver, dw, fod, net=net, dwis=dwis, b0=b0)   # inference + solve as a HIP graph
    step.launch()                                              # one replay on the current stream

All tensors are torch CUDA tensors; the graph keeps their device pointers, so they must stay alive
(and in place) for the graph's lifetime.
"""
import ctypes as C

from . import _lib
from ._lib import check

_P = C.c_void_p
_registered = False


class Report(C.Structure):
    _fields_ = [("n_candidates", C.c_int), ("values", C.c_int * 8), ("ms", C.c_float * 8), ("chosen", C.c_int),
                ("full", C.c_int)]

    def as_dict(self):
        return {"candidates": [{"value": int(self.values[i]), "ms": float(self.ms[i])} for i in range(self.n_candidates)],
                "chosen": int(self.chosen), "full": int(self.full)}


def _register():
    global _registered
    if _registered:
        return
    _lib.register("mrcsd_set_groups", C.c_int, [_P, C.c_int, C.c_int])
    _lib.register("mrcsd_get_groups", C.c_int, [_P, C.c_int])
    _lib.register("mrcsd_autotune", C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P, C.POINTER(Report)])
    _lib.register("mricls_autotune", C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, C.POINTER(Report)])
    _lib.register("mrcsd_graph_create", C.c_int, [_P, _P, _P, C.c_int, _P, C.c_size_t, _P, _P, C.c_int, _P, C.c_int,
                                                  _P, _P, C.POINTER(_P)])
    _lib.register("mrcsd_graph_launch", C.c_int, [_P, _P])
    _lib.register("mrcsd_graph_destroy", None, [_P])
    _registered = True


def _stream(t):
    import torch
    return torch.cuda.current_stream(t.device).cuda_stream


def autotune_csd(solver, dw, warm_neg=None, budget=6, valid=None):
    """Times solver.run on dw ([nvox, nvol] CUDA float32) at several persistent grid sizes and keeps
    the fastest for later runs (the warm-started instantiation when warm_neg is given)."""
    import torch
    _register()
    fod = torch.empty((dw.shape[0], solver.n_sh), dtype=torch.float32, device=dw.device)
    warm = None
    if warm_neg is not None:
        warm = _lib.Warm(None, warm_neg.data_ptr(), int(budget))
    rep = Report()
    check(_lib.load().mrcsd_autotune(solver._h, dw.data_ptr(), dw.shape[1],
                                     valid.data_ptr() if valid is not None else None, dw.shape[0],
                                     C.byref(warm) if warm is not None else None, fod.data_ptr(), _stream(dw),
                                     C.byref(rep)))
    return rep.as_dict()


def set_groups(solver, groups, warm=False):
    _register()
    check(_lib.load().mrcsd_set_groups(solver._h, int(bool(warm)), int(groups)))


def autotune_icls(problem, b, valid=None):
    """Times problem.solve on b ([nvox, >= nmeas] CUDA float32) with each first-pass capacity and
    keeps the fastest."""
    import torch
    _register()
    x = torch.empty((b.shape[0], problem.n), dtype=torch.float32, device=b.device)
    rep = Report()
    check(_lib.load().mricls_autotune(problem._h, b.data_ptr(), b.shape[1],
                                      valid.data_ptr() if valid is not None else None, b.shape[0], x.data_ptr(),
                                      _stream(b), C.byref(rep)))
    return rep.as_dict()


class SolveGraph:
    """One solve step captured as a HIP graph: [fused warm-start inference +] mrcsd_run."""

    def __init__(self, solver, dw, fod, net=None, dwis=None, b0=None, negmask=None, budget=6, valid=None,
                 stats=None):
        import torch
        _register()
        self._keep = [solver, dw, fod, net, dwis, b0, valid, stats]
        if net is not None and negmask is None:
            negmask = torch.empty((dw.shape[0], (solver.n_hr + 63) // 64), dtype=torch.int64, device=dw.device)
        self.negmask = negmask
        self._outs = _lib.Outputs(None, None, None, stats.data_ptr() if stats is not None else None)
        h = C.c_void_p()
        check(_lib.load().mrcsd_graph_create(
            solver._h, net._h if net is not None else None, dw.data_ptr(), dw.stride(0),
            valid.data_ptr() if valid is not None else None, dw.shape[0],
            dwis.data_ptr() if dwis is not None else None, b0.data_ptr() if b0 is not None else None,
            b0.numel() if b0 is not None else 0, negmask.data_ptr() if negmask is not None else None, int(budget),
            fod.data_ptr(), C.byref(self._outs), C.byref(h)))
        self._h = h
        self._dev = dw.device

    def launch(self):
        import torch
        check(_lib.load().mrcsd_graph_launch(self._h, torch.cuda.current_stream(self._dev).cuda_stream))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrcsd_graph_destroy(h)
        self._h = None
