"""ctypes binding of libmrcsd.so (the C ABI declared in include/mrcsd.h).

The library is built in-tree by mrtrix3_amd/build.py.  Loading fails loudly when it is missing:
there is no fallback implementation of the solver anywhere in this package.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmrcsd.so")

_lib = None


class MrcsdError(RuntimeError):
    pass


class Options(C.Structure):
    _fields_ = [("lmax", C.c_int), ("filter", C.POINTER(C.c_double)), ("n_filter", C.c_int),
                ("directions", C.POINTER(C.c_double)), ("n_directions", C.c_int),
                ("neg_lambda", C.c_double), ("norm_lambda", C.c_double),
                ("threshold", C.c_double), ("niter", C.c_int), ("bvalue_scaling", C.c_int),
                ("shells", C.POINTER(C.c_double)), ("n_shells", C.c_int)]


class SharedInfo(C.Structure):
    _fields_ = [("n_dw", C.c_int), ("n_sh", C.c_int), ("n_init", C.c_int), ("n_hr", C.c_int),
                ("lmax", C.c_int), ("lmax_response", C.c_int), ("lmax_data", C.c_int),
                ("niter", C.c_int), ("threshold", C.c_double), ("norm_lambda", C.c_double)]


class Params(C.Structure):
    _fields_ = [("n_dw", C.c_int), ("n_sh", C.c_int), ("n_init", C.c_int), ("n_hr", C.c_int),
                ("rconv", C.POINTER(C.c_double)), ("M", C.POINTER(C.c_double)),
                ("MtM", C.POINTER(C.c_double)), ("HR_trans", C.POINTER(C.c_double)),
                ("dwis", C.POINTER(C.c_int)), ("threshold", C.c_double), ("niter", C.c_int)]


class Warm(C.Structure):
    _fields_ = [("F0", C.c_void_p), ("neg", C.c_void_p), ("budget", C.c_int)]


class Outputs(C.Structure):
    _fields_ = [("iters", C.c_void_p), ("nneg", C.c_void_p), ("negmask", C.c_void_p),
                ("stats", C.c_void_p)]


_P = C.c_void_p
_SIG = {
    "mrcsd_last_error": (C.c_char_p, []),
    "mrcsd_version": (C.c_char_p, []),
    "mrcsd_options_default": (None, [C.POINTER(Options)]),
    "mrcsd_shared_create": (C.c_int, [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double), C.c_int,
                                      C.POINTER(Options), C.POINTER(_P)]),
    "mrcsd_shared_destroy": (None, [_P]),
    "mrcsd_shared_info_get": (C.c_int, [_P, C.POINTER(SharedInfo)]),
    "mrcsd_shared_matrix": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.c_size_t]),
    "mrcsd_create": (C.c_int, [C.POINTER(Params), C.POINTER(_P)]),
    "mrcsd_create_from_shared": (C.c_int, [_P, C.POINTER(_P)]),
    "mrcsd_destroy": (None, [_P]),
    "mrcsd_warm_supported": (C.c_int, [_P]),
    "mrcsd_run": (C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, C.POINTER(Outputs), C.POINTER(Warm), _P]),
    "mrcsd_run_host": (C.c_int, [_P, _P, C.c_int, _P, C.c_size_t, _P, _P, _P]),
}


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MrcsdError("libmrcsd.so is not built (%s); run `python -c 'import __graft_entry__ as g; "
                         "g.build()'` or `python mrtrix3_amd/build.py`" % LIB_PATH)
    # torch ships its own libamdhip64 / libhsa-runtime64 under the same sonames as ROCm's.  Loading
    # torch first makes libmrcsd bind to that runtime, so the process holds ONE HIP runtime and torch
    # device pointers are valid here; loaded the other way round, the process maps two runtimes and
    # the second one to initialise reports no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIG.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def register(name, restype, argtypes):
    lib = load()
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    _SIG[name] = (restype, argtypes)
    return fn


def check(rc):
    if rc != 0:
        msg = load().mrcsd_last_error().decode(errors="replace")
        raise MrcsdError("mrcsd error %d: %s" % (rc, msg))
    return rc


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(C.POINTER(C.c_double))
