"""ctypes binding of the CPU oracle (oracle/_build/liborc*.so) — TEST INFRASTRUCTURE ONLY.

Used by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's CPU baseline leg.
The product (mrtrix3_amd/) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_libs = {}


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


class CsdShared(C.Structure):
    _fields_ = [("n_dw", C.c_int), ("n_sh", C.c_int), ("n_init", C.c_int), ("n_hr", C.c_int),
                ("lmax", C.c_int), ("lmax_response", C.c_int), ("lmax_data", C.c_int),
                ("niter", C.c_int), ("threshold", C.c_double), ("norm_lambda", C.c_double),
                ("rconv", C.POINTER(C.c_double)), ("M", C.POINTER(C.c_double)),
                ("MtM", C.POINTER(C.c_double)), ("HR_trans", C.POINTER(C.c_double))]


class IclsProblem(C.Structure):
    _fields_ = [("nmeas", C.c_int), ("n", C.c_int), ("ncon", C.c_int), ("num_eq", C.c_int),
                ("max_niter", C.c_int), ("lambda_min_norm", C.c_double), ("tol", C.c_double),
                ("chol", C.POINTER(C.c_double)), ("b2d", C.POINTER(C.c_double)),
                ("B", C.POINTER(C.c_double)), ("t", C.POINTER(C.c_double)),
                ("H", C.POINTER(C.c_double))]


D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
FP = C.POINTER(C.c_float)
_SIG = {
    "orc_nforl": (C.c_int, [C.c_int]),
    "orc_lforn": (C.c_int, [C.c_int]),
    "orc_plm_sph": (None, [D, C.c_int, C.c_int, C.c_double]),
    "orc_sh_transform": (None, [D, C.c_int, C.c_int, D]),
    "orc_zsh2rh": (None, [D, C.c_int, D]),
    "orc_pinv": (C.c_int, [D, C.c_int, C.c_int, D]),
    "orc_llt_solve": (C.c_int, [D, C.c_int, D, D]),
    "orc_dw_scheme_normalise": (C.c_int, [D, C.c_int, C.c_int]),
    "orc_gen_direction_matrix": (None, [D, I, C.c_int, D]),
    "orc_shells_cluster": (C.c_int, [D, C.c_int, C.c_double, C.c_double, I]),
    "orc_csd_select_volumes": (C.c_int, [D, C.c_int, C.c_double, C.c_double, I, I]),
    "orc_csd_init": (C.c_int, [D, C.c_int, D, C.c_int, D, C.c_int, C.c_int, D, C.c_int, C.c_double,
                               C.c_double, C.c_double, C.c_int, C.POINTER(C.POINTER(CsdShared)),
                               C.c_char_p, C.c_int]),
    "orc_csd_free": (None, [C.POINTER(CsdShared)]),
    "orc_csd_solve": (C.c_int, [C.POINTER(CsdShared), D, D, I, I, I]),
    "orc_csd_solve_from": (C.c_int, [C.POINTER(CsdShared), D, D, D, I, I, I]),
    "orc_csd_run_volume": (C.c_int, [C.POINTER(CsdShared), C.c_void_p, C.c_int, I, C.c_void_p, C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "orc_icls_problem_create": (C.c_int, [D, C.c_int, C.c_int, D, C.c_int, D, C.c_int, C.c_double,
                                          C.c_double, C.c_int, C.c_double, C.c_int,
                                          C.POINTER(C.POINTER(IclsProblem)), C.c_char_p, C.c_int]),
    "orc_icls_problem_free": (None, [C.POINTER(IclsProblem)]),
    "orc_icls_solve": (C.c_int, [C.POINTER(IclsProblem), D, D, C.c_void_p]),
    "orc_icls_run_batch": (C.c_int, [C.POINTER(IclsProblem), C.c_void_p, C.c_size_t, C.c_void_p,
                                     C.c_int, C.c_void_p, C.c_void_p]),
    "orc_msmt_build": (C.c_int, [D, C.c_int, I, C.c_int, D, I, C.c_int, I, D, C.c_int, D, D, I, I]),
    "orc_real_name": (C.c_char_p, []),
    "orc_time_syrk_llt": (C.c_int, [C.POINTER(CsdShared), C.c_int, C.c_int, D, D]),
    "orc_plm_sph_f": (None, [FP, C.c_int, C.c_int, C.c_float, C.c_float]),
    "orc_sh_derivatives_f": (None, [FP, C.c_int, C.c_float, C.c_float, FP]),
    "orc_sh_get_peak_f": (C.c_float, [FP, C.c_int, FP, I]),
    "orc_sh2peaks_voxel": (C.c_int, [FP, C.c_int, C.c_int, FP, C.c_int, C.c_int, C.c_float, FP, FP, I]),
    "orc_sh2peaks_run": (C.c_longlong, [C.c_void_p, C.c_int, C.c_int, C.c_size_t, C.c_int, FP, C.c_int, C.c_int,
                                        C.c_float, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    "orc_sh2peaks_run_fast": (C.c_longlong, [C.c_void_p, C.c_int, C.c_int, C.c_size_t, C.c_int, FP, C.c_int, C.c_int,
                                             C.c_float, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                             C.c_void_p]),
    "orc_peaks_fast_table": (C.c_int, [C.c_int, C.c_void_p]),
}


_FILES = {"d": "liborc.so", "ld": "liborc_ld.so", "o3": "liborc_o3.so"}


def lib(long_double=False, generic=False):
    """The FP64 build (-O3 -march=native), its generic-x86-64 -O3 build (generic=True), or the
    long double build."""
    key = "ld" if long_double else ("o3" if generic else "d")
    if key in _libs:
        return _libs[key]
    path = os.path.join(BUILD, _FILES[key])
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    for name, (res, args) in _SIG.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _libs[key] = L
    return L


def _d(a):
    return a.ctypes.data_as(D)


def _arr(x, dtype=np.float64):
    return np.ascontiguousarray(x, dtype=dtype)


def sh_transform(dirs, lmax, long_double=False):
    dirs = _arr(dirs)
    n = lib().orc_nforl(lmax)
    out = np.empty((dirs.shape[0], n))
    lib(long_double).orc_sh_transform(_d(dirs), dirs.shape[0], lmax, _d(out))
    return out


def plm(lmax, m, x):
    out = np.zeros(lmax + 1)
    lib().orc_plm_sph(_d(out), lmax, m, float(x))
    return out


def zsh2rh(z):
    z = _arr(z)
    out = np.empty_like(z)
    lib().orc_zsh2rh(_d(z), z.size, _d(out))
    return out


def pinv(M):
    M = _arr(M)
    out = np.empty((M.shape[1], M.shape[0]))
    lib().orc_pinv(_d(M), M.shape[0], M.shape[1], _d(out))
    return out


def lforn(n):
    return lib().orc_lforn(n)


def normalise_grad(grad, scaling_mode=0):
    g = _arr(grad).copy()
    scaled = lib().orc_dw_scheme_normalise(_d(g), g.shape[0], scaling_mode)
    return g, bool(scaled)


def direction_matrix(grad, vols):
    grad = _arr(grad)
    v = _arr(vols, np.int32)
    out = np.empty((v.size, 2))
    lib().orc_gen_direction_matrix(_d(grad), v.ctypes.data_as(I), v.size, _d(out))
    return out


def shells_cluster(b, bzero=10.0, eps=80.0):
    b = _arr(b)
    cl = np.zeros(b.size, dtype=np.int32)
    n = lib().orc_shells_cluster(_d(b), b.size, bzero, eps, cl.ctypes.data_as(I))
    return n, cl


def csd_select_volumes(grad, bzero=10.0, eps=80.0):
    g = _arr(grad)
    vols = np.zeros(g.shape[0], dtype=np.int32)
    nv = C.c_int(0)
    rc = lib().orc_csd_select_volumes(_d(g), g.shape[0], bzero, eps, vols.ctypes.data_as(I), C.byref(nv))
    if rc:
        raise ValueError("shell selection failed (%d)" % rc)
    return vols[:nv.value].copy()


def load_dirs300():
    path = os.path.join(os.path.dirname(HERE), "mrtrix3_amd", "data", "electrostatic_repulsion_300.txt")
    return np.loadtxt(path)


class CsdOracle:
    """CSD::Shared::init + per-voxel CSD solves, FP64 (or long double) on the CPU."""

    def __init__(self, grad, response, lmax=0, filter=None, directions=None, neg_lambda=1.0,
                 norm_lambda=1.0, threshold=0.0, niter=50, long_double=False, bzero=10.0, eps=80.0,
                 generic=False):
        self.L = lib(long_double, generic)
        g, _ = normalise_grad(grad)
        self.grad = g
        self.vols = csd_select_volumes(g, bzero, eps)
        dw_dirs = direction_matrix(g, self.vols)
        hr = _arr(directions) if directions is not None else load_dirs300()
        resp = _arr(np.ravel(response))
        filt = _arr(filter) if filter is not None else None
        err = C.create_string_buffer(256)
        ptr = C.POINTER(CsdShared)()
        rc = self.L.orc_csd_init(_d(dw_dirs), dw_dirs.shape[0], _d(hr), hr.shape[0], _d(resp), resp.size,
                                 int(lmax), _d(filt) if filt is not None else None,
                                 filt.size if filt is not None else 0, neg_lambda, norm_lambda, threshold,
                                 int(niter), C.byref(ptr), err, 256)
        if rc:
            raise ValueError(err.value.decode())
        self.p = ptr
        s = ptr.contents
        self.n_dw, self.n_sh, self.n_init, self.n_hr = s.n_dw, s.n_sh, s.n_init, s.n_hr
        self.lmax, self.lmax_response, self.lmax_data = s.lmax, s.lmax_response, s.lmax_data
        self.niter, self.threshold, self.norm_lambda = s.niter, s.threshold, s.norm_lambda
        self.rconv = np.ctypeslib.as_array(s.rconv, (self.n_init, self.n_dw)).copy()
        self.M = np.ctypeslib.as_array(s.M, (self.n_dw, self.n_sh)).copy()
        self.MtM = np.ctypeslib.as_array(s.MtM, (self.n_sh, self.n_sh)).copy()
        self.HR_trans = np.ctypeslib.as_array(s.HR_trans, (self.n_hr, self.n_sh)).copy()

    def solve(self, b, F0=None):
        """One voxel (selected, clamped signal). Returns (F, n, neg_list, trace_k)."""
        b = _arr(b)
        F = np.empty(self.n_sh)
        neg = np.empty(self.n_hr, dtype=np.int32)
        nneg = C.c_int(0)
        trace = np.zeros(max(self.niter, 1), dtype=np.int32)
        if F0 is None:
            n = self.L.orc_csd_solve(self.p, _d(b), _d(F), neg.ctypes.data_as(I), C.byref(nneg),
                                     trace.ctypes.data_as(I))
        else:
            f0 = _arr(F0)
            n = self.L.orc_csd_solve_from(self.p, _d(b), _d(f0), _d(F), neg.ctypes.data_as(I),
                                          C.byref(nneg), trace.ctypes.data_as(I))
        return F, n, neg[:nneg.value].copy(), trace[:min(n, self.niter)].copy()

    def run_volume(self, dw, shape=None, mask=None, nthreads=0):
        """dw [nvox, nvol] float32 (x fastest).  Returns (fod float32 [nvox, n_sh], iters int16)."""
        dw = _arr(dw, np.float32)
        nvox, nvol = dw.shape
        if shape is None:
            shape = (nvox, 1, 1)
        nx, ny, nz = shape
        assert nx * ny * nz == nvox
        fod = np.empty((nvox, self.n_sh), dtype=np.float32)
        iters = np.empty(nvox, dtype=np.int16)
        m = _arr(mask, np.uint8) if mask is not None else None
        self.L.orc_csd_run_volume(self.p, dw.ctypes.data, nvol, self.vols.ctypes.data_as(I),
                                  m.ctypes.data if m is not None else None, nx, ny, nz, int(nthreads),
                                  fod.ctypes.data, iters.ctypes.data)
        return fod, iters

    def time_syrk_llt(self, k, reps=2000):
        """Single-thread GFLOP/s of the rank-k normal-matrix update (n(n+1)k FLOP) and the LLT
        (n^3/3 FLOP) as CSD::iterate runs them, k constraint rows."""
        ts, tl = C.c_double(0), C.c_double(0)
        if self.L.orc_time_syrk_llt(self.p, int(k), int(reps), C.byref(ts), C.byref(tl)) < 0:
            raise ValueError("bad timing arguments")
        n = self.n_sh
        return {"n": n, "k": int(k), "syrk_gflops": n * (n + 1) * k * reps / ts.value / 1e9,
                "llt_gflops": n ** 3 / 3.0 * reps / tl.value / 1e9}

    def __del__(self):
        p = getattr(self, "p", None)
        if p:
            self.L.orc_csd_free(p)
            self.p = None


class IclsOracle:
    def __init__(self, H, A, t=None, num_eq=0, sol_reg=0.0, con_reg=0.0, max_iter=0, tol=0.0,
                 standard_form=False, long_double=False, generic=False):
        self.L = lib(long_double, generic)
        H = _arr(H)
        A = _arr(A)
        tt = _arr(t) if t is not None and np.size(t) else None
        err = C.create_string_buffer(256)
        ptr = C.POINTER(IclsProblem)()
        n = H.shape[1]
        rc = self.L.orc_icls_problem_create(_d(H), H.shape[0], n, _d(A), A.shape[0],
                                            _d(tt) if tt is not None else None, int(num_eq), sol_reg,
                                            con_reg, int(max_iter), tol, int(standard_form),
                                            C.byref(ptr), err, 256)
        if rc:
            raise ValueError(err.value.decode())
        self.p = ptr
        s = ptr.contents
        self.n, self.ncon, self.nmeas, self.max_niter = s.n, s.ncon, s.nmeas, s.max_niter

    def matrices(self):
        s = self.p.contents
        return dict(chol=np.ctypeslib.as_array(s.chol, (self.n, self.n)).copy(),
                    b2d=np.ctypeslib.as_array(s.b2d, (self.nmeas, self.n)).copy(),
                    B=np.ctypeslib.as_array(s.B, (self.ncon, self.n)).copy(),
                    t=(np.ctypeslib.as_array(s.t, (self.ncon,)).copy() if s.t else None))

    def solve(self, b):
        b = _arr(b)
        x = np.empty(self.n)
        act = np.zeros(self.ncon, dtype=np.uint8)
        it = self.L.orc_icls_solve(self.p, _d(b), _d(x), act.ctypes.data)
        return x, it, act.astype(bool)

    def run_batch(self, b, valid=None, nthreads=0):
        b = _arr(b, np.float32)
        nvox = b.shape[0]
        x = np.empty((nvox, self.n), dtype=np.float32)
        it = np.empty(nvox, dtype=np.int16)
        v = _arr(valid, np.uint8) if valid is not None else None
        self.L.orc_icls_run_batch(self.p, b.ctypes.data, nvox, v.ctypes.data if v is not None else None,
                                  int(nthreads), x.ctypes.data, it.ctypes.data)
        return x, it

    def __del__(self):
        p = getattr(self, "p", None)
        if p:
            self.L.orc_icls_problem_free(p)
            self.p = None


# ------------------------------------------------------------------------------------------
# sh2peaks (cmd/sh2peaks.cpp, Math::SH::get_peak): single precision, as the reference computes it
# ------------------------------------------------------------------------------------------
def _f(a):
    return a.ctypes.data_as(FP)


def load_peak_seeds():
    """The 60 built-in search directions of sh2peaks [azimuth, elevation] as float32."""
    path = os.path.join(os.path.dirname(HERE), "mrtrix3_amd", "data", "sh2peaks_seeds_60.txt")
    return np.loadtxt(path, dtype=np.float64).astype(np.float32).reshape(-1, 2)


def plm_sph_f(lmax, m, x, sin_x):
    out = np.zeros(lmax + 1, dtype=np.float32)
    lib().orc_plm_sph_f(_f(out), lmax, m, np.float32(x), np.float32(sin_x))
    return out


def sh_derivatives_f(sh, lmax, el, az):
    sh = _arr(sh, np.float32)
    out = np.zeros(6, dtype=np.float32)
    lib().orc_sh_derivatives_f(_f(sh), lmax, np.float32(el), np.float32(az), _f(out))
    return out


def sh_get_peak_f(sh, lmax, direction):
    """Returns (amplitude, unit direction, iterations); NaN amplitude when it does not converge."""
    sh = _arr(sh, np.float32)
    d = _arr(direction, np.float32).copy()
    it = C.c_int(0)
    a = lib().orc_sh_get_peak_f(_f(sh), lmax, _f(d), C.byref(it))
    return float(a), d, it.value


def peaks_fast_table(lmax, generic=False):
    """The -fast lookup table [512, (lmax/2+1)^2] float32 (SH.h:418-434)."""
    nal = (lmax // 2 + 1) ** 2
    tab = np.empty((512, nal), dtype=np.float32)
    assert lib(False, generic).orc_peaks_fast_table(int(lmax), tab.ctypes.data) == nal
    return tab


def sh2peaks(sh, lmax=None, seeds=None, npeaks=3, threshold=-np.inf, directions=None, peaks_image=None,
             mask=None, nthreads=0, generic=False, fast=False):
    """sh [nvox, nsh] float32 -> peaks [nvox, 3 npeaks] (amplitude x direction, NaN past the last).
    fast: Legendre values from the 512-elevation lookup table (-fast).
    directions: [k, 3] directions reported for every voxel (-direction; npeaks = k);
    peaks_image: [nvox, 3 npeaks] per-voxel directions (-peaks).  Returns (peaks, Newton iterations)."""
    sh = _arr(np.atleast_2d(sh), np.float32)
    nvox, nsh = sh.shape
    if lmax is None:
        lmax = lforn(nsh)
    seeds = load_peak_seeds() if seeds is None else _arr(seeds, np.float32)
    ld_match = 0
    m = None
    if directions is not None:
        m = _arr(directions, np.float32).reshape(-1)
        npeaks = m.size // 3
    elif peaks_image is not None:
        m = _arr(peaks_image, np.float32)
        assert m.shape[0] == nvox and m.shape[1] % 3 == 0
        npeaks = m.shape[1] // 3
        ld_match = m.shape[1]
    out = np.empty((nvox, 3 * npeaks), dtype=np.float32)
    mk = _arr(mask, np.uint8) if mask is not None else None
    tab = peaks_fast_table(int(lmax), generic) if fast else None
    it = lib(False, generic).orc_sh2peaks_run_fast(sh.ctypes.data, nsh, nsh, nvox, int(lmax), _f(seeds), seeds.shape[0],
                                int(npeaks), float(threshold), m.ctypes.data if m is not None else None, ld_match,
                                mk.ctypes.data if mk is not None else None, int(nthreads),
                                tab.ctypes.data if tab is not None else None, out.ctypes.data)
    return out, int(it)


# ------------------------------------------------------------------------------------------
# MSMT_CSD::Shared (src/dwi/sdeconv/msmt_csd.h:100-218): C and A built by the oracle alone
# ------------------------------------------------------------------------------------------
def msmt_shells(grad, shells=None, bzero=10.0, eps=80.0):
    """Shell index of each volume (-1: in no retained shell), shells sorted by mean b-value
    (shells.cpp:300-342).  shells: the -shells b-values; each picks the b=0 shell (b <= bzero),
    else the shell whose [min, max] holds it, else the one within 1 of its mean
    (shells.cpp:115-180); None keeps every shell (msmt_csd.h:54)."""
    g = _arr(grad)
    b = g[:, 3]
    _, cl = shells_cluster(b, bzero, eps)
    ids = [c for c in np.unique(cl) if c > 0]
    ids.sort(key=lambda c: b[cl == c].mean())
    if shells is not None:
        keep = []
        for want in shells:
            if want < 0:
                raise ValueError("Cannot select shells corresponding to negative b-values")
            hit = None
            for c in ids:
                bc = b[cl == c]
                is_zero = bc.mean() <= bzero
                if (want <= bzero and is_zero) or (want > bzero and bc.min() <= want <= bc.max()):
                    hit = c
                    break
            if hit is None:
                near = [c for c in ids if abs(b[cl == c].mean() - want) <= 1.0]
                if len(near) != 1:
                    raise ValueError("no unambiguous shell for b = %g" % want)
                hit = near[0]
            if hit in keep:
                raise ValueError("shell selected more than once")
            keep.append(hit)
        ids = [c for c in ids if c in keep]
    shell_of = np.array([ids.index(c) if c in ids else -1 for c in cl], dtype=np.int32)
    return shell_of, len(ids)


def msmt_matrices(grad, responses, lmax=None, shells=None, directions=None):
    """(C [nvol, nparams], A [ncon, nparams], lmax list) of MSMT_CSD::Shared::init from the raw
    gradient table and per-tissue response matrices [nshell, ncoef]."""
    g, _ = normalise_grad(grad)
    shell_of, nshell = msmt_shells(g, shells)
    resp = [np.atleast_2d(_arr(r)) for r in responses]
    if lmax is None:
        lmax = []
        for r in resp:   # trailing all-zero columns dropped (msmt_csd.h:237-255), capped at 8
            nz = [j + 1 for j in range(r.shape[1]) if np.any(r[:, j] != 0)]
            lmax.append(min(8, 2 * (max(nz, default=1) - 1)))
    lmax = [int(l) for l in lmax]
    for r in resp:
        if r.shape[0] != nshell:
            raise ValueError("number of rows in response functions must match number of b-value shells")
    cat = _arr(np.concatenate([r.ravel() for r in resp]))
    cols = np.array([r.shape[1] for r in resp], dtype=np.int32)
    lm = np.array(lmax, dtype=np.int32)
    hr = _arr(directions) if directions is not None else load_dirs300()
    nparams = int(sum((l + 1) * (l + 2) // 2 for l in lmax))
    ncon = int(sum(hr.shape[0] if l > 0 else 1 for l in lmax))
    Cm = np.zeros((g.shape[0], nparams))
    A = np.zeros((ncon, nparams))
    nc, nr = C.c_int(), C.c_int()
    lib().orc_msmt_build(_d(g), g.shape[0], shell_of.ctypes.data_as(I), nshell, _d(cat),
                         cols.ctypes.data_as(I), len(resp), lm.ctypes.data_as(I), _d(hr), hr.shape[0],
                         _d(Cm), _d(A), C.byref(nc), C.byref(nr))
    return Cm, A, lmax


class MsmtOracle(IclsOracle):
    """MSMT_CSD::Shared::init + ICLS solves, every matrix built by the oracle."""

    def __init__(self, grad, responses, lmax=None, shells=None, directions=None, norm_lambda=1e-10,
                 neg_lambda=1e-10, **kw):
        self.C, self.A, self.lmax = msmt_matrices(grad, responses, lmax, shells, directions)
        super().__init__(self.C, self.A, None, 0, norm_lambda, neg_lambda, **kw)


# ------------------------------------------------------------------------------------------
# fod2fixel's FMLS segmentation (src/dwi/fmls.cpp), FP64
# ------------------------------------------------------------------------------------------
FMLS_MAX_ADJ = 8


def load_dirs1281():
    path = os.path.join(os.path.dirname(HERE), "mrtrix3_amd", "data", "tesselation_1281.txt")
    return np.loadtxt(path)


def fmls_unit_vectors(azel):
    """Set::initialise (directions/set.h:128-139): unit vectors of [azimuth, elevation] rows."""
    az, el = azel[:, 0], azel[:, 1]
    s = np.sin(el)
    return np.ascontiguousarray(np.stack([np.cos(az) * s, np.sin(az) * s, np.cos(el)], axis=1))


def fmls_adjacency(unit):
    """Neighbours of each direction: the edges of the convex hull of the directions and their
    antipodes (what Set::initialise_adjacency, set.cpp:70-313, computes), here by scipy's Qhull.
    [n, FMLS_MAX_ADJ] int32, ascending, -1 padded."""
    from scipy.spatial import ConvexHull
    n = unit.shape[0]
    hull = ConvexHull(np.concatenate([unit, -unit]))
    nb = [set() for _ in range(n)]
    for s in hull.simplices:
        for a in s:
            for b in s:
                if a != b and (a % n) != (b % n):
                    nb[a % n].add(b % n)
    out = -np.ones((n, FMLS_MAX_ADJ), dtype=np.int32)
    for i in range(n):
        v = sorted(nb[i])
        out[i, :len(v)] = v
    return out


def fmls_weights(unit):
    """IntegrationWeights (fmls.cpp:109-141): least-squares weights with which the directions
    integrate every even SH basis function up to LforN(n) + 2 (4 pi for the constant)."""
    azel = np.stack([np.arctan2(unit[:, 1], unit[:, 0]), np.arccos(unit[:, 2])], axis=1)
    A = sh_transform(azel, lforn(unit.shape[0]) + 2).T
    b = np.zeros(A.shape[0])
    b[0] = 2.0 * np.sqrt(np.pi)
    q, r = np.linalg.qr(A)
    return np.linalg.solve(r, q.T @ b)


class FmlsOracle:
    """Segmenter (fmls.cpp:144-377) on the 1281-direction set fod2fixel uses (fod2fixel.cpp:318)."""

    def __init__(self, lmax, integral=0.0, peak_value=0.1, merge_ratio=1.0, azel=None):
        L = lib()
        for name, res, args in (
                ("orc_fmls_al_table", C.c_int, [C.c_int, C.c_int, D]),
                ("orc_fmls_get_peak", C.c_double, [D, C.c_int, D, D, C.c_int]),
                ("orc_fmls_lookup_create", C.c_void_p, [D, C.c_int, I]),
                ("orc_fmls_lookup_free", None, [C.c_void_p]),
                ("orc_fmls_lookup_info", C.c_int, [C.c_void_p, I, I, I, I, C.c_int]),
                ("orc_fmls_select", C.c_int, [C.c_void_p, D, D]),
                ("orc_fmls_segment", C.c_int, [D, C.c_int, D, C.c_int, D, I, D, C.c_void_p, D, C.c_int,
                                               C.c_double, C.c_double, C.c_double, D, C.c_int, I])):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        self.L = L
        self.lmax = int(lmax)
        self.azel = load_dirs1281() if azel is None else _arr(azel)
        self.unit = fmls_unit_vectors(self.azel)
        self.n = self.unit.shape[0]
        self.adj = fmls_adjacency(self.unit)
        self.weights = fmls_weights(self.unit)
        az = np.arctan2(self.unit[:, 1], self.unit[:, 0])
        el = np.arccos(self.unit[:, 2])
        self.SHT = np.ascontiguousarray(sh_transform(np.stack([az, el], axis=1), self.lmax))
        self.al_ndir = 2 * self.n
        nal = (self.lmax // 2 + 1) ** 2
        self.al = np.zeros((self.al_ndir, nal))
        L.orc_fmls_al_table(self.lmax, self.al_ndir, _d(self.al))
        self.look = L.orc_fmls_lookup_create(_d(self.unit), self.n, self.adj.ctypes.data_as(I))
        self.integral, self.peak_value, self.merge_ratio = float(integral), float(peak_value), float(merge_ratio)

    def lookup_tables(self):
        naz, nel = C.c_int(), C.c_int()
        total = self.L.orc_fmls_lookup_info(self.look, C.byref(naz), C.byref(nel), None, None, 0)
        start = np.zeros(naz.value * nel.value + 1, dtype=np.int32)
        idx = np.zeros(total, dtype=np.int32)
        self.L.orc_fmls_lookup_info(self.look, C.byref(naz), C.byref(nel), start.ctypes.data_as(I),
                                    idx.ctypes.data_as(I), total)
        return naz.value, nel.value, start, idx

    def select(self, p):
        p = _arr(p)
        return int(self.L.orc_fmls_select(self.look, _d(self.unit), _d(p)))

    def get_peak(self, sh, direction):
        sh = _arr(sh)
        d = _arr(direction).copy()
        v = self.L.orc_fmls_get_peak(_d(sh), self.lmax, _d(d), _d(self.al), self.al_ndir)
        return float(v), d

    def segment(self, sh, max_out=64):
        """One FOD -> (lobes [k, 8]: mean_dir[3], peak_dir[3], integral, max_peak, in the reference's
        order; lut [n]: lobe of each direction, k where none)."""
        sh = _arr(sh)
        out = np.zeros((max_out, 8))
        lut = np.zeros(self.n, dtype=np.int32)
        k = self.L.orc_fmls_segment(_d(sh), self.lmax, _d(self.SHT), self.n, _d(self.unit),
                                    self.adj.ctypes.data_as(I), _d(self.weights), self.look, _d(self.al),
                                    self.al_ndir, self.integral, self.peak_value, self.merge_ratio,
                                    _d(out), max_out, lut.ctypes.data_as(I))
        if k < 0:
            raise ValueError("more than 1281 lobes before thresholding")
        return out[:min(k, max_out)].copy(), lut

    def __del__(self):
        look = getattr(self, "look", None)
        if look:
            self.L.orc_fmls_lookup_free(look)
            self.look = None
