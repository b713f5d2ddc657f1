"""Learned warm start of the GPU CSD solver (libmrcsd.so, include/mrwarm.h).

    net = WarmNet(n_in=64)                                   # 64 -> 256 -> 256 -> 256 -> 45 + 300
    train_warm_start(net, solver, dw_train, dwis, b0)         # targets from the plain GPU solver
    fod, iters, nneg, stats = warm_solve(net, solver, dw, dwis, b0)

The network sees each voxel's DW signals divided by its mean b=0 signal and predicts the SH
coefficients (also divided by that mean) and, per constraint direction, whether it belongs to
the final negative set.  The solver starts from that set and accepts a voxel only when the set
is a fixed point of CSD::iterate (certifying the unique optimum for threshold 0); other voxels
are re-solved from the plain initialisation.  All network arithmetic runs in the library's
bf16 MFMA kernels; torch only supplies device buffers and streams.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

_P = C.c_void_p
_registered = False


def _register():
    global _registered
    if _registered:
        return
    _lib.register("mrwarm_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(_P)])
    _lib.register("mrwarm_destroy", None, [_P])
    _lib.register("mrwarm_num_params", C.c_size_t, [_P])
    _lib.register("mrwarm_get_params", C.c_int, [_P, _P, C.c_size_t])
    _lib.register("mrwarm_set_params", C.c_int, [_P, _P, C.c_size_t])
    _lib.register("mrwarm_features", C.c_int, [_P, C.c_int, C.c_size_t, _P, C.c_int, _P, C.c_int, _P, _P, _P])
    _lib.register("mrwarm_forward", C.c_int, [_P, _P, C.c_size_t, _P, _P])
    _lib.register("mrwarm_train_step", C.c_int, [_P, _P, _P, _P, C.c_size_t, C.c_float, C.c_float, C.c_float,
                                                _P, _P])
    _lib.register("mrwarm_gradients", C.c_int, [_P, _P, _P, _P, C.c_size_t, C.c_float, C.c_float, _P, _P])
    _lib.register("mrwarm_propose", C.c_int, [_P, _P, C.c_size_t, C.c_int, C.c_int, _P, _P, _P])
    _lib.register("mrwarm_infer", C.c_int, [_P, _P, C.c_int, C.c_size_t, _P, _P, C.c_int, _P, _P, _P])
    _registered = True


def _stream(t):
    import torch
    return torch.cuda.current_stream(t.device).cuda_stream


class WarmNet:
    def __init__(self, n_in, hidden=256, nsh=45, nlog=300, seed=0):
        _register()
        self.n_in, self.hidden, self.nsh, self.nlog = n_in, hidden, nsh, nlog
        h = C.c_void_p()
        check(_lib.load().mrwarm_create(n_in, hidden, nsh, nlog, seed, C.byref(h)))
        self._h = h

    @property
    def nout(self):
        return self.nsh + self.nlog

    def num_params(self):
        return int(_lib.load().mrwarm_num_params(self._h))

    def get_params(self):
        out = np.empty(self.num_params(), dtype=np.float32)
        check(_lib.load().mrwarm_get_params(self._h, out.ctypes.data, out.size))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(_lib.load().mrwarm_set_params(self._h, p.ctypes.data, p.size))

    def save(self, path):
        """Writes the layer sizes and fp32 master parameters as an .npz (no pickled objects): the
        network is trained once per acquisition protocol and reused for every subject."""
        np.savez(path, dims=np.array([self.n_in, self.hidden, self.nsh, self.nlog], dtype=np.int64),
                 params=self.get_params())

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            n_in, hidden, nsh, nlog = (int(v) for v in z["dims"])
            net = cls(n_in, hidden=hidden, nsh=nsh, nlog=nlog)
            net.set_params(z["params"])
        return net

    def features(self, dw, dwis, b0):
        """dw: CUDA float32 [M, nvol]; dwis / b0: CUDA int32 index tensors.  Returns (x bf16, s0)."""
        import torch
        M, ld = dw.shape
        x = torch.empty((M, self.n_in), dtype=torch.bfloat16, device=dw.device)
        s0 = torch.empty(M, dtype=torch.float32, device=dw.device)
        check(_lib.load().mrwarm_features(dw.data_ptr(), ld, M, dwis.data_ptr(), self.n_in, b0.data_ptr(),
                                          b0.numel(), x.data_ptr(), s0.data_ptr(), _stream(dw)))
        return x, s0

    def forward(self, x, out=None):
        import torch
        M = x.shape[0]
        if out is None:
            out = torch.empty((M, self.nout), dtype=torch.float32, device=x.device)
        check(_lib.load().mrwarm_forward(self._h, x.data_ptr(), M, out.data_ptr(), _stream(x)))
        return out

    def train_step(self, x, tgt_coef, tgt_neg, lr, wc=1.0, wl=0.1, want_loss=False):
        M = x.shape[0]
        loss = (C.c_float * 2)() if want_loss else None
        check(_lib.load().mrwarm_train_step(self._h, x.data_ptr(), tgt_coef.data_ptr(), tgt_neg.data_ptr(), M,
                                            lr, wc, wl, C.cast(loss, C.c_void_p) if want_loss else None,
                                            _stream(x)))
        return (loss[0], loss[1]) if want_loss else None

    def gradients(self, x, tgt_coef, tgt_neg, wc=1.0, wl=0.1):
        g = np.empty(self.num_params(), dtype=np.float32)
        loss = (C.c_float * 2)()
        check(_lib.load().mrwarm_gradients(self._h, x.data_ptr(), tgt_coef.data_ptr(), tgt_neg.data_ptr(),
                                           x.shape[0], wc, wl, g.ctypes.data, C.cast(loss, C.c_void_p)))
        return g, (loss[0], loss[1])

    def propose(self, out, s0):
        import torch
        M = out.shape[0]
        F0 = torch.empty((M, self.nsh), dtype=torch.float32, device=out.device)
        neg = torch.empty((M, (self.nlog + 63) // 64), dtype=torch.int64, device=out.device)
        check(_lib.load().mrwarm_propose(out.data_ptr(), s0.data_ptr(), M, self.nsh, self.nlog, F0.data_ptr(),
                                         neg.data_ptr(), _stream(out)))
        return F0, neg

    def fused_ok(self):
        """Dimensions the fused inference kernel handles (hidden 256, <= 128 inputs, <= 512 outputs)."""
        return self.hidden == 256 and self.n_in <= 128 and self.nout <= 512

    def infer(self, dw, dwis, b0, want_F0=False):
        """Features + forward + proposal in one kernel (mrwarm_infer): (F0 or None, neg words)."""
        import torch
        M = dw.shape[0]
        F0 = torch.empty((M, self.nsh), dtype=torch.float32, device=dw.device) if want_F0 else None
        neg = torch.empty((M, (self.nlog + 63) // 64), dtype=torch.int64, device=dw.device)
        check(_lib.load().mrwarm_infer(self._h, dw.data_ptr(), dw.stride(0), M, dwis.data_ptr(), b0.data_ptr(),
                                       b0.shape[0], F0.data_ptr() if want_F0 else None, neg.data_ptr(),
                                       _stream(dw)))
        return F0, neg

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib._lib is not None:
            _lib._lib.mrwarm_destroy(h)
        self._h = None


def unpack_neg(mask, nlog):
    """[M, words] int64 bitmask -> [M, nlog] uint8 labels (torch, on the mask's device)."""
    import torch
    M, words = mask.shape
    bits = torch.arange(64, device=mask.device, dtype=torch.int64)
    lab = ((mask.unsqueeze(-1) >> bits) & 1).reshape(M, words * 64)[:, :nlog]
    return lab.to(torch.uint8).contiguous()


def solve_targets(solver, dw, s0):
    """Plain GPU CSD on the training voxels: normalised coefficients and negative-set labels."""
    fod, it, nn, nm = solver.run(dw, negmask=True)
    coef = (fod / s0.unsqueeze(1)).contiguous()
    return coef, unpack_neg(nm, solver.n_hr)


def train_warm_start(net, solver, dw_train, dwis, b0, epochs=12, batch=4096, lr=2e-3, wc=1.0, wl=0.1,
                     seed=0, log=None):
    """Fits the network to plain-solver answers on dw_train (CUDA float32 [M, nvol])."""
    import torch
    x, s0 = net.features(dw_train, dwis, b0)
    coef, lab = solve_targets(solver, dw_train, s0)
    M = x.shape[0]
    gen = torch.Generator(device=x.device)
    gen.manual_seed(seed)
    steps = 0
    total = epochs * (M // batch)
    history = []
    for ep in range(epochs):
        perm = torch.randperm(M, device=x.device, generator=gen)
        for b in range(M // batch):
            idx = perm[b * batch:(b + 1) * batch]
            # cosine learning-rate decay
            lr_t = 0.5 * lr * (1.0 + np.cos(np.pi * steps / max(1, total)))
            want = (b == M // batch - 1)
            loss = net.train_step(x[idx].contiguous(), coef[idx].contiguous(), lab[idx].contiguous(),
                                  float(lr_t), wc, wl, want_loss=want)
            steps += 1
            if want:
                history.append(loss)
                if log:
                    log("epoch %d: coefficient loss %.4g, set loss %.4g" % (ep, loss[0], loss[1]))
    return history


def warm_solve(net, solver, dw, dwis, b0, valid=None, budget=6, stats=None, fod=None, negmask=False, fused=True):
    """Network proposal -> warm-started CSD (fallback to the plain start).  The solver starts from the
    proposed negative set, so the proposed coefficients are not needed; the fused kernel computes the
    set directly from the DW rows.  When the solver's options void the certificate (threshold not 0,
    or a singular normal matrix), the proposal is not made and every voxel takes the plain start."""
    if not solver.warm_supported():
        return solver.run(dw, valid=valid, stats=stats, fod=fod, negmask=negmask)
    if fused and net.fused_ok():
        _, neg = net.infer(dw, dwis, b0)
        F0 = None
    else:
        x, s0 = net.features(dw, dwis, b0)
        out = net.forward(x)
        F0, neg = net.propose(out, s0)
    return solver.run(dw, valid=valid, warm_F0=F0, warm_neg=neg, warm_budget=budget, stats=stats, fod=fod,
                      negmask=negmask)
