"""float64 numpy reference of the warm-start network — TEST INFRASTRUCTURE ONLY (the CPU oracle
for the bf16 MFMA forward / backward kernels of mrtrix3_amd/csrc/kernels/mlp.hip).

Layers n_in -> h -> h -> h -> nsh + nlog with tanh-GELU, loss
    wc * mean_m sum_c (out_c - t_c)^2  +  wl * mean_m sum_r BCE(logit_r, y_r).
Parameters use the library's order: per layer W [fan_in][fan_out] then b [fan_out].
"""
import numpy as np


def gelu(x):
    c = np.sqrt(2.0 / np.pi)
    return 0.5 * x * (1.0 + np.tanh(c * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    c = np.sqrt(2.0 / np.pi)
    u = c * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * c * (1.0 + 3 * 0.044715 * x * x)


def unpack(params, dims):
    layers, o = [], 0
    for fin, fout in zip(dims[:-1], dims[1:]):
        W = params[o:o + fin * fout].reshape(fin, fout)
        o += fin * fout
        b = params[o:o + fout]
        o += fout
        layers.append((W.astype(np.float64), b.astype(np.float64)))
    return layers


def bf16(a):
    """Round to bfloat16 (round-to-nearest-even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(params, dims, x, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, dims)
    h = r(x)
    pres, acts = [], [h]
    for i, (W, b) in enumerate(layers):
        z = h @ r(W) + b
        if i < len(layers) - 1:
            pres.append(z)
            h = r(gelu(z))
            acts.append(h)
        else:
            out = z
    return out, pres, acts


def loss_and_grads(params, dims, x, tcoef, tneg, wc, wl, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, dims)
    out, pres, acts = forward(params, dims, x, emulate_bf16)
    M = x.shape[0]
    nsh = tcoef.shape[1]
    e = out[:, :nsh] - tcoef
    lg = out[:, nsh:]
    y = tneg.astype(np.float64)
    p = 1.0 / (1.0 + np.exp(-lg))
    loss_c = wc * np.sum(e * e) / M
    loss_l = wl * np.sum(np.maximum(lg, 0) - lg * y + np.log1p(np.exp(-np.abs(lg)))) / M
    dout = np.concatenate([2 * wc * e, wl * (p - y)], axis=1) / M
    grads = []
    dy = dout
    for i in range(len(layers) - 1, -1, -1):
        W, b = layers[i]
        X = acts[i]
        dyr = r(dy)
        grads.append((X.T @ dyr, dy.sum(axis=0)))
        if i == 0:
            break
        dx = dyr @ r(W).T
        dy = dx * gelu_grad(pres[i - 1])
    grads = grads[::-1]
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return (loss_c, loss_l), flat
