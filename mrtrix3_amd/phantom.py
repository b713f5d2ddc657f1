"""Seeded synthetic diffusion phantoms for the CSD workloads (numpy, float32 output).

Each voxel's fibre orientation distribution is a mixture of 1-3 "sticks" with uniform random
orientations, Dirichlet(1) volume fractions and a minimum pairwise crossing angle; the signal of
a stick is an axially symmetric tensor S0 * exp(-b (l_perp + (l_par - l_perp) cos^2 t)).  The
single-fibre response handed to the solver is that kernel's zonal SH expansion up to `lmax_resp`
(Gauss-Legendre projection), which is what dwi2response would estimate from such data.  Rician
noise with sigma = S0 / SNR is added to every volume.

Workloads:
  single_shell_small   32x32x16, all voxels, 5 b=0 + 60 b=3000, SNR 30
  single_shell_brain   128x128x72 with an ellipsoidal mask (~650k voxels), 8 b=0 + 64 b=3000,
                       15% isotropic-only voxels, SNR ~ U(15, 40) per voxel
  super_resolved       same mask, 5 b=0 + 45 b=3000, crossing angles >= 25 deg, SNR 20
  multi_shell          same mask, b=0 x10, 1000 x30, 2000 x60, 3000 x90; WM sticks + GM + CSF
"""
import math

import numpy as np

S0 = 1000.0
L_PAR = 1.7e-3
L_PERP = 0.3e-3


def fibonacci_sphere(n, rng):
    """n unit vectors on a Fibonacci lattice, randomly rotated."""
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    r = np.sqrt(1.0 - z * z)
    v = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    q, rr = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(rr))
    return v @ q.T


def gradient_table(shells, rng):
    """shells: list of (b, count).  b=0 rows get a zero direction.  Returns [nvol, 4]."""
    rows = []
    for b, cnt in shells:
        if b == 0:
            rows.append(np.zeros((cnt, 4)))
        else:
            d = fibonacci_sphere(cnt, rng)
            rows.append(np.concatenate([d, np.full((cnt, 1), float(b))], axis=1))
    return np.concatenate(rows, axis=0)


def legendre_p(l, x):
    return np.polynomial.legendre.legval(x, [0.0] * l + [1.0])


def stick_kernel(b, cos_t, l_par=L_PAR, l_perp=L_PERP):
    return S0 * np.exp(-b * (l_perp + (l_par - l_perp) * cos_t * cos_t))


def zonal_response(b, lmax, l_par=L_PAR, l_perp=L_PERP, npts=96):
    """Zonal SH coefficients r_l (l = 0, 2, ..., lmax) of the stick kernel at b-value b."""
    x, w = np.polynomial.legendre.leggauss(npts)
    s = stick_kernel(b, x, l_par, l_perp)
    out = []
    for l in range(0, lmax + 1, 2):
        y = math.sqrt((2 * l + 1) / (4 * math.pi)) * legendre_p(l, x)
        out.append(2 * math.pi * np.sum(w * s * y))
    return np.array(out)


def isotropic_response(b, adc, lmax=0):
    """An isotropic compartment: only r_0 = sqrt(4 pi) * S0 exp(-b adc) is non-zero."""
    out = np.zeros(lmax // 2 + 1)
    out[0] = math.sqrt(4 * math.pi) * S0 * math.exp(-b * adc)
    return out


def _random_units(n, rng):
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def random_sticks(nvox, rng, min_angle_deg=30.0, max_sticks=3):
    """Per voxel: count in 1..max_sticks, unit orientations [nvox, 3, 3] (unused rows zero) with
    pairwise axial angles >= min_angle, Dirichlet(1) fractions [nvox, 3]."""
    count = rng.integers(1, max_sticks + 1, size=nvox)
    dirs = np.zeros((nvox, max_sticks, 3))
    cos_max = math.cos(math.radians(min_angle_deg))
    dirs[:, 0] = _random_units(nvox, rng)
    for k in range(1, max_sticks):
        need = np.nonzero(count > k)[0]
        cand = np.zeros((need.size, 3))
        todo = np.arange(need.size)
        while todo.size:
            c = _random_units(todo.size, rng)
            ok = np.ones(todo.size, dtype=bool)
            for j in range(k):
                ok &= np.abs(np.sum(c * dirs[need[todo], j], axis=1)) <= cos_max
            cand[todo[ok]] = c[ok]
            todo = todo[~ok]
        dirs[need, k] = cand
    frac = np.zeros((nvox, max_sticks))
    g = rng.gamma(1.0, size=(nvox, max_sticks))
    mask = np.arange(max_sticks)[None, :] < count[:, None]
    g = np.where(mask, g, 0.0)
    frac = g / g.sum(axis=1, keepdims=True)
    return count, dirs, frac


def stick_signals(grad, dirs, frac, chunk=65536):
    """Noise-free mixture signal [nvox, nvol] (float64) of sticks for a gradient table."""
    nvox = dirs.shape[0]
    g = grad[:, :3]
    b = grad[:, 3]
    out = np.empty((nvox, grad.shape[0]))
    for s in range(0, nvox, chunk):
        e = min(nvox, s + chunk)
        acc = np.zeros((e - s, grad.shape[0]))
        for k in range(dirs.shape[1]):
            cos_t = dirs[s:e, k] @ g.T
            acc += frac[s:e, k:k + 1] * stick_kernel(b[None, :], cos_t)
        out[s:e] = acc
    return out


def add_rician(sig, sigma, rng):
    sigma = np.asarray(sigma, dtype=np.float64)
    if sigma.ndim == 1:
        sigma = sigma[:, None]
    n1 = rng.standard_normal(sig.shape) * sigma
    n2 = rng.standard_normal(sig.shape) * sigma
    return np.sqrt((sig + n1) ** 2 + n2 ** 2)


def ellipsoid_mask(shape, semi):
    c = [(n - 1) / 2.0 for n in shape]
    zz, yy, xx = np.meshgrid(*[np.arange(n) for n in shape[::-1]], indexing="ij")
    r = ((xx - c[0]) / semi[0]) ** 2 + ((yy - c[1]) / semi[1]) ** 2 + ((zz - c[2]) / semi[2]) ** 2
    return (r <= 1.0)   # [z, y, x]


class Phantom:
    """Voxel signals of one workload.

    dw      [nvox, nvol] float32 signals of the masked voxels (x fastest, then y, then z)
    grad    [nvol, 4] gradient table (unit directions, b in s/mm^2)
    response zonal response rows (one row per non-zero shell for single-tissue CSD)
    mask    [nz, ny, nx] bool (None: every voxel)
    shape   (nx, ny, nz)
    """

    def __init__(self, name, dw, grad, response, mask, shape, extra=None):
        self.name = name
        self.dw = dw
        self.grad = grad
        self.response = response
        self.mask = mask
        self.shape = shape
        self.extra = extra or {}

    @property
    def nvox(self):
        return self.dw.shape[0]

    def volume(self):
        """The full 4-D image [nz, ny, nx, nvol] (zeros outside the mask)."""
        nx, ny, nz = self.shape
        vol = np.zeros((nz, ny, nx, self.dw.shape[1]), dtype=np.float32)
        if self.mask is None:
            vol.reshape(-1, self.dw.shape[1])[:] = self.dw
        else:
            vol[self.mask] = self.dw
        return vol


def single_shell(name, shape, n_b0, n_dw, bval, seed, mask_semi=None, snr=30.0,
                 snr_range=None, iso_fraction=0.0, min_angle=30.0, lmax_resp=8, protocol_seed=None):
    """protocol_seed fixes the gradient table independently of the voxel data (default: seed)."""
    grad = gradient_table([(0, n_b0), (bval, n_dw)],
                          np.random.default_rng(seed if protocol_seed is None else protocol_seed))
    rng = np.random.default_rng(seed + 7919)
    mask = ellipsoid_mask(shape, mask_semi) if mask_semi is not None else None
    nvox = int(mask.sum()) if mask is not None else int(np.prod(shape))
    count, dirs, frac = random_sticks(nvox, rng, min_angle_deg=min_angle)
    sig = stick_signals(grad, dirs, frac)
    iso = np.zeros(nvox, dtype=bool)
    if iso_fraction > 0:
        iso = rng.random(nvox) < iso_fraction
        d_iso = rng.uniform(0.8e-3, 3.0e-3, size=int(iso.sum()))
        sig[iso] = S0 * np.exp(-grad[None, :, 3] * d_iso[:, None])
    if snr_range is not None:
        snr_v = rng.uniform(snr_range[0], snr_range[1], size=nvox)
    else:
        snr_v = np.full(nvox, float(snr))
    dw = add_rician(sig, S0 / snr_v, rng).astype(np.float32)
    response = zonal_response(bval, lmax_resp)[None, :]
    extra = dict(count=count, dirs=dirs, frac=frac, iso=iso, snr=snr_v)
    return Phantom(name, dw, grad, response, mask, shape, extra)


def single_shell_small(seed=1, protocol_seed=1):
    return single_shell("single_shell_small 32x32x16 5xb0+60xb3000 SNR30", (32, 32, 16), 5, 60,
                        3000, seed, snr=30.0, protocol_seed=protocol_seed)


BRAIN_SHAPE = (128, 128, 72)
BRAIN_SEMI = (63.5, 63.5, 38.5)


def single_shell_brain(seed=2, protocol_seed=2):
    return single_shell("single_shell_brain 128x128x72 ellipsoid 8xb0+64xb3000 SNR15-40 15%iso",
                        BRAIN_SHAPE, 8, 64, 3000, seed, mask_semi=BRAIN_SEMI,
                        snr_range=(15.0, 40.0), iso_fraction=0.15, protocol_seed=protocol_seed)


def super_resolved(seed=3, protocol_seed=3):
    return single_shell("super_resolved 128x128x72 ellipsoid 5xb0+45xb3000 SNR20 lmax12",
                        BRAIN_SHAPE, 5, 45, 3000, seed, mask_semi=BRAIN_SEMI, snr=20.0,
                        min_angle=25.0, lmax_resp=12, protocol_seed=protocol_seed)


def super_resolved_small(seed=3, protocol_seed=3, shape=(16, 16, 8)):
    """The super-resolved protocol (5xb0 + 45xb3000, lmax-12 response, SNR 20, crossings down to
    25 degrees) on a small unmasked block, for parity tests against the CPU solver."""
    return single_shell("super_resolved_small %dx%dx%d 5xb0+45xb3000 SNR20 lmax12" % shape, shape, 5, 45,
                        3000, seed, snr=20.0, min_angle=25.0, lmax_resp=12, protocol_seed=protocol_seed)


def multi_shell(seed=4, shape=BRAIN_SHAPE, semi=BRAIN_SEMI):
    """Three tissues: WM sticks (per-shell tensor responses, lmax 8), GM (ADC 0.8e-3) and CSF
    (ADC 3.0e-3); fractions ~ Dirichlet(2, 1, 1); SNR 25."""
    rng = np.random.default_rng(seed)
    shells = [(0, 10), (1000, 30), (2000, 60), (3000, 90)]
    grad = gradient_table(shells, rng)
    mask = ellipsoid_mask(shape, semi) if semi is not None else None
    nvox = int(mask.sum()) if mask is not None else int(np.prod(shape))
    count, dirs, frac = random_sticks(nvox, rng)
    wm = stick_signals(grad, dirs, frac)
    tf = rng.dirichlet([2.0, 1.0, 1.0], size=nvox)
    b = grad[:, 3]
    gm = S0 * np.exp(-b * 0.8e-3)
    csf = S0 * np.exp(-b * 3.0e-3)
    sig = tf[:, 0:1] * wm + tf[:, 1:2] * gm[None, :] + tf[:, 2:3] * csf[None, :]
    dw = add_rician(sig, S0 / 25.0, rng).astype(np.float32)
    bvals = [0, 1000, 2000, 3000]
    wm_r = np.stack([zonal_response(bb, 8) if bb > 0 else isotropic_response(0, 0.0, 8)
                     for bb in bvals])
    # at b = 0 the stick kernel is isotropic with amplitude S0
    gm_r = np.array([[isotropic_response(bb, 0.8e-3)[0]] for bb in bvals])
    csf_r = np.array([[isotropic_response(bb, 3.0e-3)[0]] for bb in bvals])
    extra = dict(count=count, dirs=dirs, frac=frac, tissue_frac=tf)
    return Phantom("multi_shell 128x128x72 ellipsoid b0x10+1000x30+2000x60+3000x90 SNR25",
                   dw, grad, [wm_r, gm_r, csf_r], mask, shape, extra)
