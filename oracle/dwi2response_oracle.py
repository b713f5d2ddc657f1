"""Test infrastructure only (imported by tests/ and bench.py's CPU leg, never by mrtrix3_amd/): the
dwi2response tournier and tax loops (lib/mrtrix3/dwi2response/tournier.py:77-154, tax.py:56-149)
restated on the CPU, the checker of the dwi2response command.

Each step runs, where the script runs a command, the oracle's restatement of it:
  dwi2fod csd -mask                 orc.CsdOracle.run_volume (oracle/csd_oracle.c)
  fod2fixel -fmls_no_thresholds,    orc.FmlsOracle.segment per voxel (oracle/fmls_oracle.c): the
  fixel2voxel none -number 2,       first two lobes' peak amplitudes (0 when absent) and the first
  fixel2peaks -number 1             lobe's mean direction, in float32 as the fixel files hold them
  mrcalc ... (the cost)             cost(): single precision, as mrcalc's real_type float
                                    (cmd/mrcalc.cpp:103-123, 193)
  mrstats count                     finite values only (core/stats.h:55-70)
  mrthreshold -top N                threshold_top(): the N-th largest non-NaN value and every
                                    non-NaN value at or above it (cmd/mrthreshold.cpp:155-245,
                                    292-299)
  amp2response                      resp_oracle.accumulate / solve
  maskfilter dilate -npass          dilate(): face neighbours, npass passes (filter/dilate.h:60-104)
  mrstats -allvolumes -mask (tax)   tax_init_rf(): mean and sample standard deviation of the finite
                                    values, as mrstats prints them: 6 significant digits
                                    (str<cdouble>, core/mrtrix.h:353-363)
Images are flat arrays over the voxel grid, x fastest.
"""
import math

import numpy as np

from . import orc, resp_oracle


def cost(p1, p2):
    """sqrt(p1) * (1 - p2 / p1)^2 in float32 (NaN where p1 = 0)."""
    p1 = np.asarray(p1, dtype=np.float32)
    p2 = np.asarray(p2, dtype=np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.float32(1.0) - p2 / p1
        return (np.sqrt(p1) * np.power(r, np.float32(2.0))).astype(np.float32)


def threshold_top(cf, n):
    data = cf[~np.isnan(cf)]
    if data.size == 0:
        raise ValueError("No valid input data found; unable to determine threshold")
    index = data.size - int(n)
    if index < 0 or index >= data.size:
        raise ValueError("Number of valid input image values (%d) less than number of voxels requested via -top "
                         "option (%d)" % (data.size, n))
    thr = np.partition(data, index)[index]
    with np.errstate(invalid="ignore"):
        return (~np.isnan(cf)) & (cf >= thr)


def dilate(m, shape, npass):
    """m flat over (nx, ny, nz) with x fastest."""
    nx, ny, nz = shape
    a = np.asarray(m, dtype=bool).reshape(nz, ny, nx)
    for _ in range(int(npass)):
        o = a.copy()
        o[:, :, 1:] |= a[:, :, :-1]
        o[:, :, :-1] |= a[:, :, 1:]
        o[:, 1:, :] |= a[:, :-1, :]
        o[:, :-1, :] |= a[:, 1:, :]
        o[1:, :, :] |= a[:-1, :, :]
        o[:-1, :, :] |= a[1:, :, :]
        a = o
    return a.reshape(-1)


def fixel_summary(fod, mask_in, lmax):
    """Per voxel: peak amplitudes of the first two lobes (0 when absent) and the first lobe's mean
    direction (NaN when absent), float32."""
    fm = orc.FmlsOracle(lmax, integral=0.0, peak_value=0.0)
    n = fod.shape[0]
    p = np.zeros((n, 2), dtype=np.float32)
    d = np.full((n, 3), np.nan, dtype=np.float32)
    for v in np.nonzero(mask_in)[0]:
        lobes, _ = fm.segment(fod[v].astype(np.float64), max_out=2048)
        if len(lobes):
            p[v, 0] = lobes[0, 7]
            d[v] = lobes[0, 0:3]
        if len(lobes) > 1:
            p[v, 1] = lobes[1, 7]
    return p, d


def shell_directions(grad):
    """Unit direction of each volume (gen_direction_matrix then spherical2cartesian)."""
    g = np.asarray(grad, dtype=np.float64)
    nrm = np.linalg.norm(g[:, :3], axis=1)
    az = np.arctan2(g[:, 1], g[:, 0])
    el = np.arccos(g[:, 2] / nrm)
    return np.stack([np.sin(el) * np.cos(az), np.sin(el) * np.sin(az), np.cos(el)], axis=1)


def step(dw, grad, shape, mask_in, rf_in, amp_lmax, number, nthreads=8):
    """One iteration on the single-shell signal dw [nvox, nvol] (float32) with gradient table grad
    [nvol, 4]: returns the new response, the cost, the count of finite costs, the selection and
    the FODs."""
    csd = orc.CsdOracle(grad, np.asarray(rf_in, dtype=np.float64))
    fod, _ = csd.run_volume(np.ascontiguousarray(dw, dtype=np.float32), shape,
                            mask=np.asarray(mask_in, dtype=np.uint8), nthreads=nthreads)
    p, d = fixel_summary(fod, mask_in, csd.lmax)
    cf = cost(p[:, 0], p[:, 1])
    count = int(np.isfinite(cf).sum())
    sf = threshold_top(cf, min(int(number), count))
    lm = 10 if amp_lmax is None else int(amp_lmax)
    M, b, _ = resp_oracle.accumulate(dw, d, sf, shell_directions(grad), np.arange(dw.shape[1]), lm)
    rf, _ = resp_oracle.solve(lm, M, b, True)
    return dict(rf=rf, cf=cf, count=count, sf=sf, fod=fod)


def next_mask(cf, count, iter_voxels, npass, mask, shape):
    return dilate(threshold_top(cf, min(int(iter_voxels), count)), shape, npass) & np.asarray(mask, dtype=bool)


def run(dw, grad, shape, mask, number=300, iter_voxels=0, npass=1, max_iters=10, lmax=None, nthreads=8,
        step_fn=None):
    """The loop; returns (response, selection, iterations run, converged, history of steps).
    step_fn (default step) takes the same arguments as step()."""
    if max_iters < 2:
        raise ValueError("Number of iterations must be at least 2")
    if iter_voxels == 0:
        iter_voxels = 10 * number
    elif iter_voxels < number:
        raise ValueError("Number of selected voxels (-iter_voxels) must be greater than number of voxels desired "
                         "(-number)")
    step_fn = step_fn or step
    mask = np.asarray(mask, dtype=bool)
    mask_in = mask
    rf = [1.0, -1.0, 1.0]
    prev = None
    history = []
    for it in range(max_iters):
        s = step_fn(dw, grad, shape, mask_in, rf, 4 if it == 0 else lmax, number, nthreads=nthreads)
        s["mask_in"] = mask_in
        history.append(s)
        rf = s["rf"]
        if it > 0 and not np.any(s["sf"] & ~prev):
            return rf, s["sf"], it + 1, True, history
        mask_in = next_mask(s["cf"], s["count"], iter_voxels, npass, mask, shape)
        s["next_mask"] = mask_in
        prev = s["sf"]
    return rf, history[-1]["sf"], max_iters, False, history


def printed6(x):
    """A double as an ostream prints it by default (6 significant digits), read back."""
    return float("%g" % x)


def tax_init_rf(dw, mask):
    """tax.py:66-73: [m, -s/2, s^2/(4m)] from the mean m and standard deviation s of the DW signal in
    the mask, each scaled by sqrt(4 pi)."""
    vals = np.asarray(dw, dtype=np.float64)[np.asarray(mask, dtype=bool)].ravel()
    vals = vals[np.isfinite(vals)]
    s4 = math.sqrt(4.0 * math.pi)
    m = printed6(vals.mean()) * s4
    sd = printed6(vals.std(ddof=1)) * s4
    return np.array([m, -0.5 * sd, 0.25 * sd * sd / m])


def tax_step(dw, grad, shape, mask_in, rf_in, amp_lmax, peak_ratio, nthreads=8):
    """One tax iteration: the response, the peak ratio image (float32), the selection, the FODs."""
    csd = orc.CsdOracle(grad, np.asarray(rf_in, dtype=np.float64))
    fod, _ = csd.run_volume(np.ascontiguousarray(dw, dtype=np.float32), shape,
                            mask=np.asarray(mask_in, dtype=np.uint8), nthreads=nthreads)
    p, d = fixel_summary(fod, mask_in, csd.lmax)
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = (p[:, 1] / p[:, 0]).astype(np.float32)
        sf = (ratio < np.float32(peak_ratio)) & np.asarray(mask_in, dtype=bool)
    if not sf.any():
        raise ValueError("Aborting: All voxels have been excluded from single-fibre selection")
    lm = 10 if amp_lmax is None else int(amp_lmax)
    M, b, _ = resp_oracle.accumulate(dw, d, sf, shell_directions(grad), np.arange(dw.shape[1]), lm)
    rf, _ = resp_oracle.solve(lm, M, b, True)
    return dict(rf=rf, ratio=ratio, sf=sf, fod=fod)


def tax_converged(old, new, convergence):
    """tax.py:124-133: every coefficient pair's half-difference over its mean within convergence
    percent (a negative mean never asks for another iteration)."""
    change = 0.01 * convergence
    for o, n in zip(old, new):
        if math.fabs(0.5 * (o - n)) / (0.5 * (o + n)) > change:
            return False
    return True


def tax_run(dw, grad, shape, mask, peak_ratio=0.1, max_iters=20, convergence=0.5, lmax=None, nthreads=8,
            step_fn=None, init_rf=None):
    """The tax loop; returns (response, selection, iterations run, converged, history)."""
    step_fn = step_fn or tax_step
    mask_in = np.asarray(mask, dtype=bool)
    rf = tax_init_rf(dw, mask_in) if init_rf is None else np.asarray(init_rf, dtype=np.float64)
    history = []
    for it in range(max_iters):
        s = step_fn(dw, grad, shape, mask_in, rf, lmax, peak_ratio, nthreads=nthreads)
        s["mask_in"] = mask_in
        history.append(s)
        if it > 0 and tax_converged(rf, s["rf"], convergence):
            return s["rf"], s["sf"], it + 1, True, history
        rf = s["rf"]
        mask_in = s["sf"]
    return rf, history[-1]["sf"], max_iters, False, history
