"""float64 numpy restatement of global tractography's external energy -- TEST INFRASTRUCTURE ONLY
(the CPU oracle for mrtrix3_amd/csrc/kernels/gt.hip and host/gt.cpp).

    K, Ak = kernels(grad, resp_wm, resp_iso, lmax, weight)     # externalenergy.cpp:55-101
    e, f = energy(K, Ak, dwi_row, tod_row, mu)                  # calcEnergy, :245-252

The NNLS fit is the oracle's ICLS restatement (orc.IclsOracle, pinned by the reference's Matlab
goldens) on Problem(Ak, I), as the reference sets it up (:111)."""
import numpy as np

from . import orc


def sh_delta(u, lmax):
    """Math::SH::delta (core/math/SH.h:279-303): the SH basis at a unit direction, cos / sin of m az
    by the angle-addition recurrence from (u0, u1) / rxy (1, 0 when rxy = 0)."""
    n = (lmax + 1) * (lmax + 2) // 2
    out = np.zeros(n)
    rxy = np.sqrt(u[1] ** 2 + u[0] ** 2)
    cp = u[0] / rxy if rxy else 1.0
    sp = u[1] / rxy if rxy else 0.0
    AL = orc.plm(lmax, 0, u[2])
    for l in range(0, lmax + 1, 2):
        out[l * (l + 1) // 2] = AL[l]
    c0, s0 = 1.0, 0.0
    for m in range(1, lmax + 1):
        AL = orc.plm(lmax, m, u[2])
        c = c0 * cp - s0 * sp
        s = s0 * cp + c0 * sp
        for l in range(m + 1 if m & 1 else m, lmax + 1, 2):
            out[l * (l + 1) // 2 + m] = AL[l] * np.sqrt(2.0) * c
            out[l * (l + 1) // 2 - m] = AL[l] * np.sqrt(2.0) * s
        c0, s0 = c, s
    return out


def kernels(grad, resp_wm, resp_iso, lmax, weight):
    """K [nvol, NforL(lmax)] and Ak [nvol, nf + 1] (externalenergy.cpp:55-101) from the raw
    gradient table (normalised as get_DW_scheme does) and the responses."""
    g, _ = orc.normalise_grad(grad)
    shell_of, nshell = orc.msmt_shells(g)
    resp_wm = np.atleast_2d(np.asarray(resp_wm, dtype=np.float64))
    if resp_wm.shape[0] != nshell:
        raise ValueError("WM kernel size does not match the no. b-values in the image.")
    for r in resp_iso:
        if len(r) != nshell:
            raise ValueError("Isotropic kernel size does not match the no. b-values in the image.")
    nrows, ncols, nf = g.shape[0], (lmax + 1) * (lmax + 2) // 2, len(resp_iso)
    nz = lmax // 2 + 1
    K = np.zeros((nrows, ncols))
    Ak = np.zeros((nrows, nf + 1))
    for s in range(nshell):
        zsh = np.array([resp_wm[s, i] if i < resp_wm.shape[1] else 0.0 for i in range(nz)])
        rh = orc.zsh2rh(zsh)
        wmr0 = resp_wm[s, 0] / np.sqrt(4 * np.pi)
        for r in np.nonzero(shell_of == s)[0]:
            u = g[r, :3].copy()
            n = np.linalg.norm(u)
            if n > 0:
                u /= n
            d = sh_delta(u, lmax)
            for l in range(0, lmax + 1, 2):
                d[l * (l + 1) // 2 - l:l * (l + 1) // 2 + l + 1] *= rh[l // 2]
            K[r] = d
            Ak[r, 0] = wmr0
            for j in range(nf):
                Ak[r, j + 1] = resp_iso[j][s]
    return K * weight, Ak


class Energy:
    def __init__(self, K, Ak, mu):
        self.K, self.Ak, self.mu = np.asarray(K, np.float64), np.asarray(Ak, np.float64), float(mu)
        self.nnls = orc.IclsOracle(self.Ak, np.eye(self.Ak.shape[1]))

    def voxel(self, dwi, tod):
        """calcEnergy (:245-252): (energy, NNLS solution f)."""
        y = np.asarray(dwi, np.float64) - self.K @ np.asarray(tod, np.float64)
        f, _, _ = self.nnls.solve(y)
        y = y - self.Ak[:, 1:] @ f[1:]
        return float(y @ y / self.K.shape[0] + self.mu * float(tod[0])), f
