"""The ICLS oracle against the reference's own golden vectors (Matlab quadprog solutions in
testing/cmd/testing_unit_tests_icls.cpp:46-131, checked there with isApprox(.., 1e-6)), plus
KKT conditions and brute-force enumeration on small random problems."""
import itertools
import os

import numpy as np
import pytest

from oracle import orc

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "icls_golden.npz"))


def is_approx(x, y, prec=1e-6):
    # Eigen's isApprox: ||x - y|| <= prec * min(||x||, ||y||)
    return np.linalg.norm(x - y) <= prec * min(np.linalg.norm(x), np.linalg.norm(y))


def golden_cases():
    H, Ai, Ae = GOLD["H"], GOLD["A_ineq"], GOLD["A_eq"]
    ti, te = GOLD["t_ineq"], GOLD["t_eq"]
    A = np.vstack([Ai, Ae])
    t = np.concatenate([ti, te])
    return [
        ("test1_eq", dict(H=H, A=A, t=t, num_eq=Ae.shape[0]), GOLD["x_eq"]),
        ("test2_eq_zero_t", dict(H=H, A=A, t=None, num_eq=Ae.shape[0]), GOLD["x_eq_zero"]),
        ("test3_no_eq", dict(H=H, A=Ai, t=ti), GOLD["x_no_eq"]),
        ("test4_no_eq_zero_t", dict(H=H, A=Ai, t=None), GOLD["x_no_eq_zero"]),
        ("test5_concat", dict(H=H, A=A, t=t, num_eq=Ae.shape[0]), GOLD["x_eq"]),
    ]


@pytest.mark.parametrize("name,kw,ref", golden_cases(), ids=[c[0] for c in golden_cases()])
def test_icls_golden(name, kw, ref):
    o = orc.IclsOracle(**kw)
    x, it, act = o.solve(GOLD["b"])
    assert is_approx(x, ref, 1e-6), (name, np.linalg.norm(x - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("name,kw,ref", golden_cases()[:4], ids=[c[0] for c in golden_cases()[:4]])
def test_icls_golden_long_double_agrees(name, kw, ref):
    x, _, _ = orc.IclsOracle(**kw).solve(GOLD["b"])
    xl, _, _ = orc.IclsOracle(long_double=True, **kw).solve(GOLD["b"])
    assert np.linalg.norm(x - xl) <= 1e-9 * np.linalg.norm(xl)


def kkt_residuals(H, A, t, num_eq, b, x, active):
    """Stationarity and complementarity of min ||Hx-b||^2 s.t. A_i x >= t_i (i < m-num_eq),
    A_eq x = t_eq, with multipliers recovered by least squares on the active rows."""
    m = A.shape[0]
    t = np.zeros(m) if t is None else t
    g = H.T @ (H @ x - b)
    act = np.nonzero(active)[0]
    if act.size:
        lam, *_ = np.linalg.lstsq(A[act].T, g, rcond=None)
        stat = np.linalg.norm(g - A[act].T @ lam) / max(1.0, np.linalg.norm(g))
    else:
        lam = np.zeros(0)
        stat = np.linalg.norm(g) / max(1.0, np.linalg.norm(H.T @ b))
    viol = A[: m - num_eq] @ x - t[: m - num_eq]
    ineq_act = [i for i, a in enumerate(act) if a < m - num_eq]
    return stat, viol.min(), (lam[ineq_act].min() if ineq_act else 0.0)


@pytest.mark.parametrize("seed", range(6))
def test_icls_kkt_random(seed):
    rng = np.random.default_rng(seed)
    nmeas, n, m = 40, 12, 60
    H = rng.standard_normal((nmeas, n))
    A = rng.standard_normal((m, n))
    x0 = rng.standard_normal(n)
    t = A @ x0 - rng.uniform(0.1, 1.0, m)      # x0 strictly feasible
    b = rng.standard_normal(nmeas) * 3
    # con_reg as MSMT-CSD uses it: without it a degenerate active set (more active rows than
    # unknowns) makes B_act B_act^T singular, as it does in the reference
    o = orc.IclsOracle(H, A, t, con_reg=1e-10)
    x, it, active = o.solve(b)
    stat, minviol, minlam = kkt_residuals(H, A, t, 0, b, x, active)
    # the 1e-10 constraint regularisation perturbs active rows by O(1e-10 |lambda|)
    assert stat < 1e-7
    assert minviol > -1e-7
    assert minlam > -1e-9


@pytest.mark.parametrize("seed", range(5))
def test_icls_brute_force(seed):
    """Enumerate every active set of a tiny problem; the feasible stationary point with the
    smallest objective is the unique optimum of this strictly convex QP."""
    rng = np.random.default_rng(100 + seed)
    nmeas, n, m = 8, 3, 6
    H = rng.standard_normal((nmeas, n))
    A = rng.standard_normal((m, n))
    x0 = rng.standard_normal(n)
    t = A @ x0 - rng.uniform(0.05, 0.5, m)     # x0 strictly feasible
    b = rng.standard_normal(nmeas) * 2
    best, bestx = np.inf, None
    Q = H.T @ H
    for k in range(0, n + 1):
        for S in itertools.combinations(range(m), k):
            S = list(S)
            K = np.block([[Q, -A[S].T], [A[S], np.zeros((k, k))]]) if k else Q
            rhs = np.concatenate([H.T @ b, t[S]]) if k else H.T @ b
            try:
                sol = np.linalg.solve(K, rhs)
            except np.linalg.LinAlgError:
                continue
            x = sol[:n]
            if np.all(A @ x - t >= -1e-10):
                f = 0.5 * np.sum((H @ x - b) ** 2)
                if f < best - 1e-12:
                    best, bestx = f, x
    x, _, _ = orc.IclsOracle(H, A, t).solve(b)
    assert np.allclose(x, bestx, atol=1e-8, rtol=1e-8)


def test_icls_validation_messages():
    H = np.eye(3)
    with pytest.raises(ValueError, match="solution norm regularisation is negative"):
        orc.IclsOracle(H, np.eye(3), sol_reg=-1.0)
    with pytest.raises(ValueError, match="constraint norm regularisation is negative"):
        orc.IclsOracle(H, np.eye(3), con_reg=-1.0)
    with pytest.raises(ValueError, match="tolerance is negative"):
        orc.IclsOracle(H, np.eye(3), tol=-1.0)
