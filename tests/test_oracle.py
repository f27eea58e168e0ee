"""Pins of the CPU FP64 oracle (CPU only).

The reference ships no tests or golden vectors and its Dual cannot run here (CuPy/NVRTC and an
NVIDIA GPU, `cuda_kernels.py:3-12`), so each oracle function is checked against what the
mathematics fixes: an independent formulation, finite differences, closed forms, known answers,
brute force on tiny inputs and invariances, chosen so that a dropped term, a wrong sign or a
transposed operand fails.
"""
import math

import mpmath
import numpy as np
import pytest
import scipy.optimize as so

from oracle import acceptance, dual as odual
from symplectic_stiefel_nearness_amd import inputs


def rng(seed):
    return np.random.default_rng(seed)


def rand_skew(g, m, s=1.0):
    a = g.standard_normal((m, m)) * s
    return (a - a.T) / 2


def orthosymplectic(g, n, k):
    """[[Re Q, -Im Q], [Im Q, Re Q]] for Q (n x k) with orthonormal columns."""
    z = g.standard_normal((n, k)) + 1j * g.standard_normal((n, k))
    q, _ = np.linalg.qr(z)
    return np.block([[q.real, -q.imag], [q.imag, q.real]])


def phi(C, omega, Jm):
    """phi(omega) = Re tr((I + i omega)^-1 C) + <omega, J>, evaluated directly in complex form."""
    m = omega.shape[0]
    N = np.linalg.inv(np.eye(m) + 1j * omega)
    return float(np.real(np.trace(N @ C))) + float(np.sum(omega * Jm))


def eigen_setup(seed, m):
    """Eigenbasis quantities: values (Cr diagonal), skew Ci, skew J, skew omega in the domain."""
    g = rng(seed)
    values = np.sort(g.uniform(0.5, 1.5, m))
    Ci = rand_skew(g, m, 0.3)
    Jm = rand_skew(g, m, 1.0)
    omega = rand_skew(g, m, 0.2)
    return g, values, Ci, Jm, omega


def parts(omega):
    m = omega.shape[0]
    L = np.linalg.cholesky(np.eye(m) + omega @ omega)
    P = odual.construct_inverse(L[None])[0]
    Q = -omega @ P
    return P, Q


# ------------------------------------------------------------------ Gram

def test_gram_equals_complex_form():
    """Cr + i Ci = B^H B with B = A + iD (independent formulation), across 1024-row chunks."""
    g = rng(0)
    n, m = 2500, 6
    T = g.standard_normal((2, 2 * n, m))
    Cr, Ci = odual.gram(T)
    B = T[:, :n] + 1j * T[:, n:]
    C = np.conj(np.swapaxes(B, 1, 2)) @ B
    assert np.allclose(Cr, C.real, rtol=0, atol=1e-11)
    assert np.allclose(Ci, C.imag, rtol=0, atol=1e-11)
    assert np.allclose(Cr, np.swapaxes(Cr, 1, 2)) and np.allclose(Ci, -np.swapaxes(Ci, 1, 2))


# ------------------------------------------------------------------ inverse parts, sandwich, gradient

def test_construct_inverse_and_parts():
    g = rng(1)
    m = 10
    omega = rand_skew(g, m, 0.3)
    P, Q = parts(omega)
    N = np.linalg.inv(np.eye(m) + 1j * omega)
    assert np.allclose(P, N.real, atol=1e-14) and np.allclose(Q, N.imag, atol=1e-14)


def test_sandwich_is_N_C_N():
    g, values, Ci, Jm, omega = eigen_setup(2, 8)
    P, Q = parts(omega)
    Sr, Si = odual.sandwich(values, Ci, P, Q)
    N = P + 1j * Q
    S = N @ (np.diag(values) + 1j * Ci) @ N.conj().T
    assert np.allclose(Sr, S.real, atol=1e-13) and np.allclose(Si, S.imag, atol=1e-13)


def test_gradient_matches_finite_differences():
    """grad phi = J - Si: directional derivatives by central differences of the complex form."""
    g, values, Ci, Jm, omega = eigen_setup(3, 8)
    C = np.diag(values) + 1j * Ci
    P, Q = parts(omega)
    Sr, Si = odual.sandwich(values, Ci, P, Q)
    grad = Jm - Si
    for _ in range(4):
        E = rand_skew(g, 8)
        h = 1e-5
        fd = (phi(C, omega + h * E, Jm) - phi(C, omega - h * E, Jm)) / (2 * h)
        assert abs(fd - np.sum(grad * E)) < 1e-8 * max(1.0, abs(fd))


def test_hvp_finite_differences_symmetry_and_closed_form():
    g, values, Ci, Jm, omega = eigen_setup(4, 8)
    P, Q = parts(omega)
    Sr, Si = odual.sandwich(values, Ci, P, Q)

    def grad_at(w):
        Pw, Qw = parts(w)
        return Jm - odual.sandwich(values, Ci, Pw, Qw)[1]

    E = rand_skew(g, 8)
    F = rand_skew(g, 8)
    H = odual.hvp(P[None], Q[None], Sr[None], Si[None], E[None])[0]
    h = 1e-5
    fd = (grad_at(omega + h * E) - grad_at(omega - h * E)) / (2 * h)
    assert np.linalg.norm(H - fd) < 1e-7 * np.linalg.norm(H)
    HF = odual.hvp(P[None], Q[None], Sr[None], Si[None], F[None])[0]
    assert abs(np.sum(F * H) - np.sum(HF * E)) < 1e-12 * np.linalg.norm(H) * np.linalg.norm(F)
    assert np.sum(E * H) > 0                      # positive definite inside the domain
    # omega = 0 in the eigenbasis: H(E)_ij = (lambda_i + lambda_j) E_ij
    I = np.eye(8)[None]
    H0 = odual.hvp(I, 0 * I, np.diag(values)[None], Ci[None], E[None])[0]
    assert np.allclose(H0, (values[:, None] + values[None, :]) * E, atol=1e-14)


def test_precondition_inverts_the_omega0_hessian():
    g, values, Ci, Jm, omega = eigen_setup(5, 8)
    den = values[:, None] + values[None, :]
    R = rand_skew(g, 8)
    Z = odual.precondition(R[None], den[None])[0]
    assert np.allclose(den * Z, R, atol=1e-14)


# ------------------------------------------------------------------ line search change

def test_line_change_equals_objective_difference_mpmath():
    """change = phi(omega + delta) - phi(omega) exactly; checked in 50-digit arithmetic (m = 4)."""
    mpmath.mp.dps = 50
    g, values, Ci, Jm, omega = eigen_setup(6, 4)
    P, Q = parts(omega)
    Sr, Si = odual.sandwich(values, Ci, P, Q)
    grad = Jm - Si
    delta = rand_skew(g, 4, 0.05)
    trial = odual.skew(omega + delta)
    Lt = np.linalg.cholesky(np.eye(4) + trial @ trial)
    Pt = odual.construct_inverse(Lt[None])[0]
    V = Pt @ delta
    dSr, dSi = delta @ Sr, delta @ Si
    change = odual.line_change(grad[None], delta[None], V[None], dSr[None], trial[None], dSi[None])[0]

    def mp_phi(w):
        M = mpmath.matrix(4, 4)
        C = mpmath.matrix(4, 4)
        for i in range(4):
            for j in range(4):
                M[i, j] = (1 if i == j else 0) + 1j * mpmath.mpf(w[i, j])
                C[i, j] = (mpmath.mpf(values[i]) if i == j else 0) + 1j * mpmath.mpf(Ci[i, j])
        N = M ** -1
        tr = sum((N * C)[i, i] for i in range(4))
        return mpmath.re(tr) + sum(mpmath.mpf(w[i, j]) * mpmath.mpf(Jm[i, j]) for i in range(4) for j in range(4))

    exact = mp_phi(trial) - mp_phi(omega)
    assert abs(float(exact) - change) < 1e-15 * max(1.0, abs(change)) + 1e-17


# ------------------------------------------------------------------ whole solve

@pytest.mark.parametrize("c", [1.0, 0.5, 1.7])
def test_known_answer_orthosymplectic(c):
    """T = c S (S orthosymplectic): X = S, omega = (c-1) J, dual = primal = 2k (c-1)^2."""
    g = rng(7)
    n, k = 40, 4
    S = orthosymplectic(g, n, k)
    T = (c * S)[None]
    X, omega, iters = odual.solve(T)
    J = acceptance.standard_j(k)
    assert np.linalg.norm(X[0] - S) < 1e-10
    assert np.abs(omega[0] - (c - 1) * J).max() < 1e-10
    if c == 1.0:
        assert iters[0] == 0
    rec = acceptance.inspect_one(T[0], X[0])
    assert acceptance.numeric_pass(rec)
    primal = np.sum((X[0] - T[0]) ** 2)
    assert abs(primal - 2 * k * (c - 1) ** 2) < 1e-10


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (3, 2)])
def test_brute_force_slsqp_tiny(n, seed):
    """k = 1: X^T J X = J is the single constraint x1^T J x2 = 1; SLSQP from many starts."""
    g = rng(100 + seed)
    T = g.standard_normal((2 * n, 2)) / math.sqrt(2 * n)
    X, omega, iters = odual.solve(T[None])
    Jn = np.block([[np.zeros((n, n)), np.eye(n)], [-np.eye(n), np.zeros((n, n))]])

    def f(v):
        return np.sum((v.reshape(2 * n, 2) - T) ** 2)

    def con(v):
        x = v.reshape(2 * n, 2)
        return x[:, 0] @ Jn @ x[:, 1] - 1.0

    best = None
    for s in range(40):
        x0 = g.standard_normal(4 * n)
        res = so.minimize(f, x0, constraints=[{"type": "eq", "fun": con}], method="SLSQP",
                          options={"ftol": 1e-15, "maxiter": 500})
        if res.success and abs(con(res.x)) < 1e-10 and (best is None or res.fun < best.fun):
            best = res
    assert best is not None
    primal = f(X[0].ravel())
    assert primal <= best.fun + 1e-9                 # the dual solution is the global minimum
    assert np.linalg.norm(best.x.reshape(2 * n, 2) - X[0]) < 1e-5


def test_left_and_right_orthosymplectic_invariance():
    g = rng(8)
    n, k = 30, 3
    T = g.standard_normal((2 * n, 2 * k)) / math.sqrt(2 * n)
    X, omega, _ = odual.solve(T[None])
    U = orthosymplectic(g, n, n)            # 2n x 2n
    W = orthosymplectic(g, k, k)            # 2k x 2k
    XU, omegaU, _ = odual.solve((U @ T)[None])
    XW, omegaW, _ = odual.solve((T @ W)[None])
    assert np.linalg.norm(XU[0] - U @ X[0]) < 1e-9
    assert np.linalg.norm(omegaU[0] - omega[0]) < 1e-9
    assert np.linalg.norm(XW[0] - X[0] @ W) < 1e-9
    assert np.linalg.norm(omegaW[0] - W.T @ omega[0] @ W) < 1e-9


def test_stationarity_relation_and_acceptance():
    """X - T = J X omega at the solution, and the seeded n=256, k=8 targets pass acceptance."""
    T = inputs.make_case("n256_k8", 3).numpy()
    X, omega, iters = odual.solve(T)
    for j in range(3):
        lhs = X[j] - T[j]
        rhs = odual.J(X[j][None])[0] @ omega[j]
        assert np.linalg.norm(lhs - rhs) < 1e-7 * np.linalg.norm(lhs)
    for rec in acceptance.inspect_output(T, X):
        assert acceptance.numeric_pass(rec)
        assert np.abs(rec["omega"] - omega[rec["index"]]).max() < 1e-9


def test_acceptance_rejects_planted_mistakes():
    g = rng(9)
    n, k = 20, 2
    S = orthosymplectic(g, n, k)
    T = 1.3 * S
    assert acceptance.numeric_pass(acceptance.inspect_one(T, S))
    assert not acceptance.numeric_pass(acceptance.inspect_one(T, T))            # infeasible
    Sbad = S.copy(); Sbad[:, 0] *= -1                                            # wrong sign
    assert not acceptance.numeric_pass(acceptance.inspect_one(T, Sbad))
    # feasible but not stationary: rotate S inside the manifold
    W = orthosymplectic(g, k, k)
    assert not acceptance.numeric_pass(acceptance.inspect_one(T, S @ W))


def test_batch_error_semantics(monkeypatch):
    """Batch-wide RuntimeError on line-search exhaustion and on the Newton cap (:189-191)."""
    T = inputs.make_case("n256_k8", 2).numpy()
    monkeypatch.setattr(odual, "MAX_BACKTRACK", 0)
    with pytest.raises(RuntimeError, match=odual.MSG_NO_DESCENT):
        odual.solve(T)
    monkeypatch.setattr(odual, "MAX_BACKTRACK", 40)
    monkeypatch.setattr(odual, "MAX_NEWTON", 2)
    with pytest.raises(RuntimeError, match=odual.MSG_NEWTON_CAP):
        odual.solve(T)


def test_batched_equals_per_target():
    """The masked batch gives each target its single-target trajectory (per-target engines rely on it)."""
    T = inputs.make_case("n2048_k16", 3).numpy()
    Xb, ob, ib, sb = odual.solve(T, return_stats=True)
    for j in range(3):
        Xs, os_, is_, ss = odual.solve(T[j:j + 1], return_stats=True)
        assert np.array_equal(ib[j:j + 1], is_)
        assert ss.cg[0] == sb.cg[j] and ss.backtracks[0] == sb.backtracks[j]
        assert np.linalg.norm(Xs[0] - Xb[j]) <= 1e-13 * np.linalg.norm(Xb[j])


# ------------------------------------------------------------------ adjoint (backward)

from oracle import adjoint as oadj


@pytest.mark.parametrize("n,k,seed", [(6, 1, 0), (20, 3, 1), (40, 4, 2)])
def test_adjoint_matches_central_differences(n, k, seed):
    """<dL/dT, E> against (L(T + hE) - L(T - hE)) / 2h, dual converged to rounding level."""
    g = rng(200 + seed)
    T = g.standard_normal((1, 2 * n, 2 * k)) / math.sqrt(2 * n)
    G = g.standard_normal((1, 2 * n, 2 * k))
    dT, X, omega = oadj.projection_grad_batch(T, G)
    h = 1e-6
    for _ in range(4):
        E = g.standard_normal(T.shape)
        Xp, _, _ = odual.solve(T + h * E, feas_tol=1e-12)
        Xm, _, _ = odual.solve(T - h * E, feas_tol=1e-12)
        fd = (np.sum(G * Xp) - np.sum(G * Xm)) / (2 * h)
        an = np.sum(dT * E)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


def test_adjoint_orthogonal_to_the_manifold_point_at_T_eq_cS():
    """X(cS) = S for every c in (0, 2), so <dL/dT, S> = 0 at T = cS for every G."""
    g = rng(11)
    n, k = 16, 2
    S = orthosymplectic(g, n, k)
    for c in (0.6, 1.4):
        G = g.standard_normal((1, 2 * n, 2 * k))
        dT, _, _ = oadj.projection_grad_batch((c * S)[None], G)
        assert abs(np.sum(dT[0] * S)) < 1e-9 * np.linalg.norm(dT)


def test_adjoint_hessian_solve_is_exact():
    """The PCG inside the adjoint solves H lam = g to 1e-14 (checked with a dense H)."""
    g = rng(12)
    n, k = 10, 2
    T = g.standard_normal((1, 2 * n, 2 * k)) / math.sqrt(2 * n)
    _, omega, _ = odual.solve(T, feas_tol=1e-12)
    m = 2 * k
    B = T[0, :n] + 1j * T[0, n:]
    N = np.linalg.inv(np.eye(m) + 1j * omega[0])
    C = B.conj().T @ B
    S = N @ C @ N
    rhs = _skew_np(g.standard_normal((m, m)))
    lam = oadj.solve_hessian(N.real, N.imag, S.real, S.imag, C.real, rhs)
    res = oadj.hvp_unscaled(N.real, N.imag, S.real, S.imag, lam) - rhs
    assert np.linalg.norm(res) < 1e-12 * np.linalg.norm(rhs)


def _skew_np(a):
    return (a - a.T) / 2
