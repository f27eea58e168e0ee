"""CPU FP64 implicit-differentiation gradient of the projection T -> X(T) (TEST INFRASTRUCTURE ONLY).

No reference counterpart: the reference Dual is forward-only (every entry point is
@torch.no_grad(), dual_recommended_hvp.py:79, :99, :194, :231). The formula follows from
the optimality condition of the dual, grad phi(omega; C) = J - Im(N C N) = 0 with
N = (I + i omega)^-1 = P + iQ, C = B^H B, B = A + iD, X_c = B N:

    W       = N (B^H G_c) N,                 G_c = G_top + i G_bot
    g       = skew(-Im W)                    (dL/domega)
    lam     = H^-1 g                         (H: Hessian of phi at the solution, skew -> skew)
    K       = N lam N
    dL/dB   = G_c N + 2i B K,                dL/dT = [Re dL/dB ; Im dL/dB]

H(E) = Z - Z^T with Z = P E Sr - Q E Si and S = N C N = Sr + i Si, the same operator as
dual_recommended_hvp.py:74-76 but on the unscaled C. lam is computed here by PCG to a
relative residual of 1e-14 in the eigenbasis of Cr (preconditioner (l_i + l_j)^-1, as the
solver's). Pinned by central finite differences of a tightly converged oracle solve.
"""
from __future__ import annotations

import numpy as np

from . import dual as odual


def _skew(a):
    return (a - np.swapaxes(a, -1, -2)) / 2


def hvp_unscaled(P, Q, Sr, Si, E):
    """Z - Z^T with Z = P E Sr - Q E Si."""
    Z = P @ E @ Sr - Q @ E @ Si
    H = Z - Z.T
    np.fill_diagonal(H, 0.0)
    return H


def solve_hessian(P, Q, Sr, Si, Cr, g, tol=1e-14, max_iter=500):
    """lam = H^-1 g by PCG in the eigenbasis of Cr (Jacobi-like preconditioner of the solver)."""
    values, V = np.linalg.eigh(Cr)
    den = values[:, None] + values[None, :]
    Pe, Qe, Sre, Sie = (V.T @ X @ V for X in (P, Q, Sr, Si))
    rhs = V.T @ g @ V
    x = np.zeros_like(rhs)
    r = rhs.copy()
    z = _skew(r / den)
    p = z.copy()
    rz = np.sum(r * z)
    nrhs = np.linalg.norm(rhs)
    for _ in range(max_iter):
        if np.linalg.norm(r) <= tol * nrhs:
            break
        Hp = hvp_unscaled(Pe, Qe, Sre, Sie, p)
        alpha = rz / np.sum(p * Hp)
        x = x + alpha * p
        r = r - alpha * Hp
        z = _skew(r / den)
        rz_new = np.sum(r * z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    return V @ x @ V.T


def projection_grad(T, G, omega):
    """dL/dT for L = <G, X(T)>, given the dual solution omega (original basis); one target."""
    T = np.asarray(T, np.float64)
    G = np.asarray(G, np.float64)
    n, m = T.shape[0] // 2, T.shape[1]
    B = T[:n] + 1j * T[n:]
    Gc = G[:n] + 1j * G[n:]
    N = np.linalg.inv(np.eye(m) + 1j * omega)
    P, Q = N.real, N.imag
    C = B.conj().T @ B
    S = N @ C @ N
    W = N @ (B.conj().T @ Gc) @ N
    g = _skew(-W.imag)
    lam = solve_hessian(P, Q, S.real, S.imag, C.real, g)
    K = N @ lam @ N
    dB = Gc @ N + 2j * (B @ K)
    return np.concatenate((dB.real, dB.imag), axis=0)


def projection_grad_batch(T, G, feas_tol=1e-12):
    """Batched dL/dT with the dual solved by the oracle to `feas_tol`."""
    T = np.asarray(T, np.float64)
    X, omega, _ = odual.solve(T, feas_tol=feas_tol)
    return np.stack([projection_grad(T[j], G[j], omega[j]) for j in range(T.shape[0])]), X, omega
