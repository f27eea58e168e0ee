"""NumPy FP64 restatement of the reference Dual solver (TEST INFRASTRUCTURE ONLY).

Follows `src/symplectic_projection/dual_recommended_hvp.py` operation by operation,
batched over targets with the reference's boolean masks, constants and batch-wide
error semantics. Each fused CUDA kernel of the reference is replaced by the maths
its wrapper docstring defines, in the same operation order:

- `gram_add`        <- gram_triangle_reflect.py:19-36 / gram_triangle_reflect.cu:1-104
- `recover_block`   <- recovery_fused.py:19-30 / recovery_fused.cu:1-41
- `inner_products`  <- hvp_fused.py:22-34 / hvp_fused.cu:1-37
- `outer_skew`      <- hvp_fused.py:37-53 / hvp_fused.cu:82-238
- `sandwich_front`  <- hvp_fused.py:56-69 / hvp_fused.cu:39-80
- `line_change`     <- hvp_fused.py:72-87 / hvp_fused.cu:240-367
- `precondition`    <- hvp_fused.py:110-117 / hvp_fused.cu:369-387
- `cg_update`       <- hvp_fused.py:120-130 / hvp_fused.cu:389-405
- `cg_next`         <- hvp_fused.py:133-143 / hvp_fused.cu:407-420
- `construct_inverse` <- dual_recommended_hvp.py:39-43 (cholesky_solve(I, L) = (L L^T)^-1)

torch.linalg.eigh / cholesky_ex / cholesky_solve (closed cuSOLVER/MAGMA in the
reference) are replaced by LAPACK through NumPy/SciPy; the algorithm is orthogonally
invariant, so eigenvector sign/order conventions do not change the exact trajectory.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

# dual_recommended_hvp.py:13-19
ROW_CHUNK = 1024
FEAS_TOL = 1e-8
MAX_NEWTON, MAX_CG, MAX_BACKTRACK = 100, 150, 40

MSG_NO_DESCENT = "批量线搜索未找到下降步"      # dual_recommended_hvp.py:190
MSG_NEWTON_CAP = "达到批量 Newton 迭代上限"     # dual_recommended_hvp.py:191


def mT(a):
    return np.swapaxes(a, -1, -2)


def J(x):
    """J_{2r} x = [x_bot; -x_top] along the row axis (dual_recommended_hvp.py:22-24)."""
    n = x.shape[-2] // 2
    return np.concatenate((x[..., n:, :], -x[..., :n, :]), axis=-2)


def skew(a):
    """(A - A^T) / 2 (dual_recommended_hvp.py:27-28)."""
    return (a - mT(a)) / 2


def inner(a, b):
    """Per-batch Frobenius inner product (dual_recommended_hvp.py:31-32)."""
    return np.sum(a * b, axis=(-2, -1))


def norm(a):
    """Per-batch Frobenius norm (dual_recommended_hvp.py:35-36)."""
    return np.sqrt(np.sum(a * a, axis=(-2, -1)))


# ---------------------------------------------------------------- fused-kernel maths

def gram_add(A, D, Cr, Ci):
    """Cr += A^T A + D^T D; Ci += A^T D - D^T A (gram_triangle_reflect.py:19-20)."""
    Cr += mT(A) @ A + mT(D) @ D
    Ci += mT(A) @ D - mT(D) @ A


def gram(T):
    """Gram accumulated per 1024-row chunk of each half (dual_recommended_hvp.py:79-96)."""
    batch, rows, m = T.shape
    n = rows // 2
    Cr = np.zeros((batch, m, m))
    Ci = np.zeros((batch, m, m))
    for start in range(0, n, ROW_CHUNK):
        stop = min(start + ROW_CHUNK, n)
        gram_add(T[:, start:stop], T[:, n + start:n + stop], Cr, Ci)
    return Cr, Ci


def recover_block(A, D, P, Q):
    """Ur = A P^T + D Q^T, Ui = D P^T - A Q^T (recovery_fused.cu:1-41)."""
    Pt, Qt = mT(P), mT(Q)
    return A @ Pt + D @ Qt, D @ Pt - A @ Qt


def inner_products(E, Sr, Si):
    """U = E Sr, V = E Si (hvp_fused.py:22-34)."""
    return E @ Sr, E @ Si


def outer_skew(P, Q, U, V):
    """Z - Z^T with Z = P U - Q V, zero diagonal (hvp_fused.cu:82-148)."""
    Z = P @ U - Q @ V
    H = Z - mT(Z)
    idx = np.arange(H.shape[-1])
    H[..., idx, idx] = 0.0
    return H


def sandwich_front(values, Ci, P, Q):
    """Yr = P diag(values) - Q Ci, Yi = P Ci + Q diag(values) (hvp_fused.cu:39-80)."""
    d = values[..., None, :]
    return P * d - Q @ Ci, P @ Ci + Q * d


def precondition(R, denominator):
    """skew(R / denominator), both positions read (hvp_fused.cu:374-387)."""
    a = R / denominator
    return (a - mT(a)) * 0.5


def cg_update(direction, r, p, Hp, alpha):
    """d + alpha p, r - alpha Hp (hvp_fused.cu:389-405)."""
    a = alpha[:, None, None]
    return direction + a * p, r - a * Hp


def cg_next(z, p, beta, active):
    """active ? z + beta p : 0 (hvp_fused.cu:407-420)."""
    nxt = z + beta[:, None, None] * p
    return np.where(active[:, None, None], nxt, 0.0)


def line_change(gradient, delta, V, delta_Sr, trial, delta_Si):
    """<g, delta> + <V, delta Sr + trial delta Si>, two separate sums (hvp_fused.cu:240-367)."""
    tail = np.sum(V * (delta_Sr + trial @ delta_Si), axis=(-2, -1))
    first = np.sum(gradient * delta, axis=(-2, -1))
    return first + tail


def cholesky_ex(K):
    """Batched lower Cholesky with per-matrix info (torch.linalg.cholesky_ex semantics)."""
    L = np.zeros_like(K)
    info = np.zeros(K.shape[0], dtype=np.int64)
    for b in range(K.shape[0]):
        try:
            L[b] = np.linalg.cholesky(K[b])
        except np.linalg.LinAlgError:
            info[b] = 1
    return L, info


def construct_inverse(L):
    """(L L^T)^-1 by two triangular solves, i.e. cholesky_solve(I, L) (dual_recommended_hvp.py:39-43)."""
    m = L.shape[-1]
    eye = np.eye(m)
    out = np.empty_like(L)
    for b in range(L.shape[0]):
        y = sla.solve_triangular(L[b], eye, lower=True)
        out[b] = sla.solve_triangular(L[b].T, y, lower=False)
    return out


def to_eigenbasis(Ci, vectors):
    """skew(V^T Ci V), skew(V^T J(V)) (dual_recommended_hvp.py:53-55)."""
    return skew(mT(vectors) @ Ci @ vectors), skew(mT(vectors) @ J(vectors))


def from_eigenbasis(omega, P, Q, vectors):
    """Rotate back and re-factor L = chol(I + omega^2) (dual_recommended_hvp.py:58-64)."""
    omega = skew(vectors @ omega @ mT(vectors))
    P = vectors @ P @ mT(vectors)
    Q = vectors @ Q @ mT(vectors)
    m = omega.shape[-1]
    L = np.linalg.cholesky(np.eye(m) + omega @ omega)
    return omega, L, P, Q


def sandwich(values, Ci, P, Q):
    """S = N C N^H in the eigenbasis; Sr symmetrised, Si skew (dual_recommended_hvp.py:67-71)."""
    Yr, Yi = sandwich_front(values, Ci, P, Q)
    Sr, Si = recover_block(Yr, Yi, P, Q)
    return (Sr + mT(Sr)) / 2, skew(Si)


def hvp(P, Q, Sr, Si, E):
    """Dual Hessian-vector product (dual_recommended_hvp.py:74-76)."""
    U, V = inner_products(E, Sr, Si)
    return outer_skew(P, Q, U, V)


# ---------------------------------------------------------------- the solver

@dataclass
class DualStats:
    """Per-target trajectory counters (not in the reference; used for parity logs)."""
    newton: np.ndarray
    cg: np.ndarray
    backtracks: np.ndarray
    hvps: np.ndarray
    failed: np.ndarray = field(default=None)


def dual_batch(Cr, Ci, raise_on_failure=True, feas_tol=None, max_newton=None, max_cg=None,
               max_backtrack=None):
    """Newton / PCG / Armijo on the dual (dual_recommended_hvp.py:99-191).

    Returns (omega, L, iterations, P, Q, stats); P, Q are the final inverse parts in
    the original basis (the reference's `final` cache, :131-133). `feas_tol`, `max_newton`,
    `max_cg` and `max_backtrack` override the module constants FEAS_TOL, MAX_NEWTON, MAX_CG and
    MAX_BACKTRACK (:15-16) for this call (the tests tighten the tolerance to converge the dual
    to rounding level, and lower the caps to reach the failure branches).
    """
    feas_tol = FEAS_TOL if feas_tol is None else feas_tol
    max_newton = MAX_NEWTON if max_newton is None else max_newton
    max_cg = MAX_CG if max_cg is None else max_cg
    max_backtrack = MAX_BACKTRACK if max_backtrack is None else max_backtrack
    batch, m, _ = Cr.shape
    I = np.broadcast_to(np.eye(m), (batch, m, m))
    omega = np.zeros_like(Cr)
    L = I.copy()
    P = I.copy()
    scale = np.maximum(np.diagonal(Cr, axis1=-2, axis2=-1).sum(-1) / m, 1.0)   # :106
    Cr = Cr / scale[:, None, None]
    Ci = Ci / scale[:, None, None]                                              # :107
    values, vectors = np.linalg.eigh(Cr)                                        # :108
    denominator = values[:, :, None] + values[:, None, :]                       # :109
    Ci, Jm = to_eigenbasis(Ci, vectors)                                         # :110
    Jscaled = Jm / scale[:, None, None]                                         # :111

    active = np.ones(batch, dtype=bool)
    iterations = np.zeros(batch, dtype=np.int64)
    stats = DualStats(np.zeros(batch, np.int64), np.zeros(batch, np.int64),
                      np.zeros(batch, np.int64), np.zeros(batch, np.int64),
                      np.zeros(batch, bool))
    Q = np.zeros_like(omega)
    for newton_step in range(max_newton):                                       # :118
        if newton_step == 0:
            Q = np.zeros_like(omega)
            Sr = np.zeros_like(Cr)
            idx = np.arange(m)
            Sr[:, idx, idx] = values
            Si = Ci
        else:
            Q = -omega @ P                                                      # :124
            Sr, Si = sandwich(values, Ci, P, Q)                                 # :125
        gradient = Jscaled - Si                                                 # :126
        gradient_norm = norm(gradient)
        residual = gradient_norm * scale / m ** 0.5                             # :128
        active &= residual >= feas_tol / 2                                      # :129
        if not bool(active.any()):                                              # :130
            omega, L, P, Q = from_eigenbasis(omega, P, Q, vectors)
            stats.newton[:] = iterations
            return omega, L, iterations, P, Q, stats
        iterations += active                                                    # :135

        if newton_step == 0:                                                    # :138-140
            initial_z = precondition(np.where(active[:, None, None], -gradient, 0.0), denominator)
            direction = initial_z
        else:                                                                   # :141-162
            direction = np.zeros_like(gradient)
            r = np.where(active[:, None, None], -gradient, 0.0)
            z = precondition(r, denominator)
            initial_z = z
            p, rz = z.copy(), inner(r, z)
            tolerance = np.clip(np.sqrt(residual), 1e-8, 0.1) * gradient_norm
            cg_active = active.copy()
            for _ in range(max_cg):
                Hp = hvp(P, Q, Sr, Si, p)
                stats.hvps += cg_active
                stats.cg += cg_active
                curvature = inner(p, Hp)
                cg_active &= curvature > 0
                alpha = np.where(cg_active, rz / np.where(cg_active, curvature, 1.0), 0.0)
                direction, r = cg_update(direction, r, p, Hp, alpha)
                cg_active &= norm(r) > tolerance
                if not bool(cg_active.any()):
                    break
                z = precondition(r, denominator)
                next_rz = inner(r, z)
                beta = next_rz / np.where(cg_active, rz, 1.0)
                p = cg_next(z, p, beta, cg_active)
                rz = next_rz
        fallback = active & (inner(gradient, direction) >= 0)                   # :163
        direction = np.where(fallback[:, None, None], initial_z, direction)

        slope = inner(gradient, direction)                                      # :167
        step = np.ones_like(slope)
        pending = active.copy()
        for _ in range(max_backtrack):                                          # :170
            stats.backtracks += pending
            delta = np.where(pending, step, 0.0)[:, None, None] * direction
            trial = skew(omega + delta)
            new_L, info = cholesky_ex(I + trial @ trial)                        # :173
            valid = info == 0
            safe_L = np.where(valid[:, None, None], new_L, L)
            P_trial = construct_inverse(safe_L)
            V = P_trial @ delta                                                 # :177
            delta_Sr, delta_Si = inner_products(delta, Sr, Si)
            change = line_change(gradient, delta, V, delta_Sr, trial, delta_Si)
            accept = pending & valid & (change <= 1e-4 * step * slope)          # :180
            omega = np.where(accept[:, None, None], trial, omega)
            L = np.where(accept[:, None, None], new_L, L)
            P = np.where(accept[:, None, None], P_trial, P)
            pending &= ~accept
            if not bool(pending.any()):
                break
            step = np.where(pending, step / 2, step)
        else:
            stats.failed |= pending
            if raise_on_failure:
                raise RuntimeError(MSG_NO_DESCENT)
            active &= ~pending
    stats.failed |= active
    if raise_on_failure:
        raise RuntimeError(MSG_NEWTON_CAP)
    # cap reached without raising: the state after the last update, as the GPU reports it
    omega, L, P, Q = from_eigenbasis(omega, P, Q, vectors)
    stats.newton[:] = iterations
    return omega, L, iterations, P, Q, stats


def recover(T, P, Q):
    """X = [A P^T + D Q^T; D P^T - A Q^T] (dual_recommended_hvp.py:194-228)."""
    n = T.shape[-2] // 2
    top, bot = recover_block(T[:, :n], T[:, n:], P, Q)
    return np.concatenate((top, bot), axis=-2)


def solve(T, return_stats=False, feas_tol=None, raise_on_failure=True, **caps):
    """(X, omega, iterations) for targets T [batch, 2n, 2k] (dual_recommended_hvp.py:231-242).

    caps: max_newton, max_cg, max_backtrack, as in dual_batch. raise_on_failure=False returns
    the state reached when the Newton cap is hit instead of raising (the line-search failure
    still drops the failed targets from the active set, as dual_batch does)."""
    T = np.asarray(T, dtype=np.float64)
    if T.ndim != 3 or T.shape[-2] % 2:
        raise ValueError("T must be [batch, 2n, 2k]")
    Cr, Ci = gram(T)
    omega, L, iterations, P, Q, stats = dual_batch(Cr, Ci, raise_on_failure=raise_on_failure,
                                                   feas_tol=feas_tol, **caps)
    X = recover(T, P, Q)
    if return_stats:
        return X, omega, iterations, stats
    return X, omega, iterations


# ---------------------------------------------------------------- dual objective helpers

def dual_objective(C, omega, Jm):
    """phi(omega) = tr(N C) + <omega, J>, N = (I + i omega)^-1, for one target (complex form)."""
    m = omega.shape[-1]
    N = np.linalg.inv(np.eye(m) + 1j * omega)
    return float(np.real(np.trace(N @ C))) + float(np.sum(omega * Jm))
