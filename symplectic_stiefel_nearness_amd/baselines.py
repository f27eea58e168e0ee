"""Comparison solvers of the reference experiment's single-target table, written from the
published methods (no counterpart on the hot path; the reference calls them per target at
src/symplectic_projection/experiment.py:82-87 with the options of reproduce.py:29-32):

* `solve_gao_nonmonotone`: Riemannian steepest descent on the symplectic Stiefel manifold
  Sp(2n, 2k) = {X : X^T J_2n X = J_2k} after Gao, Son, Absil and Stykel, "Riemannian optimization
  on the symplectic Stiefel manifold" (SIAM J. Matrix Anal. Appl. 42, 2021): the gradient of the
  canonical-like metric, grad f = G X^T X + J X G^T J X (G the Euclidean gradient), the Cayley
  retraction, and a non-monotone (Zhang-Hager) Armijo line search on alternating
  Barzilai-Borwein steps.
* `solve_jensen_trustregions`: a second-order Riemannian trust-region method (after Jensen and
  Zimmermann's use of second-order information on Sp(2n, 2k)): the manifold with the Euclidean
  metric, whose Riemannian Hessian of f = 1/2 ||X - T||^2 is Hess f(X)[Z] = P_X(Z - J Z Omega)
  (Omega the multiplier of the gradient's projection), the Cayley retraction, and the
  Steihaug-Toint truncated CG for the trust-region model.

Objective f(X) = 1/2 ||X - T||_F^2, FP64, any device (torch). Both report the iterate, the
multiplier omega of X - T = J X omega at it (the Dual solver's omega at the solution), a status
and the iteration count. The stopping test is the acceptance one: feasibility
||X^T J X - J|| / sqrt(2k) <= feas_tol and stationarity ||P_X(2 (X - T))|| / (2 max(1, ||T||))
<= grad_tol.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
import torch


@dataclass
class SolverOptions:
    """Budgets and tolerances of a comparison solve (field names as the reference experiment's
    protocol passes them, reproduce.py:29-32)."""
    max_iter: int = 300
    grad_tol: float = 1e-8
    feas_tol: float = 1e-8
    time_limit: float = 120.0
    history: bool = True

    def __post_init__(self) -> None:
        if self.max_iter < 0:
            raise ValueError("max_iter must be non-negative")
        if not all(math.isfinite(x) and x > 0 for x in (self.grad_tol, self.feas_tol, self.time_limit)):
            raise ValueError("tolerances and the time limit must be finite and positive")


@dataclass
class TrustRegionOptions(SolverOptions):
    """SolverOptions plus the trust-region parameters (radius bound relative to sqrt(dim),
    acceptance ratio, truncated-CG budget)."""
    radius_scale: float = 1.0
    accept_ratio: float = 0.1
    max_inner: int = 200


@dataclass
class SolveResult:
    x: torch.Tensor | None
    omega: torch.Tensor | None
    status: str
    iterations: int
    history: list = field(default_factory=list)
    stats: dict = field(default_factory=dict)

    @property
    def success(self) -> bool:
        return self.status == "converged"


# ---------------------------------------------------------------- geometry
def left_j(x):
    """J_2r x = [x_bot; -x_top] along the rows."""
    r = x.shape[-2] // 2
    return torch.cat((x[..., r:, :], -x[..., :r, :]), dim=-2)


def right_j(x):
    """x J_2c = [-x_right, x_left] along the columns (J = [[0, I], [-I, 0]])."""
    c = x.shape[-1] // 2
    return torch.cat((-x[..., c:], x[..., :c]), dim=-1)


def _skew(a):
    return 0.5 * (a - a.mT)


def _sym(a):
    return 0.5 * (a + a.mT)


def _inner(a, b):
    return float((a * b).sum())


def _decrease(X, Xn, T):
    """f(X) - f(Xn) for f = 1/2 ||. - T||^2, from the step D = Xn - X: -<D, (X - T) + D / 2>. Near
    a minimum the decrease is far below the rounding of f itself, so it is never formed as a
    difference of two values of f."""
    D = Xn - X
    return -_inner(D, (X - T) + 0.5 * D)


class Projector:
    """Euclidean-metric projection onto T_X Sp at a fixed X: the normal space is {J X W : W skew},
    and P_X(G) = G - J X W with (X^T X) W + W (X^T X) = -2 skew(X^T J G), solved in the
    eigenbasis of X^T X."""

    def __init__(self, X):
        self.X = X
        self.JX = left_j(X)
        vals, self.V = torch.linalg.eigh(_sym(X.mT @ X))
        self.den = vals[:, None] + vals[None, :]

    def multiplier(self, G):
        rhs = -2.0 * _skew(self.X.mT @ left_j(G))
        return _skew(self.V @ ((self.V.mT @ rhs @ self.V) / self.den) @ self.V.mT)

    def __call__(self, G):
        W = self.multiplier(G)
        return G - self.JX @ W, W


def cayley(X, Z):
    """Cayley retraction of the tangent vector Z at X (Gao et al.): with W = G_X Z,
    G_X = I - 1/2 X J X^T J^T and S = W (X J)^T + (X J) W^T, the point
    (I - S J / 2)^-1 (I + S J / 2) X, evaluated through S = U V^T (U = [W, XJ], V = [XJ, W]) and
    the Woodbury identity (one 4k x 4k solve)."""
    m = X.shape[-1]
    XJ = right_j(X)
    W = Z + 0.5 * (X @ left_j(X.mT @ left_j(Z)))       # G_X Z, with J^T Z = -J Z
    U = torch.cat((W, XJ), dim=-1)
    Vt = -left_j(torch.cat((XJ, W), dim=-1))          # J^T V
    Y0 = X + 0.5 * (U @ (Vt.mT @ X))
    small = torch.eye(2 * m, dtype=X.dtype, device=X.device) - 0.5 * (Vt.mT @ U)
    return Y0 + 0.5 * (U @ torch.linalg.solve(small, Vt.mT @ Y0))


def feasibility(X):
    m = X.shape[-1]
    eye = torch.eye(m, dtype=X.dtype, device=X.device)
    return float(torch.linalg.vector_norm(X.mT @ left_j(X) - left_j(eye)) / math.sqrt(m))


def stationarity(X, T, proj=None):
    proj = proj or Projector(X)
    g, _ = proj(2.0 * (X - T))
    return float(torch.linalg.vector_norm(g)) / (2.0 * max(1.0, float(torch.linalg.vector_norm(T))))


def feasible_start(T):
    """A point of Sp(2n, 2k) in the column span of T: X0 = T M with M^T (T^T J T) M = J, M from
    the eigenvectors of the Hermitian i T^T J T (each eigenvalue s > 0 with eigenvector a + ib
    gives the pair sqrt(2/s) a, -sqrt(2/s) b of a Darboux basis)."""
    K = T.mT @ left_j(T)                               # real skew, 2k x 2k
    vals, vecs = torch.linalg.eigh(1j * K.to(torch.complex128))
    k = K.shape[-1] // 2
    s, u = vals[k:].real, vecs[:, k:]                   # the k positive eigenvalues
    if not bool((s > 0).all()):
        raise ValueError("T^T J T is singular: no symplectic basis in the span of T")
    scale = torch.sqrt(2.0 / s).to(T.dtype)
    M = torch.cat((u.real * scale, -u.imag * scale), dim=-1)
    return T @ M


def _omega(X, T, proj):
    """omega with X - T = J X omega + (tangent part): the projection multiplier of X - T."""
    return proj.multiplier(X - T)


def _result(X, T, status, it, history, t0, **stats):
    proj = Projector(X)
    stats.update(seconds=time.perf_counter() - t0, feasibility=feasibility(X), stationarity=stationarity(X, T, proj))
    return SolveResult(X, _omega(X, T, proj), status, it, history, stats)


def _check_target(T):
    if not (isinstance(T, torch.Tensor) and T.ndim == 2 and T.shape[0] % 2 == 0 and T.shape[1] % 2 == 0):
        raise ValueError("T must be a 2n x 2k tensor")
    if T.dtype != torch.float64:
        raise ValueError("T must be float64")
    if not bool(torch.isfinite(T).all()):
        raise ValueError("T must be finite")


# ---------------------------------------------------------------- Riemannian steepest descent
@torch.no_grad()
def solve_gao_nonmonotone(T, options=None, x0=None, rho=1e-4, shrink=0.1, eta=0.85, step0=1.0,
                          step_min=1e-10, step_max=1e10):
    """Riemannian steepest descent with the canonical-like gradient, the Cayley retraction and
    the Zhang-Hager non-monotone Armijo test f(R(t d)) <= C + rho t <P_X grad_E f, d> with
    C_{k+1} = (eta Q_k C_k + f_{k+1}) / Q_{k+1}, Q_{k+1} = eta Q_k + 1 (carried as the gap
    C_k - f_k, updated from the step's decrease); trial steps alternate the two
    Barzilai-Borwein lengths of s = X_{k+1} - X_k, y = grad_{k+1} - grad_k."""
    options = options or SolverOptions()
    _check_target(T)
    t0 = time.perf_counter()
    X = feasible_start(T) if x0 is None else x0.clone()
    f = 0.5 * float(((X - T) ** 2).sum())
    gap, Q = 0.0, 1.0                                   # C - f of the Zhang-Hager reference value C
    history = []
    step = step0
    grad_prev = X_prev = None
    for it in range(options.max_iter + 1):
        proj = Projector(X)
        rg, rf = stationarity(X, T, proj), feasibility(X)
        if options.history:
            history.append(dict(iteration=it, objective=f, stationarity=rg, feasibility=rf))
        if rg <= options.grad_tol and rf <= options.feas_tol:
            return _result(X, T, "converged", it, history, t0)
        if it == options.max_iter:
            return _result(X, T, "max_iter", it, history, t0)
        if time.perf_counter() - t0 > options.time_limit:
            return _result(X, T, "time_limit", it, history, t0)
        # the canonical-like gradient vanishes on the normal space {J X W}, so it is formed from
        # the tangent part of G = X - T: no cancellation between its two O(||G||) terms
        Gt, _ = proj(X - T)
        grad = Gt @ (X.mT @ X) + left_j(X) @ (Gt.mT @ left_j(X))
        if grad_prev is not None:
            s, y = X - X_prev, grad - grad_prev
            sy = abs(_inner(s, y))
            if sy > 0:
                step = _inner(s, s) / sy if it % 2 else sy / _inner(y, y)
            step = min(max(step, step_min), step_max)
        d = -grad
        slope = _inner(Gt, d)                           # d/dt f(R(t d)) at t = 0 (d is tangent)
        if not slope < 0:
            return _result(X, T, "no_descent", it, history, t0)
        t = step
        while True:
            Xn = cayley(X, t * d)
            dec = _decrease(X, Xn, T)                   # f - f(Xn)
            if dec >= -gap - rho * t * slope:           # f(Xn) <= C + rho t slope
                break
            t *= shrink
            if t < 1e-20:
                return _result(X, T, "line_search_failed", it, history, t0)
        X_prev, grad_prev = X, grad
        X, f = Xn, f - dec
        Qn = eta * Q + 1.0
        gap = eta * Q * (gap + dec) / Qn                # C_{k+1} - f_{k+1}
        Q = Qn
    return _result(X, T, "max_iter", options.max_iter, history, t0)


# ---------------------------------------------------------------- Riemannian trust region
def _hessian(proj, Omega):
    """Hess f(X)[Z] = P_X(Z - J Z Omega) for f = 1/2 ||X - T||^2 with the Euclidean metric
    (Omega: multiplier of X - T; the Weingarten term of the embedded manifold)."""
    return lambda Z: proj(Z - left_j(Z) @ Omega)[0]


def _truncated_cg(grad, hess, radius, max_inner, kappa=0.1, theta=1.0):
    """Steihaug-Toint truncated CG for min <grad, eta> + 1/2 <eta, H eta>, ||eta|| <= radius."""
    eta = torch.zeros_like(grad)
    r = grad.clone()
    r0 = math.sqrt(_inner(r, r))
    d = -r
    rr = _inner(r, r)
    for inner in range(max_inner):
        Hd = hess(d)
        dHd = _inner(d, Hd)
        if dHd <= 0:                                    # negative curvature: go to the boundary
            return _to_boundary(eta, d, radius), inner + 1, "negative_curvature"
        alpha = rr / dHd
        eta_n = eta + alpha * d
        if math.sqrt(_inner(eta_n, eta_n)) >= radius:
            return _to_boundary(eta, d, radius), inner + 1, "boundary"
        eta = eta_n
        r = r + alpha * Hd
        rr_n = _inner(r, r)
        if math.sqrt(rr_n) <= r0 * min(r0 ** theta, kappa):
            return eta, inner + 1, "converged"
        d = -r + (rr_n / rr) * d
        rr = rr_n
    return eta, max_inner, "max_inner"


def _to_boundary(eta, d, radius):
    """eta + tau d with ||eta + tau d|| = radius, tau >= 0."""
    a, b, c = _inner(d, d), 2 * _inner(eta, d), _inner(eta, eta) - radius ** 2
    tau = (-b + math.sqrt(max(b * b - 4 * a * c, 0.0))) / (2 * a)
    return eta + tau * d


@torch.no_grad()
def solve_jensen_trustregions(T, options=None, x0=None):
    """Riemannian trust region with the exact Riemannian Hessian (Euclidean metric), the Cayley
    retraction and truncated CG; radius bound radius_scale sqrt(dim Sp), start at 1/8 of it,
    shrink by 4 when the model ratio is below 1/4, double at the boundary above 3/4, accept
    steps whose ratio exceeds accept_ratio."""
    options = options or TrustRegionOptions()
    if not isinstance(options, TrustRegionOptions):
        options = TrustRegionOptions(**{k: getattr(options, k) for k in
                                        ("max_iter", "grad_tol", "feas_tol", "time_limit", "history")})
    _check_target(T)
    t0 = time.perf_counter()
    X = feasible_start(T) if x0 is None else x0.clone()
    rows, m = T.shape
    dim = rows * m - m * (m - 1) // 2                   # dim Sp(2n, 2k) = 4nk - k(2k - 1)
    radius_max = options.radius_scale * math.sqrt(dim)
    radius = radius_max / 8
    f = 0.5 * float(((X - T) ** 2).sum())
    history, inner_total = [], 0
    for it in range(options.max_iter + 1):
        proj = Projector(X)
        grad, Omega = proj(X - T)
        rg, rf = stationarity(X, T, proj), feasibility(X)
        if options.history:
            history.append(dict(iteration=it, objective=f, stationarity=rg, feasibility=rf, radius=radius))
        if rg <= options.grad_tol and rf <= options.feas_tol:
            return _result(X, T, "converged", it, history, t0, inner=inner_total)
        if it == options.max_iter:
            return _result(X, T, "max_iter", it, history, t0, inner=inner_total)
        if time.perf_counter() - t0 > options.time_limit:
            return _result(X, T, "time_limit", it, history, t0, inner=inner_total)
        hess = _hessian(proj, Omega)
        eta, inner, why = _truncated_cg(grad, hess, radius, options.max_inner)
        inner_total += inner
        Xn = cayley(X, eta)
        dec = _decrease(X, Xn, T)
        model = -(_inner(grad, eta) + 0.5 * _inner(eta, hess(eta)))
        ratio = dec / model if model > 0 else -1.0
        if ratio < 0.25:
            radius /= 4
        elif ratio > 0.75 and why in ("boundary", "negative_curvature"):
            radius = min(2 * radius, radius_max)
        if ratio > options.accept_ratio:
            X, f = Xn, f - dec
        if radius < 1e-300:
            return _result(X, T, "radius_collapsed", it, history, t0, inner=inner_total)
    return _result(X, T, "max_iter", options.max_iter, history, t0, inner=inner_total)
