"""FP64 acceptance metrics of a solve, batched, on the device or the CPU.

Restates `src/symplectic_projection/experiment.py:105-137` (`inspect_output`) with the
projection of `baselines/geometry.py:22-39` and the thresholds of `reproduce.py:28`:

    r_F   = ||Y^T J Y - J|| / sqrt(m)
    r_G   = ||Proj_Y(2 (Y - T))|| / (2 max(1, ||T||))
    delta = |objective - dual| / max(1, |objective|),  dual = ||T||^2 - ||L^-1 B^H||^2 - <omega, J>

(||L^-1 B^H||^2 is evaluated as Re tr((I + i omega)^-1 B^H B) = tr(P Cr) - tr(Q Ci), the same
number, with the real inverse parts P, Q from the symmetric eigendecomposition of omega^T omega.)

with omega the multiplier of the projection / 2 and L the Cholesky factor of I + i omega.
This is a checker for reports and benchmarks (torch.linalg on FP64); it is not on the
solver's hot path.
"""
from __future__ import annotations

import math

import torch

THRESHOLDS = {"feasibility": 1e-8, "stationarity": 1e-9, "delta": 1e-8}


def _left_j(x):
    n = x.shape[-2] // 2
    return torch.cat((x[..., n:, :], -x[..., :n, :]), dim=-2)


def _skew(a):
    return 0.5 * (a - a.mT)


def _sym(a):
    return 0.5 * (a + a.mT)


@torch.no_grad()
def inspect(T, Y):
    """Per-target metrics for targets T and candidates Y, both float64 [batch, 2n, 2k].

    Returns a dict of float64 tensors [batch] (feasibility, stationarity, delta) and a
    bool tensor `passed` (finite and within the reference thresholds).
    """
    assert T.shape == Y.shape and T.ndim == 3
    T = T.to(torch.float64)
    Y = Y.to(device=T.device, dtype=torch.float64)
    batch, rows, m = T.shape
    n = rows // 2
    eye = torch.eye(m, dtype=torch.float64, device=T.device)
    jm = _left_j(eye)
    error = Y - T
    jy = _left_j(Y)
    objective = (error * error).sum((-2, -1))
    scale = torch.linalg.vector_norm(T, dim=(-2, -1)).clamp_min(1.0)
    feas = torch.linalg.vector_norm(Y.mT @ jy - jm, dim=(-2, -1)) / math.sqrt(m)
    values, vectors = torch.linalg.eigh(_sym(Y.mT @ Y))
    denominator = values[..., :, None] + values[..., None, :]
    value = 2 * error
    coupling = jy.mT @ value
    rhs = coupling - coupling.mT
    rotated = vectors.mT @ rhs @ vectors
    multiplier = _skew(vectors @ (rotated / denominator) @ vectors.mT)
    gradient = value - jy @ multiplier
    stat = torch.linalg.vector_norm(gradient, dim=(-2, -1)) / (2 * scale)
    omega = multiplier / 2
    # ||L^-1 B^H||_F^2 = Re tr((I + i omega)^-1 B^H B) = tr(P Cr) - tr(Q Ci) with
    # (I + i omega)^-1 = P + i Q, P = (I + omega^2)^-1 = (I - omega^T omega)^-1, Q = -omega P,
    # and B^H B = Cr + i Ci. I + i omega is positive definite exactly when every eigenvalue s
    # of omega^T omega is below 1 (its Cholesky factor L exists); P comes from that
    # eigendecomposition, so the check needs no complex factorisation.
    s_vals, s_vecs = torch.linalg.eigh(_sym(omega.mT @ omega))
    pd = (s_vals < 1.0).all(-1)
    P = s_vecs @ (s_vecs.mT / (1.0 - s_vals).clamp_min(1e-300)[..., :, None])
    Q = -omega @ P
    A, D = T[:, :n], T[:, n:]
    Cr = A.mT @ A + D.mT @ D
    Ci = A.mT @ D - D.mT @ A
    solved_sq = (P * Cr).sum((-2, -1)) - (Q.mT * Ci).sum((-2, -1))
    dual = (T * T).sum((-2, -1)) - solved_sq - (omega * jm).sum((-2, -1))
    delta = ((objective - dual) / objective.abs().clamp_min(1.0)).abs()
    delta = torch.where(pd, delta, torch.full_like(delta, float("nan")))
    ok = torch.isfinite(Y).flatten(1).all(1) & (values.min(-1).values > 0)
    passed = ok.clone()
    for key, tensor in (("feasibility", feas), ("stationarity", stat), ("delta", delta)):
        passed &= torch.isfinite(tensor) & (tensor.abs() <= THRESHOLDS[key])
    return dict(feasibility=feas, stationarity=stat, delta=delta, omega=omega, passed=passed)


def summary(metrics):
    """Max of each metric and the all-pass flag, as plain floats."""
    return dict(all_pass=bool(metrics["passed"].all()),
                max_rF=float(metrics["feasibility"].max()),
                max_rG=float(metrics["stationarity"].max()),
                max_gap=float(metrics["delta"].max()))
