"""CPU FP64 acceptance metrics (TEST INFRASTRUCTURE ONLY).

Restates `src/symplectic_projection/experiment.py:105-137` (`inspect_output`) and the
projection it uses, `baselines/geometry.py:12-39` (`prepare_projection`,
`project_with_multiplier`), with `baselines/common.py:97-118` (`standard_j`, `left_j`,
`skew`, `sym`). Thresholds are `reproduce.py:28`.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

THRESHOLDS = {"feasibility": 1e-8, "stationarity": 1e-9, "delta": 1e-8}


def standard_j(k):
    """J_{2k} = [[0, I], [-I, 0]] (baselines/common.py:97-104)."""
    j = np.zeros((2 * k, 2 * k))
    j[:k, k:] = np.eye(k)
    j[k:, :k] = -np.eye(k)
    return j


def left_j(x):
    """J x = [x_bot; -x_top] (baselines/common.py:107-111)."""
    n = x.shape[0] // 2
    return np.concatenate((x[n:], -x[:n]), axis=0)


def _skew(a):
    return 0.5 * (a - a.T)


def _sym(a):
    return 0.5 * (a + a.T)


def project_with_multiplier(x, value):
    """Tangent projection and its multiplier at base point x (geometry.py:22-39).

    Solves (X^T X) L + L (X^T X) = (JX)^T A - A^T JX by eigendecomposition.
    """
    values, vectors = np.linalg.eigh(_sym(x.T @ x))
    if not np.all(np.isfinite(values)) or values.min() <= 0:
        raise FloatingPointError("Gram matrix of the base point is not positive definite")
    denominator = values[:, None] + values[None, :]
    jx = left_j(x)
    coupling = jx.T @ value
    rhs = coupling - coupling.T
    rotated = vectors.T @ rhs @ vectors
    multiplier = _skew(vectors @ (rotated / denominator) @ vectors.T)
    return value - jx @ multiplier, multiplier


def inspect_one(t, y):
    """r_F, r_G, delta for one target t and candidate y (experiment.py:105-137)."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    record = dict(feasibility=None, stationarity=None, delta=None)
    m, n = t.shape[-1], t.shape[-2] // 2
    jm = left_j(np.eye(m))
    if y.shape != t.shape or not np.all(np.isfinite(y)):
        record["exception"] = "ValueError: output has the wrong shape or non-finite values"
        return record
    try:
        error, jy = y - t, left_j(y)
        objective = float(np.sum(error * error))
        scale = max(1.0, float(np.linalg.norm(t)))
        record["feasibility"] = float(np.linalg.norm(y.T @ jy - jm) / math.sqrt(m))
        gradient, multiplier = project_with_multiplier(y, 2 * error)
        omega = multiplier / 2
        if not (np.all(np.isfinite(gradient)) and np.all(np.isfinite(omega))):
            raise FloatingPointError("projected gradient or multiplier is non-finite")
        record["stationarity"] = float(np.linalg.norm(gradient) / (2 * scale))
        kc = np.eye(m, dtype=np.complex128) + 1j * omega
        b = t[:n] + 1j * t[n:]
        factor = np.linalg.cholesky(kc)
        solved = sla.solve_triangular(factor, b.conj().T, lower=True)
        dual = float(np.sum(t * t) - np.sum(np.abs(solved) ** 2) - np.sum(omega * jm))
        record["delta"] = abs((objective - dual) / max(1.0, abs(objective)))
        record["omega"] = omega
    except Exception as exc:  # recorded, as in the reference
        record["exception"] = f"{type(exc).__name__}: {exc}"
    return record


def inspect_output(target, x):
    """Per-target records for a batch (experiment.py:105-137)."""
    return [dict(index=i, **inspect_one(target[i], x[i])) for i in range(len(target))]


def numeric_pass(record, thresholds=THRESHOLDS):
    """experiment.py:100-102."""
    return all(record.get(key) is not None and math.isfinite(record[key])
               and abs(record[key]) <= threshold for key, threshold in thresholds.items())
