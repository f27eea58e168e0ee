"""Tridiagonal eigensolver (every order up to 128; replaces torch.linalg.eigh at
dual_recommended_hvp.py:108) through ssn_dual_eigh, and its Jacobi fallback.

Bar (FP64): eigenvalues within 2e-14 max(1, ||A||) of numpy.linalg.eigvalsh of A = Cr / scale,
ascending; |A V - V diag(lambda)| <= 2e-14 and |V^T V - I| <= 2e-14 entrywise (the accuracy of
a backward-stable dense eigensolver at these orders; the Jacobi forms meet the same bar).
The form used is read back from the workspace (2 = tridiagonal; 1 or 0 = the Jacobi fallback with
V in LDS or in global memory):
random Gram matrices take the tridiagonal path; a multiple of the identity (the tridiagonal
splits at once), a block-diagonal matrix (a split further down) and a matrix with eigenvalues
repeated exactly (the inverse-iteration vectors of a repeated eigenvalue coincide) take the
fallback. Whole-solve parity against the oracle runs through this path in test_gpu_parity.py.
"""
import math
import os
import sys

import numpy as np
import pytest
import torch

from symplectic_stiefel_nearness_amd import _lib, inputs

pytestmark = pytest.mark.gpu

S_COUNT = 22          # workspace slots per target (engine.h)
MAX_SWEEPS = 40


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def target_doubles(mp):
    return S_COUNT * mp * mp + mp + 16 + MAX_SWEEPS * (mp - 1) * mp + mp


def run_eigh(Cr):
    L = _lib.load()
    b, m, _ = Cr.shape
    mp = (m + 15) // 16 * 16
    nbytes = L.ssn_workspace_bytes(b, 1, m)
    ws = torch.zeros(nbytes // 8, dtype=torch.float64, device="cuda")
    rc = L.ssn_dual_eigh(torch.cuda.current_stream().cuda_stream, Cr.data_ptr(), Cr.stride(0), Cr.stride(1), b, m,
                         ws.data_ptr(), nbytes)
    _lib.check(rc, "ssn_dual_eigh")
    torch.cuda.synchronize()
    td = target_doubles(mp)
    w = ws[: b * td].view(b, td).cpu().numpy()
    V = w[:, : mp * mp].reshape(b, mp, mp)
    lam = w[:, S_COUNT * mp * mp: S_COUNT * mp * mp + mp]
    extra = w[:, S_COUNT * mp * mp + mp: S_COUNT * mp * mp + mp + 16]
    return lam, V, extra


def check_eigh(Cr, lam, V, extra, form):
    b, m, _ = Cr.shape
    mp = V.shape[1]
    for j in range(b):
        C = np.tril(Cr[j]) + np.tril(Cr[j], -1).T          # UPLO='L'
        scale = max(1.0, np.trace(C) / m)
        assert abs(extra[j, 0] - scale) <= 1e-15 * scale      # the trace is summed in another order
        A = C / extra[j, 0]
        assert extra[j, 3] == form, (j, extra[j, 3])
        assert extra[j, 1] > 0
        w = np.linalg.eigvalsh(A)
        tol = 2e-14 * max(1.0, np.abs(w).max())
        lj, Vj = lam[j, :m], V[j, :m, :m]
        assert np.all(np.diff(lj) >= 0)
        assert np.abs(lj - w).max() <= tol, np.abs(lj - w).max()
        assert np.all(lam[j, m:] == 1.0) and np.all(V[j, m:, :] == 0) and np.all(V[j, :, m:] == 0)
        res = np.abs(A @ Vj - Vj * lj).max()
        orth = np.abs(Vj.T @ Vj - np.eye(m)).max()
        assert res <= tol and orth <= 2e-14, (j, res, orth)
    say(f"  m={m} form={form} ok")


@pytest.mark.parametrize("n,m,batch", [(16, 2, 3), (40, 6, 2), (300, 16, 2), (300, 20, 2), (2048, 32, 2),
                                       (1000, 50, 2), (4096, 64, 2), (1200, 72, 2), (700, 80, 2),
                                       (4096, 96, 4), (8192, 96, 2), (900, 100, 2), (2048, 112, 2),
                                       (8192, 128, 3)])
def test_tridiagonal_path_on_gram_matrices(ssn, n, m, batch):
    say("tridiagonal eigensolver", n, m, batch)
    g = torch.Generator().manual_seed(31 * n + m)
    T = torch.randn(batch, 2 * n, m, dtype=torch.float64, generator=g) / math.sqrt(2 * n)
    Cr, _ = ssn.gram_chunks(T.cuda())
    lam, V, extra = run_eigh(Cr)
    check_eigh(Cr.cpu().numpy(), lam, V, extra, 2.0)


def _orthogonal(m, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((m, m)))
    return q


@pytest.mark.parametrize("m", [8, 48, 96, 112, 128])
def test_fallback_cases(ssn, m):
    say("fallback cases", m)
    mats = []
    mats.append(2.5 * np.eye(m))                                    # splits at the first reflector
    Q1, Q2 = _orthogonal(m // 2, 1), _orthogonal(m - m // 2, 2)
    B = np.zeros((m, m))
    B[: m // 2, : m // 2] = Q1 @ np.diag(np.linspace(1.0, 2.0, m // 2)) @ Q1.T
    B[m // 2:, m // 2:] = Q2 @ np.diag(np.linspace(1.5, 3.0, m - m // 2)) @ Q2.T
    mats.append(B)                                                  # splits halfway
    Q = _orthogonal(m, 3)
    ev = np.repeat(np.array([1.0, 1.25, 2.0, 3.5]), m // 4)
    mats.append(Q @ np.diag(ev) @ Q.T)                              # eigenvalues repeated m / 4 times
    Cr = torch.tensor(np.stack(mats), dtype=torch.float64, device="cuda")
    lam, V, extra = run_eigh(Cr)
    check_eigh(Cr.cpu().numpy(), lam, V, extra, 1.0 if m <= 96 else 0.0)   # Jacobi, V in LDS up to 96


def test_jacobi_forms_selectable(ssn, monkeypatch):
    """SSN_EIGH_TRI=0 selects the Jacobi eigensolver (form 0: V by the rotation replay)."""
    monkeypatch.setenv("SSN_EIGH_TRI", "0")
    g = torch.Generator().manual_seed(5)
    T = torch.randn(2, 2 * 512, 40, dtype=torch.float64, generator=g) / math.sqrt(1024)
    Cr, _ = ssn.gram_chunks(T.cuda())
    lam, V, extra = run_eigh(Cr)
    check_eigh(Cr.cpu().numpy(), lam, V, extra, 0.0)


def test_solve_through_both_eigensolvers_agrees(ssn, monkeypatch):
    """The whole solve with the tridiagonal eigensolver and with Jacobi: equal iteration counts and
    X, omega within 1e-12 of each other on the metric shape."""
    T = inputs.make_case("n4096_k48", 8).cuda()
    X1, o1, i1, s1 = ssn.solve_device(T, return_stats=True)
    monkeypatch.setenv("SSN_EIGH_TRI", "0")
    X2, o2, i2, s2 = ssn.solve_device(T, return_stats=True)
    assert bool((s1["counters"][:, 5] == 2).all()) and bool((s2["counters"][:, 5] != 2).all())
    assert torch.equal(i1, i2)
    assert torch.equal(s1["counters"][:, :4], s2["counters"][:, :4])
    for j in range(T.shape[0]):
        ex = float(torch.linalg.norm(X1[j] - X2[j]) / torch.linalg.norm(X2[j]))
        eo = float(torch.linalg.norm(o1[j] - o2[j]) / torch.linalg.norm(o2[j]))
        assert ex <= 1e-12 and eo <= 1e-12, (j, ex, eo)
