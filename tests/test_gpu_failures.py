"""Failure semantics of the HIP Dual solver, driven on the GPU through the runtime caps of
ssn_solve_opts (the reference's module constants, dual_recommended_hvp.py:15-16).

Mirrors the oracle's batch-semantics test: Armijo exhaustion raises the reference's
RuntimeError (:189-190), the Newton cap raises its own (:191), the fallback to the
preconditioned gradient (:163-164) gives the oracle's trajectory, and a non-finite target
ends with status SSN_ST_NONFINITE without disturbing the other targets of its batch. Every
kernel here exits normally with a per-target status; nothing faults or waits.

Tolerances: X and omega within 1e-10 relative of the oracle run with the same caps (the
parity bar of the plain solve), equal Newton / CG / backtrack counts.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import dual as odual
from symplectic_stiefel_nearness_amd import _lib, inputs

pytestmark = pytest.mark.gpu

PARITY_TOL = 1e-10


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _compare_capped(ssn, T, **caps):
    """GPU solve with caps (no raise) against the oracle with the same caps (no raise)."""
    X, omega, iters, st = ssn.solve_device(T.cuda(), return_stats=True, check=False, **caps)
    Xo, omo, ito, sto = odual.solve(T.numpy(), return_stats=True, raise_on_failure=False, **caps)
    X, omega, iters = X.cpu().numpy(), omega.cpu().numpy(), iters.cpu().numpy()
    cnt = st["counters"].cpu().numpy()
    for j in range(T.shape[0]):
        ex, eo = rel(X[j], Xo[j]), rel(omega[j], omo[j])
        say(f"  target {j}: newton {iters[j]}/{ito[j]} cg {cnt[j, 1]}/{sto.cg[j]} "
            f"bt {cnt[j, 2]}/{sto.backtracks[j]} relX {ex:.2e} relOmega {eo:.2e}")
        assert ex <= PARITY_TOL and eo <= PARITY_TOL
        assert int(iters[j]) == int(ito[j])
        assert cnt[j, 1] == sto.cg[j] and cnt[j, 2] == sto.backtracks[j]
    return st["status"].cpu().numpy(), sto


def test_line_search_exhausted_raises(ssn, monkeypatch):
    """MAX_BACKTRACK = 0: no trial step at all, so every target fails its first line search."""
    say("max_backtrack 0")
    T = inputs.make_case("n256_k8", 2)
    _, _, _, st = ssn.solve_device(T.cuda(), return_stats=True, check=False, max_backtrack=0)
    assert (st["status"].cpu().numpy() == _lib.ST_NO_DESCENT).all()
    with pytest.raises(RuntimeError, match=ssn.dual.MSG_NO_DESCENT):
        ssn.solve(T, max_backtrack=0)
    # the module constant, patched as a caller of the reference would patch it
    monkeypatch.setattr(ssn.dual, "MAX_BACKTRACK", 0)
    with pytest.raises(RuntimeError, match=ssn.dual.MSG_NO_DESCENT):
        ssn.solve(T.cuda())
    monkeypatch.setattr(odual, "MAX_BACKTRACK", 0)
    with pytest.raises(RuntimeError, match=odual.MSG_NO_DESCENT):
        odual.solve(T.numpy())


@pytest.mark.parametrize("case,count", [("n256_k8", 2), ("n4096_k48", 1)])
def test_newton_cap_matches_oracle(ssn, monkeypatch, case, count):
    """MAX_NEWTON = 2: the cap error, and the state after two Newton updates equals the oracle's."""
    say("max_newton 2", case)
    T = inputs.make_case(case, count)
    status, sto = _compare_capped(ssn, T, max_newton=2)
    assert (status == _lib.ST_NEWTON_CAP).all() and sto.failed.all()
    with pytest.raises(RuntimeError, match=ssn.dual.MSG_NEWTON_CAP):
        ssn.solve(T, max_newton=2)
    monkeypatch.setattr(ssn.dual, "MAX_NEWTON", 2)
    with pytest.raises(RuntimeError, match=ssn.dual.MSG_NEWTON_CAP):
        ssn.solve(T.cuda())


def test_newton_cap_zero_returns_the_start(ssn):
    """MAX_NEWTON = 0 reaches the cap at once: omega = 0, N = I, so X = T to rounding."""
    T = inputs.make_case("n2048_k16", 2)
    X, omega, iters, st = ssn.solve_device(T.cuda(), return_stats=True, check=False, max_newton=0)
    assert (st["status"].cpu().numpy() == _lib.ST_NEWTON_CAP).all()
    assert int(iters.abs().max()) == 0
    assert float(omega.abs().max()) <= 1e-15
    assert rel(X.cpu().numpy(), T.numpy()) <= 1e-14


@pytest.mark.parametrize("case,count", [("n256_k8", 2), ("n2048_k16", 2), ("n4096_k48", 1)])
def test_fallback_direction_matches_oracle(ssn, case, count):
    """MAX_CG = 0 leaves the Newton direction zero after step 0, so <g, d> >= 0 and every later
    step falls back to the preconditioned gradient (dual_recommended_hvp.py:163-164): the
    staged (m = 16) and on-chip (m = 32, 96) CG paths against the oracle for four updates."""
    say("max_cg 0", case)
    T = inputs.make_case(case, count)
    status, sto = _compare_capped(ssn, T, max_cg=0, max_newton=4)
    assert (sto.cg == 0).all()
    assert ((status == _lib.ST_NEWTON_CAP) | (status == _lib.ST_OK)).all()


def test_tightened_tolerance_matches_oracle(ssn):
    """FEAS_TOL = 1e-12: more Newton updates, the same trajectory as the oracle's."""
    T = inputs.make_case("n2048_k16", 2)
    status, sto = _compare_capped(ssn, T, feas_tol=1e-12)
    assert (status == _lib.ST_OK).all()


def test_nonfinite_target(ssn):
    """A NaN in one target: that target ends with SSN_ST_NONFINITE, solve raises, and the other
    targets of the batch equal those of a batch with a finite target in its place, bit for bit."""
    T = inputs.make_case("n2048_k16", 3)
    Tbad = T.clone()
    Tbad[1, 5, 3] = float("nan")
    say("nan target")
    Xb, ob, ib, sb = ssn.solve_device(Tbad.cuda(), return_stats=True, check=False)
    Xg, og, ig, sg = ssn.solve_device(T.cuda(), return_stats=True, check=False)
    st = sb["status"].cpu().numpy()
    assert st[1] == _lib.ST_NONFINITE and st[0] == _lib.ST_OK and st[2] == _lib.ST_OK
    for j in (0, 2):
        assert torch.equal(Xb[j], Xg[j]) and torch.equal(ob[j], og[j]) and torch.equal(ib[j], ig[j])
    with pytest.raises(RuntimeError):
        ssn.solve(Tbad.cuda())


def test_options_validation(ssn):
    """Out-of-range caps are argument errors (AssertionError on the Python side), not launches."""
    T = inputs.make_case("n256_k8", 1).cuda()
    for bad in (dict(max_newton=-1), dict(max_cg=-2), dict(max_backtrack=5000),
                dict(feas_tol=float("nan")), dict(feas_tol=-1.0)):
        with pytest.raises(AssertionError):
            ssn.solve_device(T, **bad)
    with pytest.raises(TypeError):
        ssn.solve_device(T, max_iterations=3)
