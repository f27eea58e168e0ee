"""Split-precision Gram / recovery (BF16 slices on MFMA, exact FP32 order sums, FP64 reassembly)
against FP64 references, and the split solve against the CPU FP64 oracle.

Tolerances, stated per output:
  Cr, Ci (stage level)         1e-11 of ||Cr|| (Frobenius) against NumPy FP64
  X (stage level)              5e-11 relative (Frobenius) of NumPy FP64 (measured 4e-12 .. 1.2e-11:
                               39 bits below each row / column maximum)
  X, omega of the split solve  1e-9 relative of the oracle (the dual is solved on a Gram
                               perturbed by ~1e-12 and amplifies it by its conditioning)
  acceptance                   r_F <= 1e-8, r_G <= 1e-9, gap <= 1e-8 (CPU FP64)
  dL/dT after a split forward  1e-7 relative of the oracle gradient
Each test prints the measured errors.
"""
import math
import sys

import numpy as np
import pytest
import torch

from oracle import acceptance, adjoint as oadj, dual as odual
from symplectic_stiefel_nearness_amd import inputs

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,m,batch", [(7, 2, 1), (64, 16, 2), (300, 20, 3), (1000, 96, 2), (513, 128, 2),
                                       (4096, 96, 2), (2000, 112, 1)])
def test_split_gram(ssn, n, m, batch):
    g = torch.Generator().manual_seed(n * 31 + m)
    T = torch.randn(batch, 2 * n, m, dtype=torch.float64, generator=g) / math.sqrt(2 * n)
    T[:, :, 0] *= 1e3                                   # columns of very different scales
    if m > 2:
        T[:, :, 1] = 0.0                                # an all-zero column
    Cr, Ci = ssn.gram_chunks(T.cuda(), split=True)
    Tn = T.numpy()
    A, D = Tn[:, :n], Tn[:, n:]
    Cr_ref = np.swapaxes(A, 1, 2) @ A + np.swapaxes(D, 1, 2) @ D
    Ci_ref = np.swapaxes(A, 1, 2) @ D - np.swapaxes(D, 1, 2) @ A
    # errors measured against the Gram's scale ||Cr|| (Ci alone can be much smaller)
    scale = np.linalg.norm(Cr_ref)
    ecr = float(np.linalg.norm(Cr.cpu().numpy() - Cr_ref) / scale)
    eci = float(np.linalg.norm(Ci.cpu().numpy() - Ci_ref) / scale)
    say(f"split gram n={n} m={m}: |dCr|/|Cr| {ecr:.2e}, |dCi|/|Cr| {eci:.2e}")
    assert ecr < 1e-11 and eci < 1e-11
    Crh = Cr.cpu().numpy()
    assert np.array_equal(Crh, np.swapaxes(Crh, 1, 2))


@pytest.mark.parametrize("n,m,batch", [(5, 2, 1), (64, 16, 2), (200, 20, 2), (1000, 96, 2), (300, 128, 1),
                                       (777, 112, 1), (4096, 96, 1)])
def test_split_recover(ssn, n, m, batch):
    lib = ssn.dual._lib.load()
    mp = lib.ssn_padded_order(m)
    g = torch.Generator().manual_seed(11 + n + m)
    T = torch.randn(batch, 2 * n, m, dtype=torch.float64, generator=g)
    T[:, 3, :] *= 1e-6                                  # rows of very different scales
    P = torch.zeros(batch, mp, mp, dtype=torch.float64)
    Q = torch.zeros_like(P)
    P[:, :m, :m] = torch.randn(batch, m, m, dtype=torch.float64, generator=g)
    Q[:, :m, :m] = torch.randn(batch, m, m, dtype=torch.float64, generator=g)
    X = ssn.recover_matmul(T.cuda(), P.cuda(), Q.cuda(), split=True).cpu().numpy()
    X_ref = odual.recover(T.numpy(), P[:, :m, :m].numpy(), Q[:, :m, :m].numpy())
    e = rel(X, X_ref)
    say(f"split recover n={n} m={m}: rel X {e:.2e}")
    assert e < 5e-11
    # the tiny row keeps its own relative accuracy (per-row scale)
    er = rel(X[:, 3], X_ref[:, 3])
    assert er < 5e-11


def _split_vs_oracle(ssn, T, tol=1e-9):
    X, omega, iters, stats = ssn.solve(T, return_stats=True, split=True)
    Xo, omo, ito, sto = odual.solve(T.numpy(), return_stats=True)
    worst_x = worst_o = 0.0
    for j in range(T.shape[0]):
        ex, eo = rel(X[j].numpy(), Xo[j]), rel(omega[j].numpy(), omo[j])
        worst_x, worst_o = max(worst_x, ex), max(worst_o, eo)
        assert ex <= tol and eo <= tol, (j, ex, eo)
    recs = acceptance.inspect_output(T.numpy(), X.numpy())
    for r in recs:
        assert acceptance.numeric_pass(r), r
    say(f"  split vs FP64 oracle: max rel X {worst_x:.2e}, max rel omega {worst_o:.2e}, "
        f"max r_F {max(r['feasibility'] for r in recs):.2e}, max r_G {max(r['stationarity'] for r in recs):.2e}, "
        f"newton {iters.tolist()} vs {list(ito)}")


def test_split_solve_config1(ssn):
    say("split solve n256_k8 x1")
    _split_vs_oracle(ssn, inputs.make_case("n256_k8", 1))


def test_split_solve_config2_subset(ssn):
    say("split solve n2048_k16 x4")
    _split_vs_oracle(ssn, inputs.make_case("n2048_k16", 4))


def test_split_solve_metric_subset(ssn):
    say("split solve n4096_k48 x4")
    _split_vs_oracle(ssn, inputs.make_case("n4096_k48", 4))


def test_split_solve_config5b_subset(ssn):
    say("split solve n8192_k64 x2")
    _split_vs_oracle(ssn, inputs.make_case("n8192_k64", 2))


def test_split_solve_metric_full_batch_acceptance(ssn):
    """All 64 metric targets: the batched split solver passes acceptance on every target."""
    from symplectic_stiefel_nearness_amd import acceptance as dacc
    T = inputs.make_case("n4096_k48", 64).cuda()
    X, omega, iters = ssn.solve_device(T, split=True)
    s = dacc.summary(dacc.inspect(T, X))
    say("  split metric batch:", s)
    assert s["all_pass"]


def test_split_forward_backward(ssn):
    T = inputs.make_targets(300, 10, 2, seeds=[41, 42])
    G = torch.randn(T.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(43))
    Td = T.cuda().requires_grad_(True)
    X, _, _ = ssn.project(Td, split=True)
    torch.autograd.backward(X, G.cuda())
    ref, _, _ = oadj.projection_grad_batch(T.numpy(), G.numpy())
    e = rel(Td.grad.cpu().numpy(), ref)
    say(f"  dL/dT after split forward: rel {e:.2e}")
    assert e < 1e-7


def test_split_refinement_step(ssn):
    """The FP64 residual-and-refinement step of the split solve: with the acceptance thresholds
    no metric target needs a re-solve (results bitwise those of the split solve); with a
    feasibility threshold no split result can meet, every target is re-solved in FP64 and the
    results are bitwise those of the FP64 solve."""
    T = inputs.make_case("n4096_k48", 8).cuda()
    Xs, os_, its = ssn.solve_device(T, split=True)
    X, om, it, st = ssn.solve_device(T, split=True, refine=True, return_stats=True)
    say("  residuals: max rF %.2e rG bound %.2e, refined %s" % (
        float(st["residuals"]["feasibility"].max()), float(st["residuals"]["stationarity"].max()),
        st["refined"].tolist()))
    assert st["refined"].numel() == 0 and bool(st["residuals"]["passed"].all())
    assert torch.equal(X, Xs) and torch.equal(om, os_) and torch.equal(it, its)
    Xf, of, itf = ssn.solve_device(T)
    X2, o2, it2, st2 = ssn.solve_device(T, split=True, refine=dict(feasibility=1e-300), return_stats=True)
    assert st2["refined"].tolist() == list(range(8))
    assert torch.equal(X2, Xf) and torch.equal(o2, of) and torch.equal(it2, itf)
    X3, _, _, st3 = ssn.solve_device(T, split=True, refine=dict(stationarity=1e-300), return_stats=True)
    assert st3["refined"].tolist() == list(range(8)) and torch.equal(X3, Xf)
    with pytest.raises(ValueError):
        ssn.solve_device(T, split=True, refine=dict(delta=1e-8))


def test_refinement_residuals_against_cpu_acceptance(ssn):
    """The refinement step's residuals (Gram and recovery kernels, FP64) against the CPU FP64
    acceptance check: r_F equal to rounding, and the KKT residual, an upper bound of r_G, equal
    to it within rounding at the split solutions (both far below the thresholds) and bounding
    it, within 10 %, when X is moved off the solution by 1e-7 (planted on the host)."""
    T = inputs.make_case("n4096_k48", 4).cuda()
    X, om, _ = ssn.solve_device(T, split=True)
    res = ssn.residuals(T, X, om)
    met = acceptance.inspect_output(T.cpu().numpy(), X.cpu().numpy())
    for b in range(4):
        rf, rg = float(res["feasibility"][b]), float(res["stationarity"][b])
        say(f"  target {b}: rF {rf:.3e} (cpu {met[b]['feasibility']:.3e})  rG bound {rg:.3e} (cpu {met[b]['stationarity']:.3e})")
        assert abs(rf - met[b]["feasibility"]) <= 1e-13
        # at a solution both are rounding (X = X(omega) meets the KKT relation exactly in exact
        # arithmetic, and r_F carries the dual's convergence): equal within a small factor
        assert 0.9 * met[b]["stationarity"] - 1e-13 <= rg <= 1.5 * met[b]["stationarity"] + 1e-13
        assert rf <= 1e-8 and rg <= 1e-9
    g = torch.Generator().manual_seed(5)
    Xp = X + 1e-7 * torch.randn(X.shape, dtype=torch.float64, generator=g).cuda()
    resp = ssn.residuals(T, Xp, om)
    metp = acceptance.inspect_output(T.cpu().numpy(), Xp.cpu().numpy())
    for b in range(4):
        rg, rgc = float(resp["stationarity"][b]), metp[b]["stationarity"]
        say(f"  moved target {b}: rG bound {rg:.3e} (cpu rG {rgc:.3e}), rF {float(resp['feasibility'][b]):.3e} (cpu {metp[b]['feasibility']:.3e})")
        assert rg >= rgc * (1 - 1e-9) and rg <= 1.1 * rgc and rg > 1e-9
        assert abs(float(resp["feasibility"][b]) - metp[b]["feasibility"]) <= 1e-9 * metp[b]["feasibility"] + 1e-13
