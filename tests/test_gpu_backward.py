"""GPU backward of the projection (dL/dT by implicit differentiation) against the CPU FP64
implicit-differentiation oracle and central finite differences.

Tolerances: 1e-8 relative (Frobenius, per target) against the oracle gradient computed at a dual
solution converged to rounding level; central differences with step 1e-6 on 16 random directions
within 1e-4 relative (the forward is converged to its residual ~1e-11, which bounds the
difference quotient's accuracy at ~1e-5).
"""
import math
import sys

import numpy as np
import pytest
import torch

from oracle import adjoint as oadj
from symplectic_stiefel_nearness_amd import inputs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,k,batch", [(20, 3, 2), (256, 8, 2), (300, 10, 1), (4096, 32, 1)])
def test_backward_matches_oracle(ssn, n, k, batch):
    say("backward vs oracle", n, k, batch)
    T = inputs.make_targets(n, k, batch, seeds=[77 + j + n for j in range(batch)])
    g = torch.Generator().manual_seed(5 + n)
    G = torch.randn(T.shape, dtype=torch.float64, generator=g)
    Td = T.cuda().requires_grad_(True)
    X, omega, iters = ssn.project(Td)
    (X * G.cuda()).sum().backward()
    dT = Td.grad.cpu().numpy()
    ref, _, _ = oadj.projection_grad_batch(T.numpy(), G.numpy())
    for j in range(batch):
        e = rel(dT[j], ref[j])
        say(f"  target {j}: rel err vs oracle {e:.2e}")
        assert e < 1e-8


def test_backward_central_differences_config4_subset(ssn):
    """n=4096, k=32 (config 4 shape), two seeded targets, 16 random directions, step 1e-6."""
    T = inputs.make_case("n4096_k32", 2).cuda()
    G = inputs.make_gradients(4096, 32, 2).cuda()
    Td = T.clone().requires_grad_(True)
    X, _, _ = ssn.project(Td)
    (X * G).sum().backward()
    dT = Td.grad
    g = torch.Generator(device="cpu").manual_seed(123)
    h = 1e-6
    worst = 0.0
    for d in range(16):
        E = torch.randn(T.shape, dtype=torch.float64, generator=g).cuda()
        Xp = ssn.solve(T + h * E)[0]
        Xm = ssn.solve(T - h * E)[0]
        fd = ((G * Xp).sum() - (G * Xm).sum()).item() / (2 * h)
        an = (dT * E).sum().item()
        err = abs(fd - an) / max(abs(an), 1e-12)
        worst = max(worst, err)
        assert err < 1e-4, (d, fd, an)
    say(f"  worst relative FD mismatch over 16 directions: {worst:.2e}")


def test_backward_invariant_at_orthosymplectic(ssn):
    """X(cS) = S for all c in (0, 2) => <dL/dT, S> = 0 at T = cS."""
    n, k = 64, 4
    rng = np.random.default_rng(9)
    Z = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    Qc, _ = np.linalg.qr(Z)
    S = np.block([[Qc.real, -Qc.imag], [Qc.imag, Qc.real]])
    for c in (0.7, 1.3):
        Td = torch.from_numpy((c * S)[None].copy()).cuda().requires_grad_(True)
        G = torch.randn(Td.shape, dtype=torch.float64, device="cuda")
        X, _, _ = ssn.project(Td)
        (X * G).sum().backward()
        dT = Td.grad
        assert abs((dT * torch.from_numpy(S).cuda()).sum().item()) < 1e-9 * dT.norm().item()


def test_backward_cpu_input_and_autograd_chain(ssn):
    """CPU input goes through autograd end to end (to device and back)."""
    T = inputs.make_targets(40, 3, 2, seeds=[1, 2]).requires_grad_(True)
    X, omega, iters = ssn.project(T)
    assert X.device.type == "cpu" and not omega.requires_grad
    loss = (X ** 2).sum()
    loss.backward()
    ref, _, _ = oadj.projection_grad_batch(T.detach().numpy(), 2 * X.detach().numpy())
    assert rel(T.grad.numpy(), ref) < 1e-8


@pytest.mark.parametrize("case,count,split", [("n4096_k48", 2, False), ("n4096_k48", 2, True),
                                              ("n8192_k64", 1, False), ("n4096_k32", 8, False)])
def test_backward_timed_shapes_match_oracle(ssn, case, count, split):
    """The backward at the shapes that are timed: n=4096, k=48 (m = 96, chained adjoint) after an
    FP64 and after a split forward; n=8192, k=64 (m = 128, staged adjoint); and config 4 as one
    batched launch of 8 n=4096, k=32 targets. dL/dT within 1e-8 of the oracle's implicit gradient."""
    n, k = (int(x[1:]) for x in case.split("_"))
    say("backward timed shape", case, count, "split" if split else "fp64")
    T = inputs.make_case(case, count)
    G = inputs.make_gradients(n, k, count)
    Td = T.cuda().requires_grad_(True)
    X, omega, iters = ssn.project(Td, split=split)
    (X * G.cuda()).sum().backward()
    dT = Td.grad.cpu().numpy()
    ref, _, _ = oadj.projection_grad_batch(T.numpy(), G.numpy())
    for j in range(count):
        e = rel(dT[j], ref[j])
        say(f"  target {j}: rel err vs oracle {e:.2e}")
        assert e < 1e-8


@pytest.mark.parametrize("n,k", [(1024, 40), (700, 56)])
def test_backward_odd_orders_match_oracle(ssn, n, k):
    """Odd tile orders (m = 80: NT = 5; m = 112: NT = 7) take the staged adjoint."""
    say("backward odd order", n, k)
    T = inputs.make_targets(n, k, 2, seeds=[31 + n, 32 + n])
    G = torch.randn(T.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(k))
    Td = T.cuda().requires_grad_(True)
    X, _, _ = ssn.project(Td)
    (X * G.cuda()).sum().backward()
    ref, _, _ = oadj.projection_grad_batch(T.numpy(), G.numpy())
    for j in range(2):
        e = rel(Td.grad[j].cpu().numpy(), ref[j])
        say(f"  target {j}: rel err vs oracle {e:.2e}")
        assert e < 1e-8


def test_backward_central_differences_metric_shape(ssn):
    """n=4096, k=48 (m = 96, the timed forward+backward), two seeded targets, 8 random directions,
    central differences with step 1e-6, within 1e-4 relative."""
    say("backward FD n4096_k48")
    T = inputs.make_case("n4096_k48", 2).cuda()
    G = inputs.make_gradients(4096, 48, 2).cuda()
    Td = T.clone().requires_grad_(True)
    X, _, _ = ssn.project(Td)
    (X * G).sum().backward()
    dT = Td.grad
    g = torch.Generator(device="cpu").manual_seed(321)
    h = 1e-6
    worst = 0.0
    for d in range(8):
        E = torch.randn(T.shape, dtype=torch.float64, generator=g).cuda()
        Xp = ssn.solve(T + h * E)[0]
        Xm = ssn.solve(T - h * E)[0]
        fd = ((G * Xp).sum() - (G * Xm).sum()).item() / (2 * h)
        an = (dT * E).sum().item()
        err = abs(fd - an) / max(abs(an), 1e-12)
        worst = max(worst, err)
        assert err < 1e-4, (d, fd, an)
    say(f"  worst relative FD mismatch over 8 directions: {worst:.2e}")


def test_backward_reports_adjoint_counters(ssn):
    """backward_device returns the adjoint's per-target counters: converged, within the CG cap."""
    T = inputs.make_case("n4096_k48", 2).cuda()
    G = inputs.make_gradients(4096, 48, 2).cuda()
    X, omega, iters, ws, _ = ssn.dual._solve_device_ws(T)
    dT, st = ssn.dual.backward_device(T, G, ws, return_stats=True)
    cnt = st["counters"].cpu().numpy()
    say("  adjoint CG iterations", cnt[:, 0].tolist())
    assert (cnt[:, 1] == 0).all() and (cnt[:, 0] > 0).all() and (cnt[:, 0] <= 150).all()
