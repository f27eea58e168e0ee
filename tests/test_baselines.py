"""The comparison solvers (Riemannian steepest descent with a non-monotone line search, and a
second-order Riemannian trust region) and their geometry, on the CPU in FP64.

The geometry is pinned by what the mathematics fixes: the Cayley retraction stays on
Sp(2n, 2k) and is X + t Z + O(t^2); the projection is idempotent, lands in the tangent space
(X^T J P(G) symmetric) and removes only normal vectors J X W; the canonical-like gradient is
tangent, a descent direction and zero on normal vectors; the Riemannian Hessian matches the
second derivative of f(R_X(t Z)) at a critical point. The solvers are pinned by the oracle's
(Dual) solution of the same targets and by the known answer T = c S -> X = S.
"""
import math

import numpy as np
import pytest
import torch

from oracle import acceptance, dual as odual
from symplectic_stiefel_nearness_amd import baselines as bl, inputs


def rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def random_point(n, k, seed):
    T = inputs.make_targets(n, k, 1, seeds=[seed])[0]
    return bl.feasible_start(T), T


def random_tangent(X, seed):
    g = torch.Generator().manual_seed(seed)
    G = torch.randn(X.shape, dtype=torch.float64, generator=g)
    return bl.Projector(X)(G)[0]


def orthosymplectic(n, k, seed):
    g = np.random.default_rng(seed)
    Z = g.standard_normal((n, k)) + 1j * g.standard_normal((n, k))
    Qc, _ = np.linalg.qr(Z)
    return torch.from_numpy(np.block([[Qc.real, -Qc.imag], [Qc.imag, Qc.real]]))


def test_feasible_start_is_on_the_manifold():
    for n, k, seed in ((10, 1, 1), (40, 3, 2), (256, 8, 3)):
        X, T = random_point(n, k, seed)
        assert bl.feasibility(X) < 1e-13


def test_projection_properties():
    X, _ = random_point(30, 3, 4)
    P = bl.Projector(X)
    G = torch.randn(X.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(5))
    Z, W = P(G)
    S = X.mT @ bl.left_j(Z)
    assert torch.allclose(S, S.mT, atol=1e-12)                       # tangent
    assert torch.allclose(W, -W.mT, atol=0)                          # skew multiplier
    assert torch.allclose(G - Z, bl.left_j(X) @ W, atol=1e-13)       # removed part is normal
    Z2, W2 = P(Z)
    assert rel(Z2, Z) < 1e-13 and float(W2.abs().max()) < 1e-12      # idempotent
    N = bl.left_j(X) @ (W - W.mT)                                    # a normal vector
    assert float(torch.linalg.vector_norm(P(N)[0])) < 1e-12


def test_cayley_retraction_is_feasible_and_first_order():
    X, _ = random_point(25, 2, 6)
    Z = random_tangent(X, 7)
    for t in (1e-3, 0.1, 1.0, 5.0):
        assert bl.feasibility(bl.cayley(X, t * Z)) < 1e-12
    errs = [float(torch.linalg.vector_norm((bl.cayley(X, t * Z) - X) / t - Z)) for t in (1e-3, 1e-4)]
    assert errs[1] < errs[0] / 5 and errs[1] < 1e-3 * float(torch.linalg.vector_norm(Z))
    assert torch.equal(bl.cayley(X, torch.zeros_like(X)), X) or rel(bl.cayley(X, torch.zeros_like(X)), X) < 1e-15


def test_canonical_gradient_descent_and_normal_kernel():
    for seed in range(4):
        X, T = random_point(20, 2, 10 + seed)
        G = X - T
        grad = G @ (X.mT @ X) + bl.left_j(X) @ (G.mT @ bl.left_j(X))
        S = X.mT @ bl.left_j(grad)
        assert torch.allclose(S, S.mT, atol=1e-11)                   # tangent
        assert float((G * grad).sum()) > 0                           # descent direction
        W = torch.randn(4, 4, dtype=torch.float64, generator=torch.Generator().manual_seed(seed))
        N = bl.left_j(X) @ (W - W.mT)
        gn = N @ (X.mT @ X) + bl.left_j(X) @ (N.mT @ bl.left_j(X))
        assert float(torch.linalg.vector_norm(gn)) < 1e-12 * float(torch.linalg.vector_norm(N))


def test_riemannian_hessian_matches_the_pullback_at_the_solution():
    """At a critical point the Hessian of f o R_X at 0 is the Riemannian Hessian for any
    retraction: <Z, Hess f[Z]> against a central second difference of f(R_X(t Z))."""
    T = inputs.make_targets(16, 2, 1, seeds=[21])[0]
    res = bl.solve_jensen_trustregions(T, bl.TrustRegionOptions(grad_tol=1e-13, feas_tol=1e-13, max_iter=100))
    X = res.x
    proj = bl.Projector(X)
    _, Om = proj(X - T)
    hess = bl._hessian(proj, Om)
    f = lambda Y: 0.5 * float(((Y - T) ** 2).sum())
    for seed in range(3):
        Z = random_tangent(X, 30 + seed)
        Z = Z / torch.linalg.vector_norm(Z)
        h = 1e-4
        fd = (f(bl.cayley(X, h * Z)) - 2 * f(X) + f(bl.cayley(X, -h * Z))) / h ** 2
        an = float((Z * hess(Z)).sum())
        assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (fd, an)
        HZ = hess(Z)
        S = X.mT @ bl.left_j(HZ)
        assert torch.allclose(S, S.mT, atol=1e-11)                   # Hess maps into T_X
        Y = random_tangent(X, 40 + seed)
        assert abs(float((Y * hess(Z)).sum()) - float((Z * hess(Y)).sum())) < 1e-10   # self-adjoint


@pytest.mark.parametrize("solver", ["gao_nonmonotone", "jensen_rtr2"])
@pytest.mark.parametrize("n,k,seed", [(20, 2, 1), (64, 4, 2), (256, 8, 3)])
def test_solvers_reach_the_dual_solution(solver, n, k, seed):
    T = inputs.make_targets(n, k, 1, seeds=[seed])[0]
    opts = dict(max_iter=1000, grad_tol=1e-9, feas_tol=1e-8, time_limit=60.0, history=False)
    if solver == "gao_nonmonotone":
        res = bl.solve_gao_nonmonotone(T, bl.SolverOptions(**opts))
    else:
        res = bl.solve_jensen_trustregions(T, bl.TrustRegionOptions(**opts))
    assert res.success, res.status
    Xo, omo, _ = odual.solve(T.numpy()[None])
    assert rel(res.x, torch.from_numpy(Xo[0])) < 1e-7
    assert rel(res.omega, torch.from_numpy(omo[0])) < 1e-6
    assert acceptance.numeric_pass(acceptance.inspect_output(T.numpy()[None], res.x.numpy()[None])[0])


@pytest.mark.parametrize("solver", [bl.solve_gao_nonmonotone, bl.solve_jensen_trustregions])
def test_known_answer(solver):
    """T = c S with S orthosymplectic: the nearest point is S and omega = (c - 1) J."""
    S = orthosymplectic(30, 3, 8)
    J = bl.left_j(torch.eye(6, dtype=torch.float64))
    for c in (0.6, 1.4):
        res = solver(c * S)
        assert res.success and rel(res.x, S) < 1e-8
        assert float((res.omega - (c - 1) * J).abs().max()) < 1e-8


def test_options_and_inputs_are_checked():
    with pytest.raises(ValueError):
        bl.SolverOptions(max_iter=-1)
    with pytest.raises(ValueError):
        bl.SolverOptions(grad_tol=0.0)
    with pytest.raises(ValueError):
        bl.SolverOptions(time_limit=float("inf"))
    with pytest.raises(ValueError):
        bl.solve_gao_nonmonotone(torch.zeros(5, 2, dtype=torch.float64))
    with pytest.raises(ValueError):
        bl.solve_jensen_trustregions(torch.zeros(6, 2, dtype=torch.float32))
    T = inputs.make_targets(10, 1, 1, seeds=[3])[0]
    res = bl.solve_gao_nonmonotone(T, bl.SolverOptions(max_iter=2, history=True))
    assert res.status == "max_iter" and res.iterations == 2 and len(res.history) == 3


def test_trust_region_near_the_optimum_accepts_its_steps():
    """Regression: on this target the decrease of the last steps is below the rounding of f, so a
    ratio formed from two values of f rejected every step and the radius collapsed; the decrease
    formed from the step converges."""
    T = inputs.make_case("n4096_k48", 3)[2]
    res = bl.solve_jensen_trustregions(T, bl.TrustRegionOptions(max_iter=100, grad_tol=1e-9, feas_tol=1e-8,
                                                                history=False))
    assert res.success, (res.status, res.stats)
