"""CPU FP64 oracle for the Dual solver of the nearest real symplectic Stiefel problem.

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this package: only
`tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`
use it, and only as the checker / host-CPU baseline, never as the thing that
is measured or shipped.

Contents
--------
- `dual`        : NumPy restatement of `src/symplectic_projection/dual_recommended_hvp.py`
                  (Gram, eigenbasis setup, Newton / PCG / Armijo, recovery), batched
                  with the reference's masks, constants and error semantics.
- `acceptance`  : restatement of `experiment.inspect_output` (r_F, r_G, primal-dual gap)
                  and `baselines/geometry.prepare_projection`.
- `adjoint`     : CPU FP64 implicit-differentiation gradient dL/dT of the projection
                  (no reference counterpart: the reference has no backward).

Pinning (what checks each function; see tests/test_oracle.py):
- Gram: complex form C = B^H B (B = A + iD), a different formulation of Cr, Ci.
- HVP: central finite differences of the gradient; self-adjointness; the closed
  form (lambda_i + lambda_j) E_ij at omega = 0 in the eigenbasis.
- line-search change: direct evaluation phi(omega+delta) - phi(omega) through a
  complex Cholesky of I + i*omega (extended precision via mpmath on tiny m).
- whole solve: known answer T = c*S (S orthosymplectic) -> X = S, omega = (c-1) J;
  brute force by SLSQP on tiny (n, k); left/right orthosymplectic invariance;
  acceptance r_F <= 1e-8, r_G <= 1e-9, gap <= 1e-8.
- acceptance metrics: planted mistakes (X = T, X = S with a wrong sign) must fail.
- adjoint: central finite differences of the oracle's own solve; the invariant
  <dL/dT, S> = 0 at T = c*S.
Trajectory quantities (iteration counts) have no reference run to pin them:
the reference Dual cannot run without CuPy/NVRTC and an NVIDIA GPU.
"""
