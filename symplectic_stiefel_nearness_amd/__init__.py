"""MI355X-native Dual solver for the nearest real symplectic Stiefel matrix.

Drop-in for the reference package `symplectic_projection` (src/symplectic_projection/__init__.py:1-19):

    import symplectic_stiefel_nearness_amd as symplectic_projection
    symplectic_projection.initialize_kernels()
    X, omega, iterations = symplectic_projection.solve(T)   # T: float64 [batch, 2n, 2k]

The arithmetic runs in libssn.so (hand-written HIP for gfx950) through a C ABI
(include/ssn.h); this package is the host-side mirror of the reference interface. The
comparison solvers of the reference experiment (Riemannian steepest descent and trust region,
`baselines`) are PyTorch code, off the hot path.
"""
from .dual import (initialize_kernels, solve, solve_device, solve_batches, project, backward_device, gram_chunks,
                   dual_batch, dual_batch_staged, recover_matmul, residuals, FEAS_TOL, MAX_NEWTON, MAX_CG,
                   MAX_BACKTRACK, REFINE_THRESHOLDS)

from .uneven import UnevenSolver, solve_uneven
from .autotune import autotune
from .baselines import (SolverOptions, TrustRegionOptions, SolveResult, solve_gao_nonmonotone,
                        solve_jensen_trustregions)

__all__ = ["initialize_kernels", "solve", "solve_device", "solve_batches", "project", "backward_device", "gram_chunks",
           "dual_batch", "dual_batch_staged", "recover_matmul", "residuals", "UnevenSolver", "solve_uneven", "autotune",
           "SolverOptions", "TrustRegionOptions", "SolveResult", "solve_gao_nonmonotone", "solve_jensen_trustregions"]
