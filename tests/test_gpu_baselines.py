"""The comparison solvers on the GPU (FP64 torch on the device): both reach the HIP Dual
solution of the same seeded targets (X within 1e-7 relative, omega within 1e-6) and pass the
acceptance check; and the experiment harness's single-target call of each returns accepted
outputs."""
import sys

import pytest
import torch

from symplectic_stiefel_nearness_amd import acceptance as dacc, baselines as bl, experiment, inputs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", ["n256_k8", "n2048_k16"])
def test_baselines_reach_the_dual_solution(ssn, case):
    T = inputs.make_case(case, 1).cuda()
    X, omega, _ = ssn.solve_device(T)
    opts = experiment.PROTOCOL["gao_options"]
    for name, res in (("gao_nonmonotone", bl.solve_gao_nonmonotone(T[0], bl.SolverOptions(**opts))),
                      ("jensen_rtr2", bl.solve_jensen_trustregions(T[0], bl.TrustRegionOptions(**opts)))):
        ex = float(torch.linalg.vector_norm(res.x - X[0]) / torch.linalg.vector_norm(X[0]))
        eo = float(torch.linalg.vector_norm(res.omega - omega[0]) / torch.linalg.vector_norm(omega[0]))
        say(f"  {case} {name}: {res.status} after {res.iterations} iterations, "
            f"{res.stats['seconds']:.3f} s, relX {ex:.1e} relOmega {eo:.1e}")
        assert res.success and res.x.is_cuda
        assert ex < 1e-7 and eo < 1e-6
        assert bool(dacc.inspect(T, res.x[None])["passed"].all())


def test_experiment_single_calls_of_the_baselines(ssn):
    T = inputs.make_case("n256_k8", 1)
    for method in ("gao_nonmonotone", "jensen_rtr2"):
        out = experiment.timed_call(method, T, T.cuda(), experiment.PROTOCOL)
        say(f"  experiment call {method}: {out['milliseconds']:.1f} ms, success {out['success']}")
        assert out["success"] and out["exception"] is None
