"""Host-side logic of the Python mirror that needs no GPU: status and counter checks."""
import pytest


def test_adjoint_status_check_raises():
    """backward_device's check of the adjoint counters: any target whose Hessian solve stopped
    unconverged (counters[:, 1] != 0) raises; converged targets pass. Host logic only."""
    import torch
    from symplectic_stiefel_nearness_amd import dual
    ok = torch.tensor([[17, 0, 5, 0], [18, 0, 6, 0]], dtype=torch.int32)
    dual._raise_for_adjoint(ok)
    bad = ok.clone()
    bad[1, 1] = 1
    with pytest.raises(RuntimeError, match=r"targets \[1\]"):
        dual._raise_for_adjoint(bad)


def test_bench_reads_traffic_of_the_timed_launch_shapes():
    """bench.py takes each stage kernel's HBM bytes per launch from the committed rocprofv3
    summary, by kernel and launch shape: every stage kernel of the 64-target step (FP64 and split)
    has a row, the value comes from the traffic table (not the trace table of the same shapes),
    and the 16-target launch of the same kernel is a different row."""
    import bench
    shapes = (("void ssn::gram_partial_kernel<6>", "4096 x 64"),
              ("void ssn::eigh_tri_kernel<6>", "32768 x 1"),
              ("void ssn::dual_kernel<6>", "32768 x 1"),
              ("void ssn::dual_basis_kernel<6, 0>", "393216 x 1"),
              ("void ssn::dual_basis_kernel<6, 1>", "589824 x 1"),
              ("void ssn::dual_setup_finish_kernel<6>", "589824 x 1"),
              ("void ssn::dual_omega_kernel<6>", "589824 x 1"),
              ("void ssn::recover_lds_kernel<6>", "2048 x 64"),
              ("void ssn::split_gram_kernel<6, 2, 5>", "8192 x 64"),
              ("void ssn::split_recover_kernel<6, 5>", "32768 x 64"))
    got = {k: bench.profiled_traffic(k, g) for k, g in shapes}
    assert all(v is not None and v > 0 for v in got.values()), got
    # the recovery reads T (403 MB at 64 targets) and writes X (403 MB): at least that much moves
    assert got["void ssn::recover_lds_kernel<6>"] >= 0.95 * 2 * 64 * 8192 * 96 * 8
    assert bench.profiled_traffic("void ssn::recover_lds_kernel<6>", "8192 x 16") < \
        0.5 * got["void ssn::recover_lds_kernel<6>"]
    assert bench.profiled_traffic("void ssn::no_such_kernel", "1 x 1") is None
