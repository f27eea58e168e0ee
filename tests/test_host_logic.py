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
