"""The eigensolver's one-launch form (rotation workgroups and eigenvector-replay workgroups in one
cooperative launch, the replay following the rotation record as it is written) against its
two-launch form (rotations, then the replay), through ssn_solve_opts.

Bar: bitwise equal X, omega and counters (both forms apply the same rotations with the same
expressions); the replay's bounded wait, forced to give up from the host (a zero spin limit),
marks every target of the launch ST_EIGH and the solve raises, with every kernel ending normally.
"""
import os
import sys

import pytest
import torch

from symplectic_stiefel_nearness_amd import _lib, inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def jacobi_eigensolver(monkeypatch):
    """These are tests of the Jacobi forms: SSN_EIGH_TRI=0 turns the tridiagonal eigensolver off
    (it would otherwise take every order up to 96)."""
    monkeypatch.setenv("SSN_EIGH_TRI", "0")


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _both(ssn, T):
    one = ssn.dual._run_solve(T, False, shared=False)
    two = ssn.dual._run_solve(T, False, shared=True)
    return one, two


@pytest.mark.parametrize("n,k,batch", [(4096, 48, 64), (4096, 48, 16), (2048, 16, 64), (700, 40, 20),
                                       (300, 5, 4), (100, 1, 3)])
def test_one_launch_equals_two_launches(ssn, n, k, batch):
    say("eigensolver forms", n, k, batch)
    T = inputs.make_targets(n, k, batch, seeds=list(range(500, 500 + batch))).cuda()
    (X1, o1, i1, _, s1), (X2, o2, i2, _, s2) = _both(ssn, T)
    assert torch.equal(X1, X2) and torch.equal(o1, o2) and torch.equal(i1, i2)
    assert torch.equal(s1["counters"][:, :5], s2["counters"][:, :5])
    assert bool((s1["status"] == 0).all())


def test_replay_wait_gives_up_cleanly(ssn):
    """SSN_REPLAY_SPIN_LIMIT=0: the following replay gives up at its first wait, every target of the
    one-launch eigensolver ends ST_EIGH and solve_device raises; the next solve is normal again."""
    T = inputs.make_case("n4096_k48", 16).cuda()
    os.environ["SSN_REPLAY_SPIN_LIMIT"] = "0"
    try:
        X, om, it, st = ssn.solve_device(T, return_stats=True, check=False)
        say("  statuses with a zero spin limit:", st["status"].tolist())
        assert bool((st["status"] == _lib.ST_EIGH).all())
        with pytest.raises(RuntimeError, match="eigensolver"):
            ssn.solve_device(T)
    finally:
        del os.environ["SSN_REPLAY_SPIN_LIMIT"]
    X, om, it, st = ssn.solve_device(T, return_stats=True)
    assert bool((st["status"] == 0).all())


def test_shared_device_option_is_validated(ssn):
    import ctypes
    L = _lib.load()
    o = _lib.SsnOptions()
    L.ssn_default_options(ctypes.byref(o))
    o.shared_device = 2
    T = inputs.make_case("n4096_k48", 2).cuda()
    X = torch.empty_like(T)
    om = torch.empty(2, 96, 96, dtype=torch.float64, device="cuda")
    it = torch.zeros(2, dtype=torch.int64, device="cuda")
    st = torch.zeros(2, dtype=torch.int32, device="cuda")
    ws = torch.empty(L.ssn_workspace_bytes(2, 4096, 96), dtype=torch.uint8, device="cuda")
    rc = L.ssn_solve_opts(torch.cuda.current_stream().cuda_stream, ctypes.byref(o), T.data_ptr(), T.stride(0),
                          T.stride(1), 2, 4096, 96, X.data_ptr(), X.stride(0), X.stride(1), om.data_ptr(),
                          om.stride(0), om.stride(1), it.data_ptr(), st.data_ptr(), None, ws.data_ptr(), ws.numel())
    assert rc == _lib.SSN_ERR_ARGS
