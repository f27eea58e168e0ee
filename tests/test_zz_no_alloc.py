"""ssn_solve_static, the allocation-free form of ssn_solve, through the C ABI.

Its workspace is the library's static device array (SSN_STATIC_SCRATCH_BYTES per GPU). Two calls
in a row, the second with a smaller input, each give the results of ssn_solve with a caller
workspace bit for bit (the same kernels on the same inputs); an input whose workspace exceeds
the array is rejected with SSN_ERR_WORKSPACE before anything runs.
"""
import ctypes

import pytest
import torch

from symplectic_stiefel_nearness_amd import _lib, inputs

pytestmark = pytest.mark.gpu


def _outputs(T):
    batch, rows, m = T.shape
    dev = T.device
    return (torch.full((batch, rows, m), float("nan"), dtype=torch.float64, device=dev),
            torch.full((batch, m, m), float("nan"), dtype=torch.float64, device=dev),
            torch.full((batch,), -1, dtype=torch.int64, device=dev),
            torch.full((batch,), -1, dtype=torch.int32, device=dev),
            torch.zeros((batch, 12), dtype=torch.int32, device=dev))


def _call(lib, T, static):
    batch, rows, m = T.shape
    X, omega, iters, status, counters = _outputs(T)
    stream = torch.cuda.current_stream().cuda_stream
    args = [stream, T.data_ptr(), T.stride(0), T.stride(1), batch, rows // 2, m,
            X.data_ptr(), X.stride(0), X.stride(1), omega.data_ptr(), omega.stride(0), omega.stride(1),
            iters.data_ptr(), status.data_ptr(), counters.data_ptr()]
    if static:
        rc = lib.ssn_solve_static(*args)
    else:
        ws = torch.empty(lib.ssn_workspace_bytes(batch, rows // 2, m), dtype=torch.uint8, device=T.device)
        rc = lib.ssn_solve(*args, ws.data_ptr(), ws.numel())
    torch.cuda.synchronize()
    return rc, (X, omega, iters, status, counters)


def test_static_solve_matches_workspace_solve(ssn):
    lib = _lib.load()
    assert lib.ssn_init() == _lib.SSN_OK
    # the second call is smaller than the first, so stale contents of the array lie past its data
    for case, count in (("n4096_k48", 8), ("n2048_k16", 3)):
        print("static solve", case, count, flush=True)
        T = inputs.make_case(case, count).cuda()
        rc_s, out_s = _call(lib, T, static=True)
        rc_w, out_w = _call(lib, T, static=False)
        assert rc_s == _lib.SSN_OK and rc_w == _lib.SSN_OK
        X, omega, iters, status, counters = out_s
        Xw, omw, itw, stw, cw = out_w
        assert int(status.abs().max()) == 0
        assert torch.equal(X, Xw) and torch.equal(omega, omw) and torch.equal(iters, itw)
        assert torch.equal(status, stw) and torch.equal(counters[:, :6], cw[:, :6])


def test_static_solve_rejects_inputs_beyond_the_array(ssn):
    lib = _lib.load()
    n, k = 4096, 48
    batch = 1
    while lib.ssn_workspace_bytes(batch, n, 2 * k) <= _lib.SSN_STATIC_SCRATCH_BYTES:
        batch += 1
    print("static solve limit", batch, flush=True)
    T = torch.zeros(batch, 2 * n, 2 * k, dtype=torch.float64, device="cuda")
    rc, (X, omega, iters, status, counters) = _call(lib, T, static=True)
    assert rc == _lib.SSN_ERR_WORKSPACE
    assert int(iters.min()) == -1 and int(status.min()) == -1       # nothing ran
