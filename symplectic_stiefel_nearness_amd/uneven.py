"""Uneven batches: targets of mixed (n, k) solved by one launch per stage, optionally captured
in a HIP graph and replayed.

The C ABI plans a batch (workspace regions and work lists, largest work first), the plan
is copied to the start of the device workspace, and ssn_solve_uneven launches the Gram
(one persistent launch per order and one reduction), the eigensolver, the Newton iteration
(one workgroup per target) and the recovery (one persistent launch). The targets of the
largest order and the others are two plans run at once on two streams (their graph captures
both). Tile shapes per target (Gram rows per work item, recovery rows per work item) come from
`tuning`, e.g. the autotuner's table.
No reference counterpart: the reference solves one (n, k) per call
(dual_recommended_hvp.py:231-242); each target's result is the one `solve` gives for it.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .dual import _raise_for_status, _require_ready, _stream, _workspace

DEFAULT_TILE = (512, 64)     # (Gram rows per work item, recovery rows per work item)


def _tile(tuning, n, m):
    if not tuning:
        return DEFAULT_TILE
    return tuning.get((n, m), DEFAULT_TILE)


class UnevenSolver:
    """Static buffers, plans and (optionally) a captured graph for one list of target shapes.

    shapes: list of (n, k); targets are [2n, 2k] float64. run(targets) copies the targets into
    the solver's input buffers and returns [(X, omega, iterations)], views of the solver's
    output buffers (overwritten by the next run).
    inputs: optional device tensors used AS the input buffers (no copy, so a planned or captured
    solve reads them in place). They are aliased, not owned: run(targets) with new targets
    copies those into these tensors, overwriting the caller's data; run() without targets
    solves whatever they hold at that moment.
    groups: with 2 (the default) the targets of the largest padded order present form one plan,
    launched on the current stream, and the other targets a second plan, launched at the same
    time on a side stream: the per-target eigensolver and Newton launches of the largest order
    set the critical path, and every other stage overlaps them. 1: one plan for all targets.
    Each target's results are those of the single-plan solve bit for bit (per-target work does
    not depend on the other targets of its plan).
    """

    def __init__(self, shapes, device=None, tuning=None, capture=False, inputs=None, groups=2):
        self.lib = _require_ready()
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.shapes = [(int(n), int(k)) for n, k in shapes]
        count = len(self.shapes)
        assert count > 0 and groups in (1, 2)
        f64 = dict(dtype=torch.float64, device=dev)
        if inputs is None:
            self.T = [torch.zeros(2 * n, 2 * k, **f64) for n, k in self.shapes]
        else:
            self.T = list(inputs)
        self.X = [torch.empty(2 * n, 2 * k, **f64) for n, k in self.shapes]
        self.omega = [torch.empty(2 * k, 2 * k, **f64) for n, k in self.shapes]
        self.iters = torch.zeros(count, dtype=torch.int64, device=dev)
        self.status = torch.zeros(count, dtype=torch.int32, device=dev)
        self.counters = torch.zeros((count, 12), dtype=torch.int32, device=dev)
        for i, (n, k) in enumerate(self.shapes):
            t = self.T[i]
            assert t.shape == (2 * n, 2 * k) and t.dtype == torch.float64 and t.device == dev
            assert t.stride(1) == 1, "targets must have unit column stride"
        padded = [self.lib.ssn_padded_order(2 * k) for _, k in self.shapes]
        top = max(padded)
        big = [i for i in range(count) if padded[i] == top]
        rest = [i for i in range(count) if padded[i] != top]
        self.groups = [big, rest] if groups == 2 and rest else [list(range(count))]
        self._plans = [self._plan(idx, tuning) for idx in self.groups]
        self._side = torch.cuda.Stream(dev) if len(self.groups) > 1 else None
        self.graph = None
        if capture:
            self._capture()

    def _plan(self, idx, tuning):
        """Plan and workspace of the targets idx (their device plan copied into the workspace)."""
        arr = (_lib.SsnTarget * len(idx))()
        for j, i in enumerate(idx):
            n, k = self.shapes[i]
            t = self.T[i]
            cr, rr = _tile(tuning, n, 2 * k)
            arr[j] = _lib.SsnTarget(t.data_ptr(), t.stride(0), self.X[i].data_ptr(), self.X[i].stride(0),
                                    self.omega[i].data_ptr(), self.omega[i].stride(0),
                                    self.iters.data_ptr() + 8 * i, self.status.data_ptr() + 4 * i,
                                    self.counters.data_ptr() + 48 * i, n, 2 * k, cr, rr)
        nbytes = self.lib.ssn_uneven_workspace_bytes(arr, len(idx))
        if nbytes == 0:
            raise ValueError("unsupported target shape in the uneven batch (k must be 1..64)")
        ws = _workspace(nbytes, self.device)
        pbytes = self.lib.ssn_uneven_plan_bytes(arr, len(idx))
        plan = (ctypes.c_uint8 * pbytes)()
        with torch.cuda.device(self.device):
            rc = self.lib.ssn_uneven_plan(arr, len(idx), ws.data_ptr(), ws.numel(), ctypes.addressof(plan), pbytes)
        _lib.check(rc, "ssn_uneven_plan")
        host = torch.frombuffer(bytearray(bytes(plan)), dtype=torch.uint8)
        ws[:pbytes].copy_(host)
        return arr, plan, ws

    def _launch_plan(self, k, stream, stages):
        _, plan, ws = self._plans[k]
        rc = self.lib.ssn_solve_uneven_stages(stream.cuda_stream, ctypes.addressof(plan), ws.data_ptr(),
                                              ws.numel(), stages)
        _lib.check(rc, "ssn_solve_uneven_stages")

    def _launch(self, stages=_lib.SSN_STAGE_ALL):
        cur = torch.cuda.current_stream(self.device)
        if self._side is None:
            self._launch_plan(0, cur, stages)
            return
        self._side.wait_stream(cur)                  # inputs ready for the side plan
        self._launch_plan(0, cur, stages)            # the largest order: the critical path
        self._launch_plan(1, self._side, stages)
        cur.wait_stream(self._side)

    def _capture(self):
        with torch.cuda.device(self.device):
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._launch()                      # warm-up outside the graph
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._launch()
            torch.cuda.synchronize()

    def run(self, targets=None, check=True, stages=_lib.SSN_STAGE_ALL):
        """Solve (or, with `stages`, run only the selected stages, eagerly); see the class doc."""
        if targets is not None:
            assert len(targets) == len(self.T)
            for dst, src in zip(self.T, targets):
                dst.copy_(src)
        with torch.cuda.device(self.device):
            if self.graph is not None and stages == _lib.SSN_STAGE_ALL:
                self.graph.replay()
            else:
                self._launch(stages)
        if check and stages == _lib.SSN_STAGE_ALL:
            _raise_for_status(self.status)
        return [(self.X[i], self.omega[i], self.iters[i:i + 1]) for i in range(len(self.T))]


@torch.no_grad()
def solve_uneven(targets, tuning=None, return_stats=False):
    """Solve a list of float64 targets [2n_i, 2k_i] (device tensors) of mixed shapes.

    Returns [(X_i, omega_i, iterations_i)]; with return_stats also dict(status, counters).
    Raises as `solve` does (RuntimeError for a line-search exhaustion or the Newton cap).
    """
    targets = [t if t.stride(-1) == 1 else t.contiguous() for t in targets]
    assert targets and all(t.is_cuda and t.dtype == torch.float64 and t.ndim == 2 for t in targets)
    shapes = [(t.shape[0] // 2, t.shape[1] // 2) for t in targets]
    solver = UnevenSolver(shapes, device=targets[0].device, tuning=tuning, inputs=targets)
    out = solver.run(check=False)
    _raise_for_status(solver.status)
    out = [(X.clone(), om.clone(), it.clone()) for X, om, it in out]
    if return_stats:
        return out, dict(status=solver.status.clone(), counters=solver.counters.clone())
    return out
