"""Autotuner of the uneven batch's tile shapes.

For every distinct (n, m) it times the Gram and recovery stages (the stages the tiles change;
ssn_solve_uneven_stages) of a small homogeneous batch of that shape under each candidate (Gram
rows per work item, recovery rows per work item) and keeps the fastest. Results can be kept in
a JSON cache (path argument or SSN_TUNING_CACHE).
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

import torch

from . import _lib
from .uneven import UnevenSolver

GRAM_ROWS_CANDIDATES = (256, 512, 1024, 2048)
REC_ROWS_CANDIDATES = (64, 128, 256)


def _key(n, m):
    return f"{n},{m}"


def load_cache(path=None):
    path = path or os.environ.get("SSN_TUNING_CACHE")
    if not path or not Path(path).exists():
        return {}
    raw = json.loads(Path(path).read_text())
    return {tuple(int(x) for x in k.split(",")): tuple(v) for k, v in raw.items()}


def save_cache(table, path=None):
    path = path or os.environ.get("SSN_TUNING_CACHE")
    if not path:
        return
    Path(path).write_text(json.dumps({_key(n, m): list(v) for (n, m), v in sorted(table.items())}, indent=1))


@torch.no_grad()
def autotune(shapes, batch=4, reps=3, gram_rows=GRAM_ROWS_CANDIDATES, rec_rows=REC_ROWS_CANDIDATES,
             cache=None, seed=0):
    """Table {(n, m): (gram_chunk_rows, rec_rows)} for the given (n, k) shapes (m = 2k)."""
    table = load_cache(cache)
    dev = torch.device("cuda", torch.cuda.current_device())
    g = torch.Generator(device="cpu").manual_seed(seed)
    for n, k in sorted(set((int(n), int(k)) for n, k in shapes)):
        m = 2 * k
        if (n, m) in table:
            continue
        T = [(torch.randn(2 * n, m, dtype=torch.float64, generator=g) / math.sqrt(2 * n)).to(dev)
             for _ in range(batch)]
        best = None
        for cr in gram_rows:
            if cr > max(n, 256) * 2:
                continue
            for rr in rec_rows:
                tuning = {(n, m): (cr, rr)}
                solver = UnevenSolver([(n, k)] * batch, device=dev, tuning=tuning, inputs=T, groups=1)
                solver.run(check=False)              # leaves P, Q for the recovery stage
                stages = _lib.SSN_STAGE_GRAM | _lib.SSN_STAGE_RECOVER
                solver.run(check=False, stages=stages)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    solver.run(check=False, stages=stages)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / reps
                if best is None or ms < best[0]:
                    best = (ms, cr, rr)
        table[(n, m)] = (best[1], best[2])
    save_cache(table, cache)
    return table
