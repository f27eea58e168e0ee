"""The package's batched FP64 acceptance checker (used by bench.py on the device) against the
oracle's restatement of the reference acceptance (experiment.py:105-137), on the CPU."""
import numpy as np
import torch

from oracle import acceptance as oacc, dual as odual
from symplectic_stiefel_nearness_amd import acceptance as dacc, inputs


def test_checker_matches_oracle_acceptance():
    T = inputs.make_targets(60, 4, 3, seeds=[1, 2, 3])
    X, _, _ = odual.solve(T.numpy())
    d = dacc.inspect(T, torch.from_numpy(X))
    for j, r in enumerate(oacc.inspect_output(T.numpy(), X)):
        assert abs(float(d["feasibility"][j]) - r["feasibility"]) <= 1e-14 + 1e-9 * r["feasibility"]
        assert abs(float(d["stationarity"][j]) - r["stationarity"]) <= 1e-14 + 1e-9 * r["stationarity"]
        assert abs(float(d["delta"][j]) - r["delta"]) <= 1e-13
        assert bool(d["passed"][j]) == oacc.numeric_pass(r)
    assert bool(d["passed"].all())


def test_checker_rejects_planted_mistakes():
    T = inputs.make_targets(40, 3, 2, seeds=[7, 8])
    X, _, _ = odual.solve(T.numpy())
    assert not dacc.inspect(T, T.clone())["passed"].any()                 # X = T
    bad = torch.from_numpy(X).clone()
    bad[:, :, 0] *= -1                                                     # a wrong sign
    assert not dacc.inspect(T, bad)["passed"].any()
    s = dacc.summary(dacc.inspect(T, torch.from_numpy(X)))
    assert s["all_pass"] and np.isfinite(s["max_gap"])
