"""The experiment protocol end to end on one GPU, at a small size."""
import json

import pytest

from symplectic_stiefel_nearness_amd import experiment

pytestmark = pytest.mark.gpu


def test_protocol_small_run(ssn, tmp_path):
    p = dict(experiment.PROTOCOL)
    p.update(cases=["n256_k8", "n2048_k16"], round_orders=[["n256_k8", "n2048_k16"], ["n2048_k16", "n256_k8"]],
             targets_per_case=4, rounds=2, single_warmups=1)
    print("experiment", p["cases"], flush=True)
    s = experiment.run(tmp_path / "out", p, log=lambda line: None)
    assert s["all_calls_successful"], s["failure_records"]
    assert s["counts"]["formal_calls"] == 2 * 2 * (4 * len(p["single_methods"]) + len(p["batch_methods"]))
    for c in s["cases"]:
        assert c["batch"]["speedup_batched_vs_sequential"] > 0
        for method in ("gao_nonmonotone", "jensen_rtr2"):      # the comparison solvers, single target
            assert c["single"][method]["all_calls_successful"] and c["single"][method]["speedup_dual_vs_this"] > 0
    env = json.loads((tmp_path / "out" / "environment.json").read_text())
    assert env["compute_units"] > 0 and env["library"]
    raw = (tmp_path / "out" / "round1" / "n256_k8" / "raw.jsonl").read_text().splitlines()
    rows = [json.loads(x) for x in raw]
    assert all(len(r["validation"]) == (4 if r["item"] is None else 1) for r in rows)
    assert s["accuracy"]["max_feasibility"] <= 1e-8
