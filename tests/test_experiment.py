"""Experiment harness logic on the CPU: call schedule and summary statistics (no GPU)."""
import json

import pytest

from symplectic_stiefel_nearness_amd import experiment


def small_protocol(**kw):
    p = dict(experiment.PROTOCOL)
    p.update(targets_per_case=3, rounds=2, cases=["n256_k8"], round_orders=[["n256_k8"]])
    p.update(kw)
    return p


def test_schedule_counts_and_batch_order():
    p = small_protocol()
    rows = list(experiment.schedule(p, 6, 1))
    warm = [r for r in rows if r[0] == "warmup"]
    meas = [r for r in rows if r[0] == "measured"]
    assert len(warm) == p["single_warmups"] * len(p["single_methods"]) + len(p["batch_methods"])
    assert [r[3] for r in meas if r[1] == "single"] == [pos for pos in range(3) for _ in p["single_methods"]]
    batch1 = [r[2] for r in meas if r[1] == "batch"]
    batch2 = [r[2] for r in experiment.schedule(p, 6, 2) if r[0] == "measured" and r[1] == "batch"]
    assert sorted(batch1) == sorted(p["batch_methods"]) and batch2 == batch1[::-1]
    # the reference's rotation of single methods (experiment.py:168-170) with three methods
    p3 = small_protocol(single_methods=["a", "b", "c"])
    first = [r[2] for r in experiment.schedule(p3, 0, 1) if r[0] == "measured" and r[1] == "single"]
    assert first == ["a", "b", "c", "b", "c", "a", "c", "a", "b"]


def _write_rows(tmp_path, p, ms):
    for r in range(1, p["rounds"] + 1):
        folder = tmp_path / f"round{r}" / "n256_k8"
        folder.mkdir(parents=True)
        with (folder / "raw.jsonl").open("w") as f:
            for phase, exp, method, pos, wi in experiment.schedule(p, 6, r):
                val = [dict(index=0, feasibility=1e-12, stationarity=2e-14, delta=3e-12)]
                row = dict(case="n256_k8", round=r, phase=phase, experiment=exp, method=method, item=pos,
                           warmup_index=wi, milliseconds=ms[(exp, method)] * r, exception=None,
                           validation=val, iterations=[3], success=True)
                f.write(json.dumps(row) + "\n")


def test_summary_statistics(tmp_path):
    p = small_protocol()
    ms = {("single", "dual"): 1.0, ("single", "gao_nonmonotone"): 30.0, ("single", "jensen_rtr2"): 10.0,
          ("batch", "dual"): 2.0, ("batch", "dual_sequential"): 8.0, ("batch", "dual_device"): 1.5}
    _write_rows(tmp_path, p, ms)
    s = experiment.summarize(tmp_path, p)
    c = s["cases"][0]
    assert c["single"]["dual"]["calls"] == 6 and c["single"]["dual"]["round_mean_ms"] == [1.0, 2.0]
    assert c["batch"]["dual"]["mean_ms"] == pytest.approx(3.0)
    assert c["batch"]["speedup_batched_vs_sequential"] == pytest.approx(4.0)
    assert c["single"]["gao_nonmonotone"]["speedup_dual_vs_this"] == pytest.approx(30.0)
    assert c["single"]["jensen_rtr2"]["speedup_dual_vs_this"] == pytest.approx(10.0)
    assert s["all_calls_successful"] and s["accuracy"]["max_delta"] == 3e-12
    assert (tmp_path / "summary.md").read_text().count("| n256_k8 |") == 1
    json.loads((tmp_path / "summary.json").read_text())


def test_summary_rejects_missing_calls(tmp_path):
    p = small_protocol()
    ms = {("single", "dual"): 1.0, ("single", "gao_nonmonotone"): 30.0, ("single", "jensen_rtr2"): 10.0,
          ("batch", "dual"): 2.0, ("batch", "dual_sequential"): 8.0, ("batch", "dual_device"): 1.5}
    _write_rows(tmp_path, p, ms)
    path = tmp_path / "round2" / "n256_k8" / "raw.jsonl"
    lines = path.read_text().splitlines()
    path.write_text("\n".join(lines[:-1]) + "\n")
    with pytest.raises(ValueError):
        experiment.summarize(tmp_path, p)

