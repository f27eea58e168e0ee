"""Uneven batches on the GPU: targets of mixed (n, k) in one launch per stage, against the
CPU FP64 oracle (X, omega within 1e-10 relative, equal Newton / CG / backtrack counts), the
128-target uneven workload with acceptance on every target, a captured graph replayed
against eager launches, and the autotuner."""
import sys

import numpy as np
import pytest
import torch

from oracle import acceptance, dual as odual
from symplectic_stiefel_nearness_amd import inputs

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_vs_oracle(T_list, outs, stats):
    cnt = stats["counters"].cpu().numpy()
    for j, (T, (X, om, it)) in enumerate(zip(T_list, outs)):
        Xo, omo, ito, sto = odual.solve(T.cpu().numpy()[None], return_stats=True)
        ex, eo = rel(X.cpu().numpy(), Xo[0]), rel(om.cpu().numpy(), omo[0])
        say(f"  target {j} {tuple(T.shape)}: newton {int(it)}/{int(ito[0])} relX {ex:.1e} relOmega {eo:.1e}")
        assert ex <= 1e-10 and eo <= 1e-10
        assert int(it) == int(ito[0]) and cnt[j, 1] == sto.cg[0] and cnt[j, 2] == sto.backtracks[0]


def test_uneven_mixed_small(ssn):
    shapes = [(40, 1), (100, 3), (300, 8), (777, 20), (513, 48), (260, 64), (64, 16)]
    T = [inputs.make_targets(n, k, 1, seeds=[300 + j])[0].cuda() for j, (n, k) in enumerate(shapes)]
    outs, stats = ssn.solve_uneven(T, return_stats=True)
    _check_vs_oracle(T, outs, stats)


def test_uneven_tiles_do_not_change_results_beyond_rounding(ssn):
    shapes = [(1000, 8), (700, 48), (1500, 32)]
    T = [inputs.make_targets(n, k, 1, seeds=[400 + j])[0].cuda() for j, (n, k) in enumerate(shapes)]
    tuning = {(1000, 16): (256, 256), (700, 96): (1024, 128), (1500, 64): (2048, 64)}
    outs, stats = ssn.solve_uneven(T, tuning=tuning, return_stats=True)
    _check_vs_oracle(T, outs, stats)


def test_uneven_config3_full(ssn):
    """The 128-target uneven workload with the autotuner's tiles (two plans on two streams):
    acceptance on every target (device FP64 checker, then the CPU oracle checker on every 16th),
    oracle parity on the smallest targets."""
    from symplectic_stiefel_nearness_amd import acceptance as dacc
    shapes = inputs.uneven_shapes(128)
    T = [t.cuda() for t in inputs.make_uneven(128)]
    table = ssn.autotune(shapes)
    say("  autotuned tiles", sorted(table.items())[:6], "...")
    outs, stats = ssn.solve_uneven(T, tuning=table, return_stats=True)
    assert int((stats["status"] != 0).sum()) == 0
    worst = {"feasibility": 0.0, "stationarity": 0.0, "delta": 0.0}
    for t, (X, _, _) in zip(T, outs):
        met = dacc.inspect(t[None], X[None])
        assert bool(met["passed"].all())
        for k in worst:
            worst[k] = max(worst[k], float(met[k].max()))
    say("  uneven 128: worst", worst)
    for j in range(0, 128, 16):
        r = acceptance.inspect_output(T[j].cpu().numpy()[None], outs[j][0].cpu().numpy()[None])[0]
        assert acceptance.numeric_pass(r), (j, r)
    small = sorted(range(128), key=lambda j: shapes[j][0] * shapes[j][1])[:4]
    sub = [T[j] for j in small]
    _check_vs_oracle(sub, [outs[j] for j in small],
                     dict(counters=stats["counters"][small]))


def test_uneven_graph_replay_matches_eager(ssn):
    """A captured graph, replayed several times with eager work and synchronisations between
    replays, gives bitwise the eager launch's results."""
    shapes = [(300, 8), (1024, 48), (200, 64), (2048, 16)]
    T1 = [inputs.make_targets(n, k, 1, seeds=[500 + j])[0].cuda() for j, (n, k) in enumerate(shapes)]
    T2 = [inputs.make_targets(n, k, 1, seeds=[600 + j])[0].cuda() for j, (n, k) in enumerate(shapes)]
    ref1 = ssn.solve_uneven(T1)
    ref2 = ssn.solve_uneven(T2)
    solver = ssn.UnevenSolver(shapes, capture=True)
    for rep in range(3):
        for Tin, ref in ((T1, ref1), (T2, ref2)):
            out = solver.run(Tin)
            torch.cuda.synchronize()
            for (X, om, it), (Xr, omr, itr) in zip(out, ref):
                assert torch.equal(X, Xr) and torch.equal(om, omr) and torch.equal(it, itr)
            _ = (torch.randn(1000, device="cuda") * 2).sum().item()     # eager work between replays


def test_autotune_returns_valid_tiles(ssn, tmp_path):
    table = ssn.autotune([(1024, 8), (2048, 48)], batch=2, reps=1, gram_rows=(256, 1024), rec_rows=(64, 256),
                         cache=str(tmp_path / "tune.json"))
    assert set(table) == {(1024, 16), (2048, 96)}
    for cr, rr in table.values():
        assert cr in (256, 1024) and rr in (64, 256)
    assert (tmp_path / "tune.json").exists()
    T = [inputs.make_targets(1024, 8, 1, seeds=[7])[0].cuda(), inputs.make_targets(2048, 48, 1, seeds=[8])[0].cuda()]
    outs, stats = ssn.solve_uneven(T, tuning=table, return_stats=True)
    _check_vs_oracle(T, outs, stats)


def test_two_plans_equal_one_plan(ssn):
    """The two-plan solve (largest order on the current stream, the rest on a side stream, both
    eager and captured) gives the single-plan solve's results bit for bit."""
    shapes = [(300, 8), (1024, 48), (200, 64), (2048, 16), (512, 64), (100, 3)]
    T = [inputs.make_targets(n, k, 1, seeds=[700 + j])[0].cuda() for j, (n, k) in enumerate(shapes)]
    one = ssn.UnevenSolver(shapes, inputs=T, groups=1)
    ref = [(X.clone(), om.clone(), it.clone()) for X, om, it in one.run()]
    two = ssn.UnevenSolver(shapes, inputs=T)
    assert len(two.groups) == 2 and sorted(two.groups[0]) == [2, 4]
    cap = ssn.UnevenSolver(shapes, inputs=T, capture=True)
    for solver in (two, cap, cap):
        out = solver.run()
        torch.cuda.synchronize()
        for (X, om, it), (Xr, omr, itr) in zip(out, ref):
            assert torch.equal(X, Xr) and torch.equal(om, omr) and torch.equal(it, itr)
