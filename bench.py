"""Benchmark of the Dual solver on one MI355X.

Workload (one "step"): one batched solve of 64 seeded targets with n = 4096, k = 48
(T of shape 8192 x 96, FP64, entries N(0, 1/(2n)), seeds as the reference experiment),
inputs already resident in HBM, all arithmetic FP64. Beside it: the same step with the
Gram and recovery as split BF16 products, the per-target loop, one C call, the
CPU-in/CPU-out boundary, forward+backward through the autograd projection (FP64 and
split forward) and the CPU oracle. Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "solves/s at n=4096,k=48,batch 64 (r_F≤1e-8) batched vs plain; fwd+bwd ms"
N, K, BATCH, CASE = 4096, 48, 64, "n4096_k48"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def dual_flops(m, counters):
    """Algorithmic FP64 FLOPs of the dual phase from per-target counters (newton, cg, bt, hvp)."""
    m3 = float(m) ** 3
    c = counters.double()
    return float((8 * m3 * c[:, 3] + 14 * m3 * c[:, 0] + 13 * m3 * c[:, 2] + 22 * m3).sum())


def cpu_baseline(budget_s=12.0):
    """The oracle (CPU FP64 restatement of the reference Dual) on 1 numeric thread."""
    from threadpoolctl import threadpool_limits
    import torch
    from symplectic_stiefel_nearness_amd import inputs
    from oracle import dual as odual
    torch.set_num_threads(1)
    sample = inputs.make_case(CASE, 2).numpy()
    solved = 0
    with threadpool_limits(limits=1):
        odual.solve(sample[:1])            # warm-up
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            odual.solve(sample)
            solved += sample.shape[0]
        dt = time.perf_counter() - t0
    return {"value": solved / dt, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle Dual (NumPy FP64, 1 BLAS thread) on the first 2 seeded {CASE} targets, "
                      f"{solved} solves in {dt:.1f} s, CPU in -> CPU out, no PCIe"}


PROFILE = ROOT / "profiles" / "solve_fp64_n4096_k48_b64_kernel_stats.md"


def profiled_traffic(kernel, grid):
    """Copy-calibrated HBM bytes per launch of `kernel` at launch shape `grid` ("x x y" threads, as
    the kernel trace records it), from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (None if it is not there)."""
    try:
        section = None
        for line in PROFILE.read_text(encoding="utf-8").splitlines():
            if line.startswith("## "):
                section = line
                continue
            if section is None or not section.startswith("## HBM traffic"):
                continue
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 5 and cells[0] == f"`{kernel}`" and cells[1] == grid:
                return float(cells[4]) * 1e6
    except (OSError, ValueError):
        pass
    return None


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_worker(budget_s):
    """One single-thread oracle worker (a spawned process): (solves, seconds)."""
    from threadpoolctl import threadpool_limits
    import torch
    from symplectic_stiefel_nearness_amd import inputs
    from oracle import dual as odual
    torch.set_num_threads(1)
    sample = inputs.make_case(CASE, 2).numpy()
    solved = 0
    with threadpool_limits(limits=1):
        odual.solve(sample[:1])            # warm-up
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            odual.solve(sample)
            solved += sample.shape[0]
        dt = time.perf_counter() - t0
    return solved, dt


def cpu_baseline_all_cores(budget_s=10.0):
    """The oracle on every core of this process's CPU share (at most 16): one spawned
    single-thread worker per core, each solving the same 2-target sample for budget_s."""
    import multiprocessing as mp
    cores = min(len(os.sched_getaffinity(0)), 16)
    with mp.get_context("spawn").Pool(cores) as pool:
        res = pool.map(_cpu_worker, [budget_s] * cores)
    rate = sum(c / dt for c, dt in res)
    return {"value": rate, "unit": "solves/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"oracle Dual (NumPy FP64), {cores} processes x 1 BLAS thread, each on the first 2 "
                      f"seeded {CASE} targets for {budget_s:.0f} s ({sum(c for c, _ in res)} solves); "
                      "value = sum of the per-process rates, CPU in -> CPU out"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-reps", type=int, default=1)
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-m128", action="store_true",
                    help="skip the 16 x (16384 x 128) configuration (order 128: FP64, split, forward+backward)")
    ap.add_argument("--no-in-flight", action="store_true",
                    help="skip the batches-in-flight measurement (its launches overlap in a kernel trace)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with one 256 MiB device-to-device copy (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the solve shards by "
              "independent targets and needs no collectives, but this build runs and reports a single "
              "GPU", file=sys.stderr)
        sys.exit(2)

    import torch
    import symplectic_stiefel_nearness_amd as ssn
    from symplectic_stiefel_nearness_amd import acceptance, inputs

    ssn.initialize_kernels()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    T_cpu = inputs.make_case(CASE, BATCH)
    T = T_cpu.to(dev)
    m = 2 * K
    log(f"inputs ready: {BATCH} x {tuple(T.shape[1:])}")

    stream = torch.cuda.current_stream()

    def step(split, ev=None):
        if ev:
            ev[0].record(stream)
        Cr, Ci = ssn.gram_chunks(T, split=split)
        omega, P, Q, iters, st = ssn.dual_batch_staged(Cr, Ci, events=ev[1:4] if ev else None, check=False)
        X = ssn.recover_matmul(T, P, Q, split=split)
        if ev:
            ev[4].record(stream)
        return X, omega, iters, st

    def timed(split):
        """K timed steps bracketed by device synchronisation; per-stage HIP event averages."""
        for _ in range(args.warmup):
            step(split)
        torch.cuda.synchronize()
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s_ in range(args.steps):
            out = step(split, ev[s_])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = {name: sum(e[i].elapsed_time(e[i + 1]) for e in ev) / args.steps
              for i, name in enumerate(("gram", "eigh", "newton", "recover"))}
        return out, wall * 1e3 / args.steps, st

    # headline: the batched FP64 solve; beside it the batched split-precision solve
    (X, omega, iters, st), ms_per_step, stages = timed(False)
    log(f"batched fp64:  {ms_per_step:.3f} ms/step " + ", ".join(f"{k} {v:.3f}" for k, v in stages.items()))
    (Xs, _, _, sts), ms_split, stages_split = timed(True)
    log(f"batched split: {ms_split:.3f} ms/step " + ", ".join(f"{k} {v:.3f}" for k, v in stages_split.items()))
    gram_ms, eigh_ms, newton_ms, rec_ms = (stages[k] for k in ("gram", "eigh", "newton", "recover"))

    status = st["status"].cpu()
    counters = st["counters"].cpu()
    ok_status = bool((status == 0).all())
    met = acceptance.inspect(T, X)
    acc = acceptance.summary(met)
    acc["status_ok"] = ok_status
    acc_split = acceptance.summary(acceptance.inspect(T, Xs))
    acc_split["status_ok"] = bool((sts["status"] == 0).all().item())
    split_vs_fp64 = float(((Xs - X).norm() / X.norm()).item())
    log("acceptance fp64", acc, "split", acc_split, f"split vs fp64 X: {split_vs_fp64:.2e}")
    del Xs

    # split with the FP64 residual-and-refinement step: every target's FP64 residuals (r_F and
    # the KKT residual, from the library's Gram and recovery kernels) on the device, and an FP64
    # re-solve of any target outside the thresholds
    for _ in range(args.warmup):
        ssn.solve_device(T, split=True, refine=True, check=False)
    torch.cuda.synchronize()
    t_r = time.perf_counter()
    for _ in range(args.steps):
        out_r = ssn.solve_device(T, split=True, refine=True, check=False, return_stats=True)
    torch.cuda.synchronize()
    ms_refine = (time.perf_counter() - t_r) * 1e3 / args.steps
    res_r = out_r[3]["residuals"]
    split_refine = {"ms_per_step": ms_refine, "solves_per_s": BATCH / (ms_refine * 1e-3),
                    "refined_targets": int(out_r[3]["refined"].numel()),
                    "max_rF": float(res_r["feasibility"].max()),
                    "max_kkt_residual": float(res_r["stationarity"].max()),
                    "acceptance": acceptance.summary(acceptance.inspect(T, out_r[0]))}
    log(f"split + FP64 residual-and-refinement: {ms_refine:.3f} ms/step, refined {split_refine['refined_targets']}")
    del out_r

    # the same solve through the single C call
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(3):
        X_call = ssn.solve_device(T, check=False)[0]
    torch.cuda.synchronize()
    fused_ms = (time.perf_counter() - t1) * 1e3 / 3

    # Batches in flight (ssn.solve_batches): consecutive 64-target solves go round-robin over the
    # current stream and nb - 1 side streams, so the per-target kernels of one batch (one workgroup
    # per target, 64 of the 256 CUs) run beside those of the next ones. Not `value`: this is the
    # rate of a caller that hands over several batches at once.
    flight = {}
    fl_count = 12
    for nb in (() if args.no_in_flight else (2, 4)):
        ssn.solve_batches([T] * (2 * nb), in_flight=nb)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        fl_out = ssn.solve_batches([T] * fl_count, in_flight=nb)
        torch.cuda.synchronize()
        fms = (time.perf_counter() - t1) * 1e3 / fl_count
        fok = all(torch.equal(o[0], X_call) for o in fl_out)
        del fl_out
        flight[str(nb)] = {"ms_per_batch": fms, "solves_per_s": BATCH / (fms * 1e-3), "results_equal": fok}
        log(f"{nb} batches in flight: {fms:.3f} ms per 64-target batch, results equal {fok}")
    del X_call

    # plain: per-target loop of the FP64 solver (experiment.py:75-80), device-resident
    plain_ms = None
    if args.plain_reps > 0:
        runs = []
        for rep in range(args.plain_reps + 1):
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            for j in range(BATCH):
                ssn.solve_device(T[j:j + 1], check=False)
            torch.cuda.synchronize()
            if rep:
                runs.append((time.perf_counter() - t2) * 1e3)
        plain_ms = sum(runs) / len(runs)
        log(f"plain per-target loop: {plain_ms:.2f} ms for {BATCH} targets")

    # forward + backward through the autograd projection, upstream gradient G ~ N(0, 1)
    G = inputs.make_gradients(N, K, BATCH).to(dev)
    Tg = T.clone().requires_grad_(True)

    def fwd_bwd(split):
        Tg.grad = None
        X_, _, _ = ssn.project(Tg, split=split)
        torch.autograd.backward(X_, G)
        return Tg.grad

    fb = {}
    for split in (True, False):
        for _ in range(2):
            fwd_bwd(split)
        torch.cuda.synchronize()
        fb_reps = 5
        t4 = time.perf_counter()
        for i in range(fb_reps):
            fwd_bwd(split)
        torch.cuda.synchronize()
        fb[split] = (time.perf_counter() - t4) * 1e3 / fb_reps
    fb_ms = fb[False]
    # backward alone (same forward state), HIP events on the launch stream
    X_, om_, it_, ws_, _ = ssn.dual._solve_device_ws(T)
    eb = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ssn.backward_device(T, G, ws_)
    torch.cuda.synchronize()
    eb[0].record(stream)
    for _ in range(fb_reps):
        dT, bst = ssn.backward_device(T, G, ws_, return_stats=True)
    eb[1].record(stream)
    torch.cuda.synchronize()
    bwd_ms = eb[0].elapsed_time(eb[1]) / fb_reps
    adj_it = bst["counters"][:, 0].double().mean().item()
    adj_ok = bool((bst["counters"][:, 1] == 0).all().item())
    del X_, om_, it_, ws_, dT, Tg
    log(f"fwd+bwd: fp64 {fb[False]:.3f} ms, split forward {fb[True]:.3f} ms (backward alone {bwd_ms:.3f} ms, "
        f"adjoint CG {adj_it:.1f} it)")

    # uneven batch: 128 seeded targets of mixed (n, k) with autotuned Gram / recovery tiles:
    # the two-plan solve (largest order and the rest on two streams) eager and as graph
    # replays, beside the single-plan eager solve with the default tiles
    uneven = None
    if not args.no_uneven:
        Tu = [t.to(dev) for t in inputs.make_uneven(128)]
        shapes_u = [(t.shape[0] // 2, t.shape[1] // 2) for t in Tu]

        def timed_runs(solver, reps=3):
            solver.run(check=False)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                solver.run(check=False)
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / reps

        single_ms = timed_runs(ssn.UnevenSolver(shapes_u, inputs=Tu, groups=1))
        t_tune = time.perf_counter()
        table = ssn.autotune(shapes_u)
        tune_s = time.perf_counter() - t_tune
        eager_ms = timed_runs(ssn.UnevenSolver(shapes_u, inputs=Tu, tuning=table))
        solver_g = ssn.UnevenSolver(shapes_u, inputs=Tu, tuning=table, capture=True)
        graph_ms = timed_runs(solver_g)
        ok = bool((solver_g.status == 0).all().item())
        uneven = {"targets": 128, "tiles": "autotuned", "plans": len(solver_g.groups),
                  "eager_ms": eager_ms, "graph_ms": graph_ms, "single_plan_default_tiles_eager_ms": single_ms,
                  "solves_per_s": 128 / (graph_ms * 1e-3), "autotune_s": tune_s, "all_status_ok": ok}
        log(f"uneven 128: single plan {single_ms:.2f} ms; two plans, autotuned: eager {eager_ms:.2f} ms, "
            f"graph {graph_ms:.2f} ms (autotune {tune_s:.1f} s)")
        del Tu, solver_g

    # the second split-precision configuration: 16 targets of 16384 x 128 (n = 8192, k = 64, order
    # 128), single C calls device in / device out, FP64 and split, and forward+backward with an
    # N(0, 1) upstream gradient
    m128 = None
    if not args.no_m128:
        T8 = inputs.make_case("n8192_k64", 16).to(dev)

        def calls_ms(fn, reps=5):
            for _ in range(2):
                fn()
            torch.cuda.synchronize()
            t8 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t8) * 1e3 / reps

        ms8 = calls_ms(lambda: ssn.solve_device(T8, check=False))
        ms8s = calls_ms(lambda: ssn.solve_device(T8, split=True, check=False))
        X8, _, _, st8 = ssn.solve_device(T8, check=False, return_stats=True)
        acc8 = acceptance.summary(acceptance.inspect(T8, X8))
        acc8["status_ok"] = bool((st8["status"] == 0).all().item())
        c8 = st8["counters"].double().mean(0).cpu()
        G8 = inputs.make_gradients(8192, 64, 16).to(dev)
        T8g = T8.clone().requires_grad_(True)

        def fwd_bwd8(split):
            T8g.grad = None
            X8_, _, _ = ssn.project(T8g, split=split)
            torch.autograd.backward(X8_, G8)

        fb8 = calls_ms(lambda: fwd_bwd8(False))
        fb8s = calls_ms(lambda: fwd_bwd8(True))
        m128 = {"targets": 16, "n": 8192, "k": 64, "ms_per_step": ms8, "solves_per_s": 16 / (ms8 * 1e-3),
                "split_ms_per_step": ms8s, "split_solves_per_s": 16 / (ms8s * 1e-3),
                "fwd_bwd_ms": fb8, "fwd_bwd_ms_split_forward": fb8s, "acceptance": acc8,
                "newton_kernel_us_per_target": float(c8[10]) / 100.0,
                "iterations": {"newton_mean": float(c8[0]), "cg_mean": float(c8[1]), "backtracks_mean": float(c8[2])}}
        log(f"16 x 16384 x 128: fp64 {ms8:.3f} ms, split {ms8s:.3f} ms, fwd+bwd {fb8:.3f} / {fb8s:.3f} ms")
        del T8, X8, T8g, G8

    # reference protocol boundary: CPU tensor in -> CPU tensors out (PCIe included), with the
    # chunked transfer pipeline (default for this size) and without it
    def host_solve_ms(pipeline):
        for _ in range(2):                  # pinned buffers and streams come up on first use
            ssn.solve(T_cpu, pipeline=pipeline)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(3):
            ssn.solve(T_cpu, pipeline=pipeline)
        torch.cuda.synchronize()
        return (time.perf_counter() - t3) * 1e3 / 3
    pcie_ms = host_solve_ms(True)
    pcie_plain_ms = host_solve_ms(False)
    log(f"CPU->CPU batched solve: {pcie_ms:.2f} ms pipelined, {pcie_plain_ms:.2f} ms unpipelined")

    fd = dual_flops(m, counters)
    f_eigh = 9.0 * m ** 3 * BATCH               # dense symmetric eigensolver with vectors
    # Gram: Cr = [A; D]^T [A; D] is symmetric (2n m (m + 1) flops for its upper triangle), and
    # Ci = A^T D - (A^T D)^T needs the one product A^T D (2 n m^2)
    f_gram = (2.0 * N * m * (m + 1) + 2.0 * N * m * m) * BATCH
    f_rec = 8.0 * N * m * m * BATCH
    b_gram = 16.0 * N * m * BATCH               # read T once
    b_rec = 32.0 * N * m * BATCH                # read T, write X
    bytes_fwd = 48.0 * N * m * BATCH

    def roof_mfma(kernel, flops, ms):
        achieved = flops / (ms * 1e-3) / 1e12
        return {"kernel": kernel, "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                "algorithmic_flops_per_launch": flops, "avg_ms": ms}

    def roof_hbm(kernel, nbytes, flops, ms):
        achieved = nbytes / (ms * 1e-3) / 1e9
        return {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "algorithmic_bytes_per_launch": nbytes,
                "fp64_equivalent_tflops": flops / (ms * 1e-3) / 1e12, "avg_ms": ms}

    kernels = [roof_mfma("ssn::gram_partial_kernel<6> (+ gram_reduce_tiles_kernel)", f_gram, gram_ms),
               roof_mfma("ssn::eigh_tri_kernel<6> (Householder tridiagonalisation, multisection, twisted "
                         "vectors, back-transformation, Newton-Schulz step)", f_eigh, eigh_ms),
               roof_mfma("ssn::dual_kernel<6> (+ dual_basis_kernel<6, 0|1>, dual_setup_finish_kernel<6>, "
                         "dual_omega_kernel<6>: the Newton stage)", fd, newton_ms),
               roof_mfma("ssn::recover_lds_kernel<6>", f_rec, rec_ms)]
    kernels_split = [roof_hbm("ssn::split_gram_kernel<6,2,5> (+ gram_reduce_tiles_kernel)", b_gram, f_gram,
                              stages_split["gram"]),
                     roof_hbm("ssn::split_recover_kernel<6,5> (+ split_recovery_fragments)", b_rec, f_rec,
                              stages_split["recover"])]
    # HBM traffic per launch of each kernel (the main kernel of a stage) at this workload's grid
    # (the Newton stage's traffic is the sum over its five launches, each at its own grid)
    w = 9216 * BATCH
    newton_launches = [("void ssn::dual_basis_kernel<6, 0>", f"{w * 2 // 3} x 1"),
                       ("void ssn::dual_setup_finish_kernel<6>", f"{w} x 1"),
                       ("void ssn::dual_kernel<6>", f"{512 * BATCH} x 1"),
                       ("void ssn::dual_basis_kernel<6, 1>", f"{w} x 1"),
                       ("void ssn::dual_omega_kernel<6>", f"{w} x 1")]
    for k, launches in zip(kernels, ([("void ssn::gram_partial_kernel<6>", f"4096 x {BATCH}")],
                                     [("void ssn::eigh_tri_kernel<6>", f"{512 * BATCH} x 1")],
                                     newton_launches,
                                     [("void ssn::recover_lds_kernel<6>", f"2048 x {BATCH}")])):
        parts = [profiled_traffic(pname, grid) for pname, grid in launches]
        k["traffic"] = None if None in parts else sum(parts)
        if k["traffic"] is not None:
            names = " + ".join(pname for pname, _ in launches)
            k["traffic_source"] = (f"{PROFILE.relative_to(ROOT)}: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 "
                                   f"(factors from a 256 MiB device copy) of {names}, bytes per launch")
    for k, (pname, grid) in zip(kernels_split, (("void ssn::split_gram_kernel<6, 2, 5>", f"8192 x {BATCH}"),
                                                ("void ssn::split_recover_kernel<6, 5>", f"32768 x {BATCH}"))):
        k["traffic"] = profiled_traffic(pname, grid)
        if k["traffic"] is not None:
            k["traffic_source"] = (f"{PROFILE.relative_to(ROOT)}: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 "
                                   f"(factors from a 256 MiB device copy) of {pname}, bytes per launch")
    dominant = max(kernels, key=lambda k: k["avg_ms"])
    # the dual engine runs one workgroup (one CU) per target: also state the fraction of the FP64
    # MFMA peak of the CUs the launch occupies (the one-launch eigensolver occupies every CU)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    for k in kernels:
        if "dual_kernel" in k["kernel"]:
            k["occupied_cus"] = min(BATCH, n_cu)
            k["frac_of_occupied_cus"] = k["achieved"] / (FP64_PEAK_TFLOPS * min(BATCH, n_cu) / n_cu)

    out = {
        "metric": METRIC,
        "value": BATCH / (ms_per_step * 1e-3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched FP64 Dual, 64 targets, n=4096, k=48 (T 8192x96), device-resident",
                   "batch": BATCH, "n": N, "k": K},
        "split_batched_solves_per_s": BATCH / (ms_split * 1e-3),
        "split_stage_ms": stages_split,
        "split_acceptance": acc_split,
        "split_vs_fp64_rel_X": split_vs_fp64,
        "split_refine": split_refine,
        "fwd_bwd_ms_split_forward": fb[True],
        "plain_solves_per_s": BATCH / (plain_ms * 1e-3) if plain_ms else None,
        "single_call_solves_per_s": BATCH / (fused_ms * 1e-3),
        "batches_in_flight": flight,
        "pcie_inclusive_solves_per_s": BATCH / (pcie_ms * 1e-3),
        "pcie_inclusive_unpipelined_solves_per_s": BATCH / (pcie_plain_ms * 1e-3),
        "fwd_bwd_ms": fb_ms,
        "fwd_bwd_solves_per_s": BATCH / (fb_ms * 1e-3),
        "bwd_ms": bwd_ms,
        "adjoint": {"cg_mean": adj_it, "all_converged": adj_ok},
        "uneven_128": uneven,
        "n8192_k64_b16": m128,
        "stage_ms": stages,
        "dual_phase_ms": eigh_ms + newton_ms,
        "forward_bytes_per_step": bytes_fwd,
        "dual_flops_per_step": fd,
        "iterations": {"newton_mean": float(counters[:, 0].double().mean()),
                       "cg_mean": float(counters[:, 1].double().mean()),
                       "backtracks_mean": float(counters[:, 2].double().mean())},
        "acceptance": acc,
        "roofline": dominant,
        "roofline_kernels": kernels,
        "roofline_kernels_split": kernels_split,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_baseline"]["cpu_model"] = cpu_model()
        log("cpu baseline, all cores ...")
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    if args.calibrate_copy:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 268435456 bytes read, 268435456 bytes written")


if __name__ == "__main__":
    main()
