#!/usr/bin/env python3
"""Benchmark of the MI355X dual simplex hot path.

Headline (the metric: solves/s, learned-warm vs slack-start on the same GPU and the same LPs):
the LP family of the single-LP recipe at m=2,000, n=5,000 with 5 nonzeros per column (base
seed 4), b and c times lognormal(0, 0.1), FP64.  The per-variable network (BF16 MFMA kernels) is
trained on the optimal bases the GPU solver finds for --train instances; a step is one held-out
instance solved on the single-LP GPU path (--threads at a time, one HIP stream each).  --warmup W
held-out instances run untimed, then --steps K instances are solved from the slack basis and from
the learned warm start; `value` is the learned-warm solves/s over those K, with the slack-start
solves/s, iterations and objective agreement beside it.  Each solving thread creates its context
and uploads its LP (about 1 MB) inside the timed wall, which also includes the host's per-chunk
synchronisation.

Also reported: the single large LP (m=20,000, n=40,000, 10 nonzeros per column, refactor every
64) over a fixed window of iterations 501..2,500 with the pivot-row SpMV roofline (`roofline`,
HIP events on the solver's stream), the FP64 MFMA use of the refactorization's trailing update,
the batched small-LP family, the SpMV size sweep, the transportation LP with the Harris ratio test,
and the batched learned warm start at m=256.  CPU baselines: the oracle (a C++ restatement of the
reference's sequential solver), single thread unless stated.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# The SpMV of every 8th iteration of the timed window is bracketed by start/stop events (a uniform
# sample of the window's launches): events on every launch slow the iteration loop by about 5%.
SPMV_TIME_STRIDE = 8
METRIC = "solves/s + ms/iter, learned-warm vs slack-start on same GPU, obj rel err≤1e-9"

# HBM bytes per k_spmv launch, from two rocprofv3 counter runs (FETCH_SIZE, then WRITE_SIZE), each
# scaled by what it reads for k_copy moving 256 MiB (FETCH_SIZE reads 131,087.8 KiB of the
# 262,144 KiB read: x 1.99976; WRITE_SIZE reads the copy's writes exactly: x 1.0).  Counters
# cannot be read inside a timed run, so other workloads report null.
#   single LP m=20,000 n=40,000, 10 nnz/col, seed 1, refactor every 64 (iterations 1..300):
#     3,633.3 KiB x 1.99976 + 316.7 KiB
#   transportation LP 500 x 1000, seed 2 (all 2,294 iterations): 9,916.4 KiB x 1.99976 + 3,849.1 KiB
SPMV_TRAFFIC_BYTES = {(20000, 40000, 10, 1, 64): 7764450}
TRANSPORT_SPMV_TRAFFIC_BYTES = {(500, 1000, 2): 24247818}
#   4,096 batched packing LPs (seeds 3000+k), the timed k_batch<512> launch of 479 ms:
#     584,897,945 KiB x 1.99976 + 906,623,171 KiB
BATCH_TRAFFIC_BYTES = {4096: 2126109373945}


def log(msg):
    """Progress on stderr (the JSON result is the only stdout line)."""
    print("[bench %.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(lp, refactor_freq, warmup, steps):
    """Oracle ms/iter over iterations (warmup, warmup+steps] of the same solve, 1 thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as ora  # test-side binding of the oracle
    t0 = time.perf_counter()
    ora.solve(lp, refactor_freq=refactor_freq, max_iterations=warmup)
    t1 = time.perf_counter()
    res, _, _ = ora.solve(lp, refactor_freq=refactor_freq, max_iterations=warmup + steps)
    t2 = time.perf_counter()
    window = (t2 - t1) - (t1 - t0)
    done = min(res.iterations, warmup + steps) - warmup
    return {"value": round(done / window, 3) if window > 0 else None, "unit": "iter/s",
            "ms_per_iter": round(1000.0 * window / max(done, 1), 4), "cores": 1, "kind": "port",
            "sample": "oracle on the same instance, iterations %d..%d from the slack basis "
                      "(time of a %d-iteration run minus a %d-iteration run), 1 thread, %s"
                      % (warmup + 1, warmup + done, warmup + steps, warmup, cpu_model())}


def _oracle_solve_one(args):
    import time as _t
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as ora
    from parallel_ds_amd import lp as lpmod
    k = args
    lp = lpmod.config3_instance(k)
    t0 = _t.perf_counter()
    r, _, _ = ora.solve(lp)
    return _t.perf_counter() - t0, r.status, r.Z


def batch_roofline(ms, m_of, iters, lds_inv_max_m):
    """HBM roofline of a batched solve: the dense B^-1 of every LP whose inverse is not kept in
    LDS is read and written once per basis change (16 m^2 bytes per LP-iteration, the fused
    FTRAN + DSE + product-form pass); vectors, pricing and refactors are not counted."""
    import numpy as np
    m = np.asarray(m_of, np.float64)
    glob = m > lds_inv_max_m
    byts = float(np.sum(16.0 * m[glob] ** 2 * np.asarray(iters, np.float64)[glob]))
    ach = byts / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_batch": round(byts),
            "per_lp_iteration": "16 m^2 (B^-1 read + write)"}


def batched_family(count, tune_sample, reps, cpu_sample):
    """4,096 packing LPs, m ~ U{32..256}, n = 2m, density 10%, one persistent launch."""
    from parallel_ds_amd import lp as lpmod
    from parallel_ds_amd.batch import BatchSolver, default_batch_params
    lps = [lpmod.config3_instance(k) for k in range(count)]
    log("batched: %d LPs generated; tuning" % count)
    with BatchSolver(lps) as B:
        best, tune_ms, tried = B.tune(sample=tune_sample, base=default_batch_params(max_iterations=100000))
        log("batched: tuned (%d candidates); solving" % tried)
        best.use_graph = 1
        first_ms = B.solve(best)
        times = [B.replay() for _ in range(reps)]
        st, obj, it, bases, _ = B.results()
    import numpy as np
    ms = float(np.median(times))
    out = {"workload": "%d packing LPs, m~U{32..256}, n=2m, density 10%%, seeds 3000+k" % count,
           "solves_per_s": round(count / (ms * 1e-3), 2), "ms_per_batch": round(ms, 3),
           "first_launch_ms": round(first_ms, 3), "replays": reps, "optimal": int((st == 0).sum()),
           "mean_iterations": round(float(it.mean()), 1),
           "ms_per_iter_per_lp": round(ms / max(1, float(it.sum())) * count, 5),
           "tuned": {"block": best.block, "lds_inv_max_m": best.lds_inv_max_m, "wgs_per_cu": best.wgs_per_cu,
                     "refactor_freq": best.refactor_freq,
                     "candidates": tried, "sample": tune_sample, "sample_ms": round(tune_ms, 3)},
           "roofline": dict(batch_roofline(ms, [lp.m for lp in lps], it, best.lds_inv_max_m),
                            traffic=BATCH_TRAFFIC_BYTES.get(count))}
    try:
        import json as _j
        g = _j.load(open(os.path.join(ROOT, "tests", "golden", "highs_objectives.json")))["config3"]
        rel = [abs(obj[k] - r["objective"]) / abs(r["objective"]) for k, r in enumerate(g) if k < count]
        out["max_rel_err_vs_highs_first_%d" % len(rel)] = float(max(rel))
    except Exception:
        pass
    if cpu_sample > 0:
        log("batched: CPU oracle on %d instances" % cpu_sample)
        import multiprocessing as mp
        cores = min(16, os.cpu_count() or 1)
        ks = list(range(cpu_sample))
        t0 = time.perf_counter()
        with mp.Pool(cores) as pool:
            res = pool.map(_oracle_solve_one, ks)
        wall = time.perf_counter() - t0
        single = sum(r[0] for r in res)
        out["cpu_baseline"] = {"value": round(cpu_sample / wall, 2), "unit": "solves/s", "cores": cores,
                               "kind": "port", "single_thread_solves_per_s": round(cpu_sample / single, 2),
                               "sample": "oracle on instances 0..%d of the same family, one process per core"
                                         % (cpu_sample - 1)}
    return out


def _oracle_family_solve(args):
    import time as _t
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as ora
    from parallel_ds_amd import lp as lpmod
    m, n, k, seed, sigma, s = args
    lp = lpmod.perturbed_instance(lpmod.gen_kcol(m, n, k, seed), sigma, s)
    t0 = _t.perf_counter()
    r, _, _ = ora.solve(lp)
    return _t.perf_counter() - t0, r.status, r.Z


def spmv_sweep(reps=100, flush_mib=640):
    """Pivot-row SpMV alone at three sizes (C4 base -> C1 -> C2), slack basis (every structural
    column nonbasic), warm and cold (caches flushed by a copy of flush_mib MiB before each launch)."""
    import numpy as np
    from parallel_ds_amd import DualSimplex, default_params
    from parallel_ds_amd import lp as lpmod
    cases = [("family base m=2000 n=5000 5 nnz/col", lpmod.config1(seed=4, m=2000, n=5000, k=5)),
             ("single LP m=20000 n=40000 10 nnz/col", lpmod.config1(seed=1)),
             ("transportation LP 500x1000 (m=1500, n=500000)", lpmod.config2(seed=2))]
    rows = []
    for name, lp in cases:
        log("SpMV sweep: %s" % name)
        rho = np.random.default_rng(0).standard_normal(lp.m)
        with DualSimplex(lp, default_params()) as s:
            s.initDualSimplex()
            w, c, by = s.time_pivot_row_ext(rho, reps, flush_mib << 20)
        rows.append({"workload": name, "bytes_per_launch": round(by),
                     "warm_us": round(w * 1e3, 3), "warm_gbs": round(by / (w * 1e-3) / 1e9, 1),
                     "warm_frac": round(by / (w * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "cold_us": round(c * 1e3, 3), "cold_gbs": round(by / (c * 1e-3) / 1e9, 1),
                     "cold_frac": round(by / (c * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
    return {"kernel": "k_spmv (pivot-row SpMV)", "reps": reps, "flush_mib": flush_mib, "peak_gbs": HBM_PEAK_GBS,
            "sizes": rows}


def transport(cpu, S=500, D=1000, seed=2, ratio_test=0, mark=200):
    """Balanced transportation LP S x D (n = S*D, m = S+D, nnz = 2n) solved to optimality on the GPU.

    SpMV roofline at this size (iterations after `mark`), objective against the HiGHS value of
    the same instance, and the oracle's full solve of the same LP on one host core."""
    from parallel_ds_amd import DualSimplex, default_params
    from parallel_ds_amd import lp as lpmod
    lp = lpmod.config2(seed=seed, S=S, D=D)
    log("transport %dx%d: GPU solve" % (S, D))
    prm = default_params(ratio_test=ratio_test, mark_iteration=mark, time_kernels=SPMV_TIME_STRIDE)
    with DualSimplex(lp, prm) as s:
        s.initDualSimplex()
        st = s.presolve("minInfeas")
        if st == 0:
            s.solve()
        r = s.result()
    spmv_avg_ms = r.spmv_ms / max(r.spmv_launches, 1)
    nnzN = 0.5 * (r.nnz_nonbasic_mark + r.nnz_nonbasic_end)
    spmv_bytes = 12.0 * nnzN + 4.0 * (lp.n + 1) + 8.0 * lp.n + 8.0 * lp.m
    achieved = spmv_bytes / (spmv_avg_ms * 1e-3) / 1e9 if spmv_avg_ms > 0 else 0.0
    out = {"workload": "balanced transportation LP %d sources x %d sinks (m=%d, n=%d, nnz=%d), seed %d, "
                       "solved from the slack basis" % (S, D, lp.m, lp.n, lp.nnz, seed),
           "ratio_test": "harris" if ratio_test else "textbook",
           "status": r.status, "objective": r.Z, "iterations": r.iterations, "perturbed": r.perturbed,
           "solve_ms": round(r.solve_ms, 2), "ms_per_iter": round(r.solve_ms / max(r.iterations, 1), 4),
           "solves_per_s": round(1e3 / r.solve_ms, 3) if r.solve_ms > 0 else None,
           "spmv_roofline": {"kernel": "k_spmv (pivot-row SpMV)", "bound": "hbm", "achieved": round(achieved, 2),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                             "traffic": TRANSPORT_SPMV_TRAFFIC_BYTES.get((S, D, seed)),
                             "bytes_per_launch": round(spmv_bytes), "avg_launch_us": round(spmv_avg_ms * 1e3, 3),
                             "launches": r.spmv_launches}}
    gold = os.path.join(ROOT, "tests", "golden", "highs_objectives.json")
    key = "transport_%dx%d_seed%d" % (S, D, seed)
    if os.path.exists(gold):
        g = json.load(open(gold)).get(key)
        if g:
            out["highs_objective"] = g["objective"]
            out["rel_err_vs_highs"] = abs(r.Z - g["objective"]) / max(abs(g["objective"]), 1.0)
    if cpu:
        log("transport %dx%d: CPU oracle full solve" % (S, D))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as ora  # test-side binding of the oracle
        t0 = time.perf_counter()
        ro, _, _ = ora.solve(lp, ratio_test=ratio_test)
        sec = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(1.0 / sec, 4), "unit": "solves/s", "solve_s": round(sec, 2),
                               "ms_per_iter": round(1e3 * sec / max(ro.iterations, 1), 4), "status": ro.status,
                               "objective": ro.Z, "iterations": ro.iterations, "cores": 1, "kind": "port",
                               "sample": "oracle full solve of the same LP, 1 thread, %s" % cpu_model()}
    return out


def learned_warm(n_train, n_eval, epochs, reps, cpu_sample, plain_sample=20, m=256, n=640, k=5, seed=4, sigma=0.1):
    """Learned warm start vs slack start on the same GPU: an LP family of the single-LP recipe
    (k nonzeros per column) whose b and c are perturbed by lognormal(0, sigma); the network is
    trained on the bases the batched GPU solver finds for n_train instances and evaluated on
    n_eval held-out instances, all solved in one batch per start."""
    import numpy as np
    from parallel_ds_amd import lp as lpmod
    from parallel_ds_amd.batch import BatchSolver, default_batch_params
    from parallel_ds_amd.warm import Mlp
    base = lpmod.gen_kcol(m, n, k, seed)
    train = [lpmod.perturbed_instance(base, sigma, 10000 + s) for s in range(n_train)]
    held = [lpmod.perturbed_instance(base, sigma, 20000 + s) for s in range(n_eval)]
    out = {"workload": "LP family m=%d n=%d, %d nnz/col (seed %d), b and c x lognormal(0, %g); train on %d "
                       "GPU-solved instances (seeds 10000+), evaluate on %d held-out (seeds 20000+)"
                       % (m, n, k, seed, sigma, n_train, n_eval)}
    log("learned warm start: %d + %d instances generated; tuning on the held-out batch" % (n_train, n_eval))
    with BatchSolver(held) as E:
        best, _, _ = E.tune(sample=min(256, n_eval), base=default_batch_params(max_iterations=200000))
        log("learned warm start: slack-start solves")
        prm = best
        prm.use_graph = 1
        # slack start
        E.set_warm(None)
        E.solve(prm)
        t_slack = [E.replay() for _ in range(reps)]
        st_s, obj_s, it_s, _, _ = E.results()
        with BatchSolver(train) as T, Mlp(seed=1) as net:
            log("learned warm start: solving the training instances")
            T.solve(prm)
            st_t, _, it_t, _, _ = T.results()
            log("learned warm start: training %d epochs" % epochs)
            losses, train_ms = net.fit(T, epochs=epochs, minibatch=8192, lr=1e-3, seed=3)
            log("learned warm start: losses %s; predicting" % [round(float(v), 4) for v in losses])
            tr_ms, tr_flops = net.train_stats()
            # learned warm start: features -> inference -> top-m -> repaired starting bases
            pred = [E.warm_start(net) for _ in range(reps)]
        E.solve(prm)
        t_warm = [E.replay() for _ in range(reps)]
        st_w, obj_w, it_w, _, _ = E.results()
        rep = E.repairs()
    ms_s = float(np.median(t_slack))
    ms_pred = float(np.median([p[0] for p in pred]))
    ms_inf = float(np.median([p[1] for p in pred]))
    ms_w = float(np.median(t_warm)) + ms_pred
    ok = (st_w == st_s) & (st_s == 0)
    rel = np.abs(obj_w - obj_s) / np.maximum(1.0, np.abs(obj_s))
    # inference kernel: 2 flops per MAC over the three hidden layers and the output
    rows = (m + n) * n_eval
    inf_flops = 2.0 * rows * (32 * 256 + 2 * 256 * 256 + 256)
    out.update({
        "slack": {"solves_per_s": round(n_eval / (ms_s * 1e-3), 2), "ms_per_batch": round(ms_s, 3),
                  "mean_iterations": round(float(it_s.mean()), 1),
                  "us_per_lp_iteration": round(1e3 * ms_s / max(1.0, float(it_s.sum())), 4)},
        "learned_warm": {"solves_per_s": round(n_eval / (ms_w * 1e-3), 2), "ms_per_batch": round(ms_w, 3),
                         "predict_ms": round(ms_pred, 3), "solve_ms": round(ms_w - ms_pred, 3),
                         "mean_iterations": round(float(it_w.mean()), 1),
                         "us_per_lp_iteration": round(1e3 * (ms_w - ms_pred) / max(1.0, float(it_w.sum())), 4),
                         "repaired_columns_mean": round(float(rep.mean()), 2)},
        "iterations_percentiles": {
            "slack": {q: round(float(np.percentile(it_s, v)), 1) for q, v in (("p50", 50), ("p90", 90), ("p99", 99), ("max", 100))},
            "learned_warm": {q: round(float(np.percentile(it_w, v)), 1) for q, v in (("p50", 50), ("p90", 90), ("p99", 99), ("max", 100))},
            "note": "the held-out LPs run concurrently (one workgroup each), so the batch time follows the slowest LP"},
        "roofline": {"slack": batch_roofline(ms_s, [m] * n_eval, it_s, prm.lds_inv_max_m),
                     "learned_warm": batch_roofline(ms_w - ms_pred, [m] * n_eval, it_w, prm.lds_inv_max_m)},
        "speedup_solves_per_s": round(ms_s / ms_w, 3),
        "iteration_ratio": round(float(it_s.mean()) / max(1e-9, float(it_w.mean())), 3),
        "status_match_and_optimal": int(ok.sum()), "of": n_eval,
        "max_obj_rel_err_warm_vs_slack": float(rel.max()),
        "training": {"instances_optimal": int((st_t == 0).sum()), "mean_iterations": round(float(it_t.mean()), 1),
                     "epochs": epochs, "epoch_loss": [round(float(v), 4) for v in losses],
                     "train_ms": round(train_ms, 1),
                     "train_gemm_tflops": round(tr_flops / (tr_ms * 1e-3) / 1e12, 2) if tr_ms > 0 else None},
        "inference_roofline": {"kernel": "k_infer (fused 3-layer BF16 MFMA)", "bound": "mfma",
                               "achieved": round(inf_flops / (ms_inf * 1e-3) / 1e12, 2), "peak": 2500.0,
                               "unit": "TFLOP/s", "frac": round(inf_flops / (ms_inf * 1e-3) / 1e12 / 2500.0, 4),
                               "avg_launch_us": round(ms_inf * 1e3, 2), "rows": rows},
        "tuned": {"block": prm.block, "lds_inv_max_m": prm.lds_inv_max_m, "wgs_per_cu": prm.wgs_per_cu,
                  "refactor_freq": prm.refactor_freq},
    })
    if plain_sample > 0:
        # The plain single-LP GPU solver (slack start, one LP at a time) on the first held-out LPs.
        from parallel_ds_amd import DualSimplex, default_params
        log("learned warm start: plain single-LP GPU solver on %d held-out instances" % plain_sample)
        t0 = time.perf_counter()
        zs, its, sts = [], [], []
        for lp in held[:plain_sample]:
            with DualSimplex(lp, default_params()) as s:
                s.initDualSimplex()
                if s.presolve("minInfeas") == 0:
                    s.solve()
                r = s.result()
            zs.append(r.Z)
            its.append(r.iterations)
            sts.append(r.status)
        sec = time.perf_counter() - t0
        zs = np.array(zs)
        agree = np.abs(zs - obj_s[:plain_sample]) <= 1e-9 * np.maximum(1.0, np.abs(obj_s[:plain_sample]))
        out["plain_single_lp_gpu"] = {
            "solves_per_s": round(plain_sample / sec, 3), "instances": plain_sample,
            "mean_iterations": round(float(np.mean(its)), 1), "optimal": int(np.sum(np.array(sts) == 0)),
            "objective_agrees_with_batched_slack": int(agree.sum()),
            "learned_warm_batched_speedup": round(out["learned_warm"]["solves_per_s"] / (plain_sample / sec), 1),
            "note": "wall time of load + solve per LP through the single-LP C-ABI path, one LP at a time"}
    if cpu_sample > 0:
        log("learned warm start: CPU oracle on %d instances" % cpu_sample)
        import multiprocessing as mp
        cores = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        with mp.Pool(cores) as pool:
            res = pool.map(_oracle_family_solve, [(m, n, k, seed, sigma, 20000 + s) for s in range(cpu_sample)])
        wall = time.perf_counter() - t0
        zs = np.array([r[2] for r in res])
        out["cpu_baseline"] = {"value": round(cpu_sample / wall, 2), "unit": "solves/s", "cores": cores,
                               "kind": "port", "single_thread_solves_per_s": round(cpu_sample / sum(r[0] for r in res), 2),
                               "max_obj_rel_err_gpu_slack_vs_oracle":
                                   float(np.max(np.abs(obj_s[:cpu_sample] - zs) / np.maximum(1.0, np.abs(zs)))),
                               "sample": "oracle (slack start) on held-out instances 0..%d, one process per core"
                                         % (cpu_sample - 1)}
    return out


def _solve_many(items, prm_of, threads):
    """Solve LPs on the single-LP path, `threads` contexts (streams) at a time; items are
    (lp, start basis or None). Returns (wall s, per-LP (s, status, Z, iterations, repairs, basis,
    solve_ms, phase1 iterations, restarts))."""
    import concurrent.futures as cf
    from parallel_ds_amd import solve_lp

    def run(item):
        lp, start, prm = item
        t = time.perf_counter()
        sol, basis, r = solve_lp(lp, prm, start=start)
        return (time.perf_counter() - t, r.status, sol.Z, r.iterations, r.repairs, basis, r.solve_ms,
                r.phase1_iterations, r.restarts, r.primal_iterations, r.shifted_bounds)

    t0 = time.perf_counter()
    out = [None] * len(items)
    with cf.ThreadPoolExecutor(threads) as ex:
        futs = {ex.submit(run, (lp, st, prm_of(k))): k for k, (lp, st) in enumerate(items)}
        for f in cf.as_completed(futs):
            k = futs[f]
            out[k] = f.result()
            log("  %s: status %d, %d iterations, %.2f s" % (items[k][0].name, out[k][1], out[k][3], out[k][0]))
    return time.perf_counter() - t0, out


def learned_warm_c4(steps, warmup, n_train, threads, refactor_freq, epochs, cpu_iters, cpu_from=1500, m=2000,
                    n=5000, k=5, seed=4, sigma=0.1, slack_limit_s=30.0, warm_primal=1, profile_part=None):
    """Learned warm start against slack start at the family's stated size: the single-LP recipe at
    m=2,000, n=5,000 with 5 nonzeros per column (base seed 4), b and c times lognormal(0, 0.1).
    The network is trained on the optimal bases the GPU solver finds for n_train instances (seeds
    10000+); `warmup` held-out instances (seeds 20000+) run untimed, then `steps` held-out instances
    (the next seeds) are solved from the slack basis and from the learned warm start, `threads` LPs
    at a time (one context and HIP stream each). A warm solve that has not ended optimal within
    twice the mean training solve time is re-solved from the slack basis (the fallback the slack
    start provides); its time counts in full.  profile_part "train" stops after training, "warm"
    solves every held-out instance from the learned start only (kernel profiles of the two parts:
    the warm solves' kernels are the difference of the two traces)."""
    import numpy as np
    from parallel_ds_amd import default_params
    from parallel_ds_amd import lp as lpmod
    from parallel_ds_amd.batch import BatchSolver
    from parallel_ds_amd.warm import Mlp
    base = lpmod.gen_kcol(m, n, k, seed)
    train = [lpmod.perturbed_instance(base, sigma, 10000 + s) for s in range(n_train)]
    held = [lpmod.perturbed_instance(base, sigma, 20000 + s) for s in range(warmup + steps)]
    # The reference's restart loops have no exit; a family solve that runs past slack_limit_s ends
    # with the iteration-limit status and is reported as not optimal.
    slack_prm = default_params(refactor_freq=refactor_freq, time_limit_ms=1e3 * slack_limit_s)
    log("learned warm start m=%d n=%d: solving %d training instances, %d at a time" % (m, n, n_train, threads))
    t_tr, tr = _solve_many([(lp, None) for lp in train], lambda i: slack_prm, threads)
    tr_ok = [r for r in tr if r[1] == 0]
    keep = [i for i, r in enumerate(tr) if r[1] == 0]
    train = [train[i] for i in keep]
    bases = [tr[i][5] for i in keep]
    limit_ms = 2.0 * float(np.mean([r[6] for r in tr_ok]))
    warm_prm = default_params(refactor_freq=refactor_freq, time_limit_ms=limit_ms, warm_primal=warm_primal)
    with BatchSolver(train) as T, Mlp(seed=1) as net:
        T.set_bases(bases)
        losses, fit_ms = net.fit(T, epochs=epochs, minibatch=2048, lr=1e-3, seed=3)
        log("learned warm start: losses %s" % [round(float(v), 4) for v in losses])
        tr_ms, tr_flops = net.train_stats()
        with BatchSolver(held) as E:
            t0 = time.perf_counter()
            E.warm_start(net)
            pred_s = time.perf_counter() - t0
            warm = E.get_warm()
    overlap = []
    part = {"workload": "LP family m=%d n=%d: %s only (profiling run, no headline)" % (m, n, profile_part),
            "learned_warm": {"solves_per_s": None, "wall_s": 0.0}, "slack": {"solves_per_s": None},
            "speedup_solves_per_s": None}
    if profile_part == "train":
        return part

    def warm_phase(lps, starts):
        t0 = time.perf_counter()
        wall_w, rw = _solve_many(list(zip(lps, starts)), lambda i: warm_prm, threads)
        redo = [i for i, r in enumerate(rw) if r[1] != 0]
        if redo:
            _, rf = _solve_many([(lps[i], None) for i in redo], lambda i: slack_prm, threads)
            for i, r in zip(redo, rf):
                rw[i] = r
        return time.perf_counter() - t0, rw, len(redo)

    if profile_part == "warm":
        log("learned warm start: %d held-out instances from the learned warm start only" % len(held))
        wall_w, rw, fallbacks = warm_phase(held, warm)
        part["learned_warm"].update({"wall_s": round(wall_w, 3), "solves": len(held), "fallbacks_to_slack": fallbacks,
                                     "mean_iterations": round(float(np.mean([r[3] for r in rw])), 1)})
        return part
    if warmup > 0:
        log("learned warm start: %d untimed held-out instances" % warmup)
        _solve_many([(lp, None) for lp in held[:warmup]], lambda i: slack_prm, threads)
        warm_phase(held[:warmup], warm[:warmup])
    ev, ew = held[warmup:], warm[warmup:]
    log("learned warm start: %d held-out instances from the slack basis" % steps)
    wall_s, rs = _solve_many([(lp, None) for lp in ev], lambda i: slack_prm, threads)
    log("learned warm start: the same %d from the learned warm start" % steps)
    wall_w, rw, fallbacks = warm_phase(ev, ew)
    wall_w += pred_s * steps / len(held)
    overlap = [len(set(w.tolist()) & set(r[5].tolist())) / m for w, r in zip(ew, rs)]
    both = [(a, b) for a, b in zip(rs, rw) if a[1] == 0 and b[1] == 0]
    ok = len(both) == len(rs)
    rel = max([abs(a[2] - b[2]) / max(1.0, abs(a[2])) for a, b in both] or [0.0])
    out = {
        "workload": "LP family m=%d n=%d, %d nnz/col (base seed %d), b and c x lognormal(0, %g); trained on %d "
                    "GPU-solved instances (seeds 10000+), %d untimed + %d timed held-out instances (seeds 20000+), "
                    "%d solves at a time, refactor every %d" % (m, n, k, seed, sigma, n_train, warmup, steps, threads,
                                                               refactor_freq),
        "slack": {"solves_per_s": round(steps / wall_s, 4), "wall_s": round(wall_s, 3),
                  "mean_iterations": round(float(np.mean([r[3] for r in rs])), 1),
                  "mean_phase1_iterations": round(float(np.mean([r[7] for r in rs])), 1),
                  "mean_solve_ms": round(float(np.mean([r[6] for r in rs])), 1),
                  "ms_per_iter": round(float(np.sum([r[6] for r in rs]) / np.sum([r[3] for r in rs])), 4)},
        "learned_warm": {"solves_per_s": round(steps / wall_w, 4), "wall_s": round(wall_w, 3),
                         "mean_iterations": round(float(np.mean([r[3] for r in rw])), 1),
                         "mean_phase1_iterations": round(float(np.mean([r[7] for r in rw])), 1),
                         "mean_restarts": round(float(np.mean([r[8] for r in rw])), 1),
                         "primal_phase": bool(warm_primal),
                         "mean_primal_iterations": round(float(np.mean([r[9] for r in rw])), 1),
                         "mean_shifted_bounds": round(float(np.mean([r[10] for r in rw])), 1),
                         "fallbacks_to_slack": fallbacks, "time_limit_ms": round(limit_ms, 1),
                         "repaired_columns_mean": round(float(np.mean([r[4] for r in rw])), 2),
                         "predict_s": round(pred_s, 4)},
        "speedup_solves_per_s": round(wall_s / wall_w, 4),
        # like-for-like per solve: summed solve times over the instances both starts solved to optimality
        "speedup_solve_ms_both_optimal": round(sum(a[6] for a, b in both) / max(1e-9, sum(b[6] for a, b in both)), 4),
        "predicted_basis_overlap_with_optimal": round(float(np.mean(overlap)), 4),
        "all_optimal": bool(ok), "optimal_both": len(both), "slack_not_optimal": sum(a[1] != 0 for a in rs),
        "max_obj_rel_err_warm_vs_slack": float(rel),
        "training": {"instances": n_train, "optimal": len(keep), "solve_wall_s": round(t_tr, 2), "minibatch": 2048,
                     "mean_iterations": round(float(np.mean([r[3] for r in tr])), 1),
                     "epochs": epochs, "epoch_loss": [round(float(v), 4) for v in losses], "fit_ms": round(fit_ms, 1),
                     "train_gemm_tflops": round(tr_flops / (tr_ms * 1e-3) / 1e12, 2) if tr_ms > 0 else None},
    }
    if cpu_iters > 0:
        # The oracle's cost per iteration grows with the basis kernel (dense LU of k x k): it is
        # timed over two bounded windows of the first timed instance, the first cpu_iters
        # iterations from the slack basis (k ~ 0) and cpu_iters iterations from a training
        # instance's optimal basis (k ~ 1,600, the end of a solve); a full solve is estimated as
        # the GPU's mean slack iterations times the mean of the two rates (cost linear in the
        # iteration count between the two ends).
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as ora  # test-side binding of the oracle
        log("learned warm start: CPU oracle, %d iterations at each end of a solve" % cpu_iters)
        t0 = time.perf_counter()
        r1, _, _ = ora.solve(ev[0], refactor_freq=refactor_freq, max_iterations=cpu_iters, tie_rule=1)
        s1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        r2, _, _ = ora.solve_from(ev[0], bases[0], refactor_freq=refactor_freq, max_iterations=cpu_iters, tie_rule=1)
        s2 = time.perf_counter() - t0
        ms_early = 1e3 * s1 / max(r1.iterations, 1)
        ms_late = 1e3 * s2 / max(r2.iterations, 1)
        est_s = out["slack"]["mean_iterations"] * 0.5 * (ms_early + ms_late) * 1e-3
        out["cpu_baseline"] = {
            "value": round(1.0 / est_s, 6), "unit": "solves/s", "cores": 1, "kind": "port",
            "ms_per_iter_early": round(ms_early, 4), "ms_per_iter_late": round(ms_late, 4),
            "estimated_solve_s": round(est_s, 1),
            "sample": "oracle on held-out instance %d, 1 thread, %s: %d iterations from the slack basis (%.1f s) "
                      "and %d from a training instance's optimal basis (%.1f s); solve time estimated as the "
                      "GPU's mean slack iterations x the mean of the two per-iteration costs"
                      % (warmup, cpu_model(), r1.iterations, s1, r2.iterations, s2)}
    return out


def single_lp_window(args):
    """The single large LP (m=20,000, n=40,000, 10 nonzeros per column, refactor every 64) over a
    fixed iteration window from the slack basis (iterations c1_from+1 .. c1_from+c1_window, default
    501..2,500), its pivot-row SpMV roofline from event pairs on every 8th SpMV launch of the window,
    the kernel size and refactor time, and the oracle over the same window."""
    from parallel_ds_amd import DualSimplex, default_params
    from parallel_ds_amd import lp as lpmod
    lp = lpmod.config1(seed=args.seed, m=args.m, n=args.n, k=args.k)
    w0, w = args.c1_from, args.c1_window
    log("single LP m=%d n=%d: iterations %d..%d" % (lp.m, lp.n, w0 + 1, w0 + w))
    prm = default_params(refactor_freq=args.refactor, chunk=args.chunk, max_iterations=w0 + w,
                         mark_iteration=w0, time_kernels=SPMV_TIME_STRIDE)
    with DualSimplex(lp, prm) as s:
        s.initDualSimplex()
        if s.presolve("minInfeas") == 0:
            s.solve()
        r = s.result()
    timed = (r.iterations - w0) if (r.iterations < w0 + w and r.status != 3) else w
    window_ms = r.solve_ms - r.mark_ms
    spmv_avg_ms = r.spmv_ms / max(r.spmv_launches, 1)
    nnzN = 0.5 * (r.nnz_nonbasic_mark + r.nnz_nonbasic_end)
    spmv_bytes = 12.0 * nnzN + 4.0 * (lp.n + 1) + 8.0 * lp.n + 8.0 * lp.m
    achieved = spmv_bytes / (spmv_avg_ms * 1e-3) / 1e9 if spmv_avg_ms > 0 else 0.0
    out = {"workload": "single sparse LP m=%d n=%d, %d nnz/col, seed %d, refactor every %d, iterations %d..%d from "
                       "the slack basis" % (lp.m, lp.n, args.k, args.seed, args.refactor, w0 + 1, w0 + timed),
           "iter_per_s": round(timed / (window_ms * 1e-3), 2), "ms_per_iter": round(window_ms / max(timed, 1), 5),
           "status_at_end": r.status, "kernel_size_at_end": r.kernel_k, "refactors": r.refactors,
           "refactor_ms_total": round(r.refactor_ms, 2),
           "roofline": {"kernel": "k_spmv (pivot-row SpMV)", "bound": "hbm", "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": SPMV_TRAFFIC_BYTES.get((args.m, args.n, args.k, args.seed, args.refactor)),
                        "bytes_per_launch": round(spmv_bytes), "avg_launch_us": round(spmv_avg_ms * 1e3, 3),
                        "launches": r.spmv_launches}}
    if not args.no_cpu_baseline:
        log("single LP: CPU oracle over the same window")
        out["cpu_baseline"] = cpu_baseline(lp, args.refactor, w0, timed)
    return out


def refactor_mfma(k_target=1600, reps=3):
    """FP64 MFMA use of the refactorization's trailing update (k_gemm_sub) on a basis of the
    learned-warm-start family's size: the slack solve's final basis (kernel k ~ 1,600), factored
    `reps` times with an event pair around every trailing-update launch. Executed flops per launch:
    2 (k - nb) k nb for a panel of nb columns."""
    from parallel_ds_amd import DualSimplex, default_params, solve_lp
    from parallel_ds_amd import lp as lpmod
    lp = lpmod.gen_kcol(2000, 5000, 5, 4)
    sol, basis, r = solve_lp(lp, default_params(refactor_freq=150))
    with DualSimplex(lp, default_params()) as s:
        s.initDualSimplex()
        s.set_basis(basis)
        t = s.time_refactor(reps)
    ach = t["gemm_flops"] / (t["gemm_ms"] * 1e-3) / 1e12 if t["gemm_ms"] > 0 else 0.0
    tot = 2.0 * t["k"] ** 3 / (t["refactor_ms"] * 1e-3) / 1e12 if t["refactor_ms"] > 0 else 0.0
    return {"kernel": "k_gemm_sub (FP64 MFMA trailing update)", "bound": "mfma", "k": t["k"],
            "achieved": round(ach, 3), "peak": 78.6, "unit": "TFLOP/s", "frac": round(ach / 78.6, 4),
            "gemm_ms_per_refactor": round(t["gemm_ms"], 3), "refactor_ms": round(t["refactor_ms"], 3),
            "refactor_tflops_2k3": round(tot, 3), "refactor_frac_2k3": round(tot / 78.6, 4),
            "sample": "basis of the optimal slack solve of the family base LP (m=2000, n=5000), %d factorizations"
                      % reps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="held-out family LPs timed (per start)")
    ap.add_argument("--warmup", type=int, default=1, help="held-out family LPs solved untimed first")
    ap.add_argument("--train", type=int, default=16, help="GPU-solved training instances of the family")
    ap.add_argument("--threads", type=int, default=20,
                    help="family LPs solved at a time (contexts/streams, one hardware queue each)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="GPU_MAX_HW_QUEUES for this process (0: one per concurrent family solve)")
    ap.add_argument("--family-refactor", type=int, default=150)
    ap.add_argument("--epochs", type=int, default=48, help="training epochs of the family network (minibatch 2048)")
    ap.add_argument("--family-cpu-iters", type=int, default=300)
    ap.add_argument("--family-profile", choices=["train", "warm"], default=None,
                    help="profiling runs of the family leg: stop after training, or solve the held-out "
                         "instances from the learned start only (no headline value)")
    ap.add_argument("--warm-primal", type=int, default=1,
                    help="1: primal simplex iterations from the bound-shifted predicted basis before the dual simplex")
    ap.add_argument("--m", type=int, default=20000)
    ap.add_argument("--n", type=int, default=40000)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--refactor", type=int, default=64)
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--c1-from", type=int, default=500)
    ap.add_argument("--c1-window", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--batch-cpu-sample", type=int, default=512)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--warm-train", type=int, default=10000)
    ap.add_argument("--warm-eval", type=int, default=1000)
    ap.add_argument("--warm-epochs", type=int, default=8)
    ap.add_argument("--warm-cpu-sample", type=int, default=64)
    ap.add_argument("--no-warm", action="store_true")
    ap.add_argument("--no-transport", action="store_true")
    ap.add_argument("--no-spmv-sweep", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-family", action="store_true")
    ap.add_argument("--no-mfma", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy a device buffer of this size once (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    # Hardware queues the family solves' streams are spread over, set before the first HIP call
    # of the process (HIP's default is 4; at most 32): by default one per concurrent solve.
    # Measured at the family size with 20 solves at a time: 20 queues 1.97 / 0.91 solves/s
    # (learned / slack), 4 queues 1.73 / 0.80, 5 queues 1.00 / 0.66.
    q = args.hw_queues if args.hw_queues > 0 else max(4, args.threads)
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, q))
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    if args.no_family:  # profiling runs of the other legs only: no headline value
        fam = {"learned_warm": {"solves_per_s": None, "wall_s": 0.0}, "slack": {"solves_per_s": None},
               "speedup_solves_per_s": None, "workload": "not run (--no-family)"}
    else:
        fam = learned_warm_c4(args.steps, args.warmup, args.train, args.threads, args.family_refactor, args.epochs,
                              0 if args.no_cpu_baseline else args.family_cpu_iters, warm_primal=args.warm_primal,
                              profile_part=args.family_profile)
    single = None if args.no_single else single_lp_window(args)
    out = {
        "metric": METRIC,
        "value": fam["learned_warm"]["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * fam["learned_warm"]["wall_s"] / max(args.steps, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 LP generator)",
        "config": {"workload": fam["workload"], "m": 2000, "n": 5000, "nnz_per_col": 5, "sigma": 0.1,
                   "train_instances": args.train, "held_out_timed": args.steps, "held_out_untimed": args.warmup,
                   "concurrent_solves": args.threads, "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
                   "refactor_freq": args.family_refactor,
                   "slack_solves_per_s": fam["slack"]["solves_per_s"], "speedup": fam["speedup_solves_per_s"]},
        "roofline": single["roofline"] if single else None,
        "cpu_baseline": fam.get("cpu_baseline"),
        "learned_warm_family_m2000": fam,
    }
    if single:
        out["single_lp"] = single
    if not args.no_mfma:
        log("refactorization: FP64 MFMA timing")
        out["refactor_mfma"] = refactor_mfma()
    if not args.no_batch:
        out["batched"] = batched_family(args.batch_count, 256, args.batch_reps,
                                        0 if args.no_cpu_baseline else args.batch_cpu_sample)
    if not args.no_spmv_sweep:
        out["spmv_sweep"] = spmv_sweep()
    if not args.no_transport:
        out["transport"] = transport(not args.no_cpu_baseline, ratio_test=1)
    if not args.no_warm:
        out["learned_warm_start_m256"] = learned_warm(args.warm_train, args.warm_eval, args.warm_epochs,
                                                      args.batch_reps,
                                                      0 if args.no_cpu_baseline else args.warm_cpu_sample)
    if args.copy_calibration_mib > 0:
        out["copy_calibration"] = copy_calibration(args.copy_calibration_mib)
    print(json.dumps(out), flush=True)


def copy_calibration(mib):
    """One plain copy kernel over `mib` MiB (k_copy): its FETCH_SIZE and WRITE_SIZE counters should
    each read `mib` MiB, which gives the scale of those counters."""
    import ctypes as C
    from parallel_ds_amd._native import lib
    ms = C.c_double()
    rc = lib().ds_debug_copy(mib * 1024 * 1024, C.byref(ms))
    if rc != 0:
        raise RuntimeError("ds_debug_copy failed: %d" % rc)
    return {"bytes": mib * 1024 * 1024, "ms": round(ms.value, 4),
            "gbs": round(2 * mib * 1024 * 1024 / (ms.value * 1e-3) / 1e9, 1)}

if __name__ == "__main__":
    main()
