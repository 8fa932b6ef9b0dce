"""Benchmark of the batched candidate-fitness path on one MI355X.

Workload (the batched candidate-eval configuration): 4096 random expression
trees of 3-50 nodes, ops uniform over {+,-,*,/,sin,cos,ln,exp,sqrt,sqr},
leaves 50/50 a feature or a N(0,1) constant, each rewritten and split into at
most 8 additive factors by the product's host engine and OLS-fit; dataset
n = 100,000 rows x 10 features uniform in [-5, 5], target a random 15-node
tree plus N(0, 0.01) noise.  Seeds: trees 12345, data 23654.

One step = one rr_fitness_batch call over all 4096 candidates (program upload,
kernels, result download), with the dataset resident in HBM.  Both forms run:
the plain FP64 form (bit-for-bit the canonical algorithm) and the FP32
precision form (single pass: FP32 evaluation with per-value error bounds, Gram
sums in FP64, FP64 solve + one refinement step; a result is certified only
when its R2 provably lies within 1e-9 of the FP64 form's); an FP32 step also
re-scores in FP64 every candidate the FP32 kernels leave uncertified, so each
step yields a trusted fitness for every candidate.  `value` is the evals/s of
the faster of the two forms (named in config.form); the other is in `forms`.

Roofline (both forms VALU-bound; see the roofline objects): algorithmic flops
F = n (E + k^2 + 5k + 3) per candidate (E operator nodes, k columns with the
free term), credited only for work a launch executed -- a candidate the
factor-0 pre-pass settles (RR_F_EARLY) counts factor 0 on its 64 rows.  The
dominant kernel of the FP64 form is the factor evaluation (each distinct
factor of a live candidate over n rows), of the FP32 form its bucket kernels;
their times are HIP events around those launches on the stream they run on
(rr_last_phase_ms, option phase_timing), averaged over the timed steps.
`traffic` comes from the rocprofv3 PMC summary under profiles/ for the FP64
step (FETCH_SIZE and WRITE_SIZE of separate runs, scaled by the same counters
on a plain copy of known size: `python bench.py --copy-probe`), and is set
against the algorithmic bytes 2 n (d + 1) 8 of the two-pass form.
The timed FP64 results are compared bit for bit with the oracle's on the CPU
baseline's sample of candidates (fp64_vs_oracle).

Prints one JSON line.  --gpus N with N > 1 exits: the program measures one GPU.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "rils-rols_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

N_ROWS, N_FEAT, N_TREES = 100_000, 10, 4096
BARE = re.compile(r"^x\d+$")   # a factor that is one variable
RR_F_EARLY = 8                  # rr_fit.flags: settled by the factor-0 pre-pass
PEAK_HBM_GBS = 8000.0
PEAK_FP64_VALU_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3

UN = ["sin", "cos", "ln", "exp", "sqrt", "sqr"]
BI = ["+", "-", "*", "/"]


def gen_tree(rng, size, nvars):
    if size <= 1:
        if rng.random() < 0.5:
            return ["c" + float(rng.normal()).hex()]
        return ["x%d" % rng.integers(nvars)]
    op = UN[rng.integers(len(UN))] if size == 2 else (BI + UN)[rng.integers(len(BI) + len(UN))]
    if op in UN:
        return [op] + gen_tree(rng, size - 1, nvars)
    left = int(rng.integers(1, size - 1))
    return [op] + gen_tree(rng, left, nvars) + gen_tree(rng, size - 1 - left, nvars)


def eval_prefix(toks, X):
    """numpy evaluation of a prefix tree (synthetic target generation only)."""
    pos = [0]

    def rec():
        t = toks[pos[0]]
        pos[0] += 1
        if t.startswith("x"):
            return X[:, int(t[1:])]
        if t.startswith("c") and t != "cos":
            return np.full(X.shape[0], float.fromhex(t[1:]))
        if t in UN:
            a = rec()
            with np.errstate(all="ignore"):
                return {"sin": np.sin, "cos": np.cos, "ln": np.log, "exp": np.exp, "sqrt": np.sqrt,
                        "sqr": np.square}[t](a)
        a = rec()
        b = rec()
        with np.errstate(all="ignore"):
            return {"+": a + b, "-": a - b, "*": a * b, "/": a / b}[t]
    return rec()


def workload():
    import rils_rols_cpp
    rng = np.random.default_rng(23654)
    X = rng.uniform(-5, 5, (N_ROWS, N_FEAT))
    while True:
        tgt = gen_tree(rng, 15, N_FEAT)
        yt = eval_prefix(tgt, X)
        if np.all(np.isfinite(yt)) and np.std(yt) > 1e-3 and np.max(np.abs(yt)) < 1e6:
            break
    y = yt + rng.normal(0, 0.01, N_ROWS)
    trng = np.random.default_rng(12345)
    trees, factors = [], []
    while len(trees) < N_TREES:
        t = " ".join(gen_tree(trng, int(trng.integers(3, 51)), N_FEAT))
        fs = rils_rols_cpp._ols_factors(t)
        if len(fs) <= 8:
            trees.append(t)
            factors.append(fs)
    return X, y, " ".join(tgt), trees, factors


def ops_in(prefix):
    return sum(1 for t in prefix.split() if not (t.startswith("x") or (t.startswith("c") and t != "cos")))


def timed_steps(runner, rescorer, precision, warmup, steps):
    """K steps of one batched call.  In the FP32 form ("fp32") a step also
    re-scores in FP64, with a second call, every candidate the FP32 kernels
    could not certify (RR_F_NEEDS64); the checked form ("fp32c") does both in
    one call.  Either way each step delivers a trusted fitness for every
    candidate."""
    fp32_fit = {}

    phase_sum = {}

    def one():
        coefs, fit = runner.run(precision)
        ms = runner.last_kernel_ms()
        for k, v in runner.last_phase_ms().items():
            phase_sum[k] = phase_sum.get(k, 0.0) + v
        n64 = 0
        if precision == "fp32c":
            n64 = int(np.count_nonzero(fit[:, 4].astype(np.int64) & 4))
        if precision == "fp32":
            fp32_fit["fit"] = fit.copy()          # the FP32 form's own flags and bounds
            need = np.nonzero(fit[:, 4].astype(np.int64) & 2)[0]
            n64 = len(need)
            if n64:
                rescorer.select(need)
                c2, f2 = rescorer.run_selected("fp64")
                ms += rescorer.last_kernel_ms()
                fp32_fit["rescore_ms"] = rescorer.last_kernel_ms()
                fit[need] = f2
        return coefs, fit, ms, n64
    for _ in range(warmup):
        one()
    phase_sum.clear()
    lat, kms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        coefs, fit, ms, n64 = one()
        lat.append(time.perf_counter() - s)
        kms.append(ms)
    total = time.perf_counter() - t0
    phases = {k: v / max(steps, 1) for k, v in phase_sum.items()}
    return total, np.array(lat), np.array(kms), coefs, fit, n64, fp32_fit, phases


def timed_steps_multi(runners, precision, warmup, steps):
    """K steps of one batched call spread over len(runners) contexts, one host
    thread each: every context has its own stream and buffers, so one thread
    checks and plans its next batch while another context's kernels run.
    Returns the wall time of all K steps and each call's latency."""
    import threading
    for r in runners:
        for _ in range(warmup):
            r.run(precision)
    share = [steps // len(runners) + (1 if i < steps % len(runners) else 0) for i in range(len(runners))]
    lat = [[] for _ in runners]
    errors = []
    gate = threading.Barrier(len(runners) + 1)

    def work(i):
        gate.wait()
        try:
            for _ in range(share[i]):
                s = time.perf_counter()
                runners[i].run(precision)
                lat[i].append(time.perf_counter() - s)
        except Exception as e:   # reported by the main thread
            errors.append(e)
    th = [threading.Thread(target=work, args=(i,)) for i in range(len(runners))]
    for t in th:
        t.start()
    gate.wait()
    t0 = time.perf_counter()
    for t in th:
        t.join()
    total = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return total, np.array(sum(lat, []))


def cpu_baseline(X, y, trees, seconds_target=15.0, threads=1):
    """The oracle (CPU restatement of the reference path, g++ -O3) on a bounded
    prefix of the same candidates: tune_constants + fitness per candidate,
    `threads` host threads over candidates, with Eigen's QR order on the data
    (the reference's algorithm and cost, not the chunked order the GPU
    matches)."""
    import oracle_lib as O
    od = O.Dataset(X, y)
    O.set_qr_direct(True)
    try:
        probe = 8 * threads
        secs, _ = od.batch_tune(trees[:probe], threads=threads)
        per = secs / probe
        count = int(min(len(trees), max(probe, seconds_target / max(per, 1e-6))))
        secs, _ = od.batch_tune(trees[:count], threads=threads)
    finally:
        O.set_qr_direct(False)
    return count / secs, count, secs


def oracle_fit(X, y, trees, threads):
    """The oracle's results in the canonical order, for the bit-for-bit check."""
    import oracle_lib as O
    return O.Dataset(X, y).batch_tune(trees, threads=threads)[1]


def host_threads():
    """Host threads this job may use: the CPU share of the box (at most 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def pmc_traffic(kernel_tag):
    """HBM bytes per launch of a kernel from a committed PMC summary
    (profiles/pmc_<tag>.json: fetch_kib and write_kib per dispatch, and the
    copy-kernel scale factors), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_%s.json" % kernel_tag)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return 1024.0 * (d["fetch_kib"] * d["fetch_scale"] + d["write_kib"] * d["write_scale"])


def pmc_counters(kernel_tag):
    """The dominant kernel's VALU counter figures from a committed PMC summary
    (profiles/pmc_<tag>.json: eval_kernel_counters), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_%s.json" % kernel_tag)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("eval_kernel_counters")


def copy_probe():
    """A plain device copy of known size, for scaling the PMC byte counters."""
    import torch
    nbytes = 1 << 30
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "launches": 5}))


def serving_workload(n_runs=64, rate=4.0, seed=31337, keep=None):
    """The serving mix: n_runs independent regression runs arriving as a Poisson
    stream (rate runs/s); per run n log-uniform in [1e3, 1e6], features uniform
    in [2, 20], target a random 5-20-node tree plus 1% noise (sigma = 0.01 RMS(y),
    the reference's noisefy convention, rils_rols/utils.py:13-24); a quarter of
    the runs are binary classification (target thresholded at its median).
    keep: indices of the runs to return (all by default); the others are drawn
    all the same, so a kept run is the same whatever else is kept."""
    rng = np.random.default_rng(seed)
    t = 0.0
    runs = []
    cls_idx = set(rng.choice(n_runs, n_runs // 4, replace=False).tolist())
    for i in range(n_runs):
        t += rng.exponential(1.0 / rate)
        n = int(round(10 ** rng.uniform(3, 6)))
        d = int(rng.integers(2, 21))
        X = rng.uniform(-3, 3, (n, d))
        while True:
            tgt = gen_tree(rng, int(rng.integers(5, 21)), d)
            yt = eval_prefix(tgt, X)
            if np.all(np.isfinite(yt)) and np.std(yt) > 1e-6 and np.max(np.abs(yt)) < 1e8:
                break
        y = yt + rng.normal(0, 0.01 * np.sqrt(np.mean(yt ** 2)), n)
        if i in cls_idx:
            y = (y > np.median(y)).astype(np.float64)
        if keep is None or i in keep:
            runs.append(dict(index=i, arrival=t, X=X, y=y, classification=i in cls_idx, seed=1000 + i,
                             target=" ".join(tgt)))
    return runs


def serving_plan(n_runs=64, rate=4.0, seed=31337):
    """(index, arrival, n, d, classification) of every run of serving_workload,
    without keeping the datasets."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    cls_idx = set(rng.choice(n_runs, n_runs // 4, replace=False).tolist())
    for i in range(n_runs):
        t += rng.exponential(1.0 / rate)
        n = int(round(10 ** rng.uniform(3, 6)))
        d = int(rng.integers(2, 21))
        X = rng.uniform(-3, 3, (n, d))
        while True:
            tgt = gen_tree(rng, int(rng.integers(5, 21)), d)
            yt = eval_prefix(tgt, X)
            if np.all(np.isfinite(yt)) and np.std(yt) > 1e-6 and np.max(np.abs(yt)) < 1e8:
                break
        rng.normal(0, 0.01 * np.sqrt(np.mean(yt ** 2)), n)
        out.append((i, t, n, d, i in cls_idx))
    return out


def precision_workload(n=1_000_000, ntrees=16384, d=20):
    """The precision-path batch: ntrees trees of 5-50 nodes (seed 4242) with at
    most 8 additive factors, over n x d uniform in [-5, 5] (seed 2424) with a
    15-node target plus N(0, 0.01) noise."""
    import rils_rols_cpp
    rng = np.random.default_rng(2424)
    X = rng.uniform(-5, 5, (n, d))
    while True:
        tgt = gen_tree(rng, 15, d)
        yt = eval_prefix(tgt, X)
        if np.all(np.isfinite(yt)) and np.std(yt) > 1e-3 and np.max(np.abs(yt)) < 1e6:
            break
    y = yt + rng.normal(0, 0.01, n)
    trng = np.random.default_rng(4242)
    trees = []
    while len(trees) < ntrees:
        t = " ".join(gen_tree(trng, int(trng.integers(5, 51)), d))
        if len(rils_rols_cpp._ols_factors(t)) <= 8:
            trees.append(t)
    return X, y, trees


def search_workload(n=1_000_000, d=5):
    """The full-search dataset: n x d uniform in [-3, 3] (seed 777), target a
    seeded 15-node tree plus noise of SD 0.01 RMS(y) (the reference's noisefy
    convention, rils_rols/utils.py:13-24)."""
    rng = np.random.default_rng(777)
    X = rng.uniform(-3, 3, (n, d))
    while True:
        tgt = gen_tree(rng, 15, d)
        yt = eval_prefix(tgt, X)
        if np.all(np.isfinite(yt)) and np.std(yt) > 1e-3 and np.max(np.abs(yt)) < 1e6:
            break
    y = yt + rng.normal(0, 0.01 * np.sqrt(np.mean(yt ** 2)), n)
    return X, y, tgt


def serving_main(args):
    import rils_rols_cpp
    t0 = time.perf_counter()
    runs = serving_workload(args.runs)
    t_gen = time.perf_counter() - t0
    srv = rils_rols_cpp.Server(0, 0, args.max_active, 1, 0, args.gather_us)
    start = time.perf_counter()
    ids = []
    for r in runs:
        wait = r["arrival"] - (time.perf_counter() - start)
        if wait > 0:
            time.sleep(wait)
        ids.append(srv.submit(r["X"], r["y"], classification=r["classification"], max_fit_calls=args.run_fit_calls,
                              max_seconds=10 ** 6, random_state=r["seed"], precision=args.serving_precision))
        print("[serving] submitted %d n=%d d=%d at %.2fs" % (len(ids) - 1, r["X"].shape[0], r["X"].shape[1],
                                                            time.perf_counter() - start), file=sys.stderr, flush=True)
    res = [srv.wait(i) for i in ids]
    wall = time.perf_counter() - start
    bad = [r["error"] for r in res if not r["ok"]]
    if bad:
        raise RuntimeError("serving runs failed: %s" % bad[:3])
    sweeps = np.concatenate([np.asarray(r["sweep_ms"], dtype=np.float64) for r in res])
    fit_calls = sum(r["fit_calls"] for r in res)
    st = srv.stats()
    lat = [r["finished_s"] - r["submitted_s"] for r in res]
    line = {
        "metric": "serving mix: aggregate fitness evals/s; p50/p99 step latency",
        "value": fit_calls / wall,
        "unit": "fitness evals/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64" if args.serving_precision == "fp64" else "f32+f64refine",
        "data": "synthetic",
        "config": {"workload": "serving mix: %d runs, Poisson arrivals 4/s, n log-uniform [1e3,1e6], d in [2,20], "
                               "5-20-node targets + 1%% noise, 1/4 classification, %d fit calls per run (the "
                               "reference's default max_fit_calls)" % (args.runs, args.run_fit_calls),
                   "seed": 31337, "max_active": args.max_active},
        "p50_step_ms": float(np.percentile(sweeps, 50)),
        "p99_step_ms": float(np.percentile(sweeps, 99)),
        "steps": int(len(sweeps)),
        "wall_s": wall,
        "fit_calls": fit_calls,
        "run_latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "dispatcher": st,
        "mean_runs_per_launch": st["requests"] / max(st["launches"], 1),
        "gpu_busy_share": st["kernel_ms"] / 1e3 / wall,
        "gather_us": args.gather_us,
        "host": {"workload_gen_s": t_gen},
    }
    print(json.dumps(line))


def precision_main(args):
    """16384 trees of 5-50 nodes over n = 1e6 x 20 (seeds: trees 4242, data
    2424): FP32 form timed, FP64 form once for the error report."""
    import rils_rols_cpp
    n, d, ntrees = args.precision_rows, 20, args.precision_trees
    X, y, trees = precision_workload(n, ntrees)
    runner = rils_rols_cpp.BatchRunner(np.ascontiguousarray(X.T), y, 0)
    nf = np.array(runner.prepare_ols(trees))
    total, lat, kms, coefs32, fit32, n64, f32, _ = timed_steps(runner, runner, "fp32", args.warmup, args.steps)
    own = f32["fit"]
    t64 = time.perf_counter()
    coefs64, fit64 = runner.run("fp64")
    t64 = time.perf_counter() - t64
    ms64 = runner.last_kernel_ms()
    fl32, fl64 = own[:, 4].astype(int), fit64[:, 4].astype(int)
    ok = ((fl64 & 1) == 0) & ((fl32 & 3) == 0)
    cs = runner.cand_start()
    cerr = []
    for c in np.nonzero(ok)[0]:
        o = cs[c] + c
        a, b = coefs64[o:o + nf[c] + 1], coefs32[o:o + nf[c] + 1]
        cerr.append(float(np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-300))))
    r2a, r2b = 1 - fit64[ok, 0], 1 - own[ok, 0]
    rm = np.abs(fit64[ok, 1] - own[ok, 1]) / np.maximum(fit64[ok, 1], 1e-300)

    def stats(v):
        v = np.asarray(v)
        return {"max": float(np.max(v)), "median": float(np.median(v))} if len(v) else None
    line = {
        "metric": "precision path: FP32+refine fitness evals/s at n=1e6, errors vs the FP64 form",
        "value": ntrees * args.steps / total, "unit": "fitness evals/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32+f64refine", "data": "synthetic",
        "config": {"workload": "precision path: %d trees of 5-50 nodes, n=%d x %d, FP32 evaluation, FP64 solve + "
                               "refinement" % (ntrees, n, d), "seeds": {"trees": 4242, "data": 2424}},
        "ms_per_step": 1e3 * total / args.steps, "steps": args.steps,
        "fp32_kernel_ms": float(np.mean(kms)), "fp64_launch_ms": ms64, "fp64_evals_per_s": ntrees / t64,
        "needs_fp64_frac": float(np.mean((fl32 & 2) > 0)), "sentinel_frac": float(np.mean(fl64 & 1)),
        "certified": int(ok.sum()),
        "errors_vs_fp64": {"coef_rel": stats(cerr), "rmse_rel": stats(rm), "r2_abs": stats(np.abs(r2a - r2b))},
    }
    print(json.dumps(line))


def search_main(args):
    """Full RILS-ROLS search at scale: n = 1e6 x 5 uniform in [-3, 3], target a
    seeded 15-node tree + 1% noise, max_complexity 30; the CPU oracle runs the
    same search for a short prefix of fit calls and its rate is extrapolated."""
    import rils_rols_cpp
    n, d = args.search_rows, 5
    X, y, tgt = search_workload(n)
    # verbose: the search prints its state every 10000 fit calls (rils_rols_cpp.cpp:563-579) on stdout;
    # the JSON line stays the last line
    rr = rils_rols_cpp.rils_rols(False, args.search_fit_calls, 10 ** 7, 0.001, 30, 1.0, True, 12345)
    rr.set_precision(args.search_precision)
    rr.set_context_option("graph", args.graph)
    t = time.perf_counter()
    rr.fit(np.ascontiguousarray(X).reshape(-1), y, n, d)
    wall = time.perf_counter() - t
    pred = rr.predict(np.ascontiguousarray(X).reshape(-1), n, d)
    r2 = 1 - np.sum((y - pred) ** 2) / np.sum((y - y.mean()) ** 2)
    st = rr.get_stats()
    cpu = None
    if not args.no_cpu:
        import oracle_lib as O
        O.set_qr_direct(True)   # timed with Eigen's QR order on the data: the reference's algorithm and cost
        try:
            t = time.perf_counter()
            want = O.fit(X, y, max_fit_calls=args.search_cpu_calls, max_time=10 ** 7, random_state=12345,
                         max_complexity=30)
            cpu_s = time.perf_counter() - t
        finally:
            O.set_qr_direct(False)
        cpu = {"fit_calls": want["fit_calls"], "seconds": cpu_s, "evals_per_s": want["fit_calls"] / cpu_s, "cores": 1,
               "extrapolated_full_run_s": args.search_fit_calls / (want["fit_calls"] / cpu_s),
               "qr": "Eigen's order on the data"}
    line = {
        "metric": "full search at scale: wall time and final R2 (n=1e6 x 5)",
        "value": rr.get_fit_calls() / wall, "unit": "fitness evals/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64" if args.search_precision == "fp64" else "f32+f64refine", "data": "synthetic",
        "config": {"workload": "full search: n=%d x %d uniform [-3,3], 15-node target + 1%% noise, max_fit_calls=%d, "
                               "max_complexity=30" % (n, d, args.search_fit_calls), "seeds": {"data": 777, "search": 12345},
                   "target": " ".join(tgt)},
        "wall_s": wall, "fit_calls": rr.get_fit_calls(), "final_r2": float(r2), "model": rr.get_model_string(),
        "p50_step_ms": float(np.percentile(rr.get_sweep_ms(), 50)) if len(rr.get_sweep_ms()) else None,
        "p99_step_ms": float(np.percentile(rr.get_sweep_ms(), 99)) if len(rr.get_sweep_ms()) else None,
        "stats": st, "cpu_prefix": cpu,
    }
    print(json.dumps(line))


def small_main(args):
    """The reference's own small case (test_example recipe, committed as
    tests/golden/config1_train.csv): 150 rows x 2 features, max_fit_calls
    100000, random_state 12345.  The GPU search and the CPU oracle's search
    run to the end; their move sequences must be identical."""
    import rils_rols_cpp
    import oracle_lib as O
    d = np.loadtxt(os.path.join(ROOT, "tests", "golden", "config1_train.csv"), delimiter=",", skiprows=1)
    X, y = d[:, :2], d[:, 2]
    n, f = X.shape
    rr = rils_rols_cpp.rils_rols(False, 100000, 10 ** 7, 0.001, 50, 1.0, False, 12345)
    rr.set_context_option("graph", args.graph)
    t = time.perf_counter()
    rr.fit(np.ascontiguousarray(X).reshape(-1), y, n, f)
    gpu_s = time.perf_counter() - t
    stats_cold, moves_cold = rr.get_stats(), [tuple(m) for m in rr.get_moves()]
    # the same fit again on the same object: the device context and its work buffers already exist
    # (the first fit's set-up includes the process's HIP start-up); the dataset is uploaded again
    t = time.perf_counter()
    rr.fit(np.ascontiguousarray(X).reshape(-1), y, n, f)
    warm_s = time.perf_counter() - t
    line = {
        "metric": "plain method, small: full search wall time (150 x 2, 100000 fit calls)",
        "value": rr.get_fit_calls() / gpu_s, "unit": "fitness evals/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64", "data": "reference test_example recipe (committed fixture)",
        "config": {"workload": "plain method, small: 150 x 2, max_fit_calls=100000, random_state=12345",
                   "graph": args.graph},
        "gpu_wall_s": gpu_s, "fit_calls": rr.get_fit_calls(), "model": rr.get_model_string(), "stats": stats_cold,
        "warm": {"gpu_wall_s": warm_s, "stats": rr.get_stats(),
                 "same_moves": [tuple(m) for m in rr.get_moves()] == moves_cold},
    }
    if not args.no_cpu:
        O.set_qr_direct(True)   # timed with Eigen's QR order on the data: the reference's algorithm and cost
        try:
            t = time.perf_counter()
            timed = O.fit(X, y, max_fit_calls=100000, max_time=10 ** 7, random_state=12345)
            cpu_s = time.perf_counter() - t
        finally:
            O.set_qr_direct(False)
        want = O.fit(X, y, max_fit_calls=100000, max_time=10 ** 7, random_state=12345)   # the canonical order
        line["cpu_oracle"] = {"wall_s": cpu_s, "cores": 1, "evals_per_s": timed["fit_calls"] / cpu_s,
                              "qr": "Eigen's order on the data"}
        line["moves_identical"] = moves_cold == want["moves"]
        line["speedup_vs_cpu"] = cpu_s / gpu_s
        line["speedup_vs_cpu_warm"] = cpu_s / warm_s
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "fp32"])
    ap.add_argument("--fp32-form", default="checked", choices=["checked", "split"],
                    help="checked: one RR_FP32_CHECKED call per step (FP64 for the candidates it needs, in the "
                         "same call); split: an RR_FP32_REFINED call, then an RR_FP64 call for the flagged ones")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--fp64-eval-rows", type=int, default=None, help="samples per thread of the FP64 evaluation (1, 2, 4)")
    ap.add_argument("--fp64-eval-form", type=int, default=None,
                    help="FP64 factor interpreter: 0 stack in registers, 1 stack in LDS with inlined math")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1],
                    help="search workloads: replay each plain FP64 fitness call as an executable graph")
    ap.add_argument("--fp32-rows", type=int, default=None, help="samples per lane per tile, FP32 kernels (1, 2)")
    ap.add_argument("--fp32-chunk", type=int, default=None, help="samples per FP32 workgroup (multiple of 512)")
    ap.add_argument("--fp32-streams", type=int, default=None, help="1: FP32 buckets on concurrent streams")
    ap.add_argument("--host-threads", type=int, default=None, help="host threads that check a large batch (1..64)")
    ap.add_argument("--pipeline", type=int, default=1, choices=[0, 1],
                    help="batched workload, plain FP64 form: time the steps pipelined (submit the next step "
                         "before waiting for the current one); the one-call-at-a-time figures stay beside them")
    ap.add_argument("--contexts", type=int, default=1,
                    help="batched workload: spread the timed steps over this many contexts, one host thread each")
    ap.add_argument("--slab-budget-mb", type=int, default=None,
                    help="device memory of one FP64 launch's work columns (candidates are launched in chunks that fit)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default kernel schedule instead of the autotuner's pick")
    ap.add_argument("--copy-probe", action="store_true", help="run only a plain 1 GiB device copy (PMC scaling)")
    ap.add_argument("--timed-only", action="store_true",
                    help="launch nothing on the GPU beyond the warmup and the one-call-at-a-time timed steps "
                         "(for kernel traces and counters: no pipelined leg, no extra untimed calls)")
    ap.add_argument("--workload", default="batched", choices=["batched", "serving", "precision", "search", "small"],
                    help="batched: the headline batched candidate eval; serving: the serving mix; precision: 16384 "
                         "trees at n=1e6 x 20 in FP32; search: a full search at n=1e6 x 5")
    ap.add_argument("--precision-rows", type=int, default=1_000_000)
    ap.add_argument("--precision-trees", type=int, default=16384)
    ap.add_argument("--search-rows", type=int, default=1_000_000)
    ap.add_argument("--search-fit-calls", type=int, default=1_000_000)
    ap.add_argument("--search-cpu-calls", type=int, default=300)
    ap.add_argument("--search-precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--runs", type=int, default=64, help="serving: number of runs")
    ap.add_argument("--run-fit-calls", type=int, default=100000,
                    help="serving: fit-call budget per run (default: the reference's max_fit_calls default)")
    ap.add_argument("--gather-us", type=int, default=300,
                    help="serving: how long a launch waits for the sweeps of other searching runs (microseconds)")
    ap.add_argument("--max-active", type=int, default=64, help="serving: resident runs at most")
    ap.add_argument("--serving-precision", default="fp64", choices=["fp64", "fp32"])
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: the fitness path would shard across GPUs by slices of each candidate "
              "batch, which this program does not do", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serving":
        serving_main(args)
        return
    if args.workload == "precision":
        precision_main(args)
        return
    if args.workload == "search":
        search_main(args)
        return
    if args.workload == "small":
        small_main(args)
        return
    batched_main(args)


def factor_reads_variable(prefix):
    return any(t.startswith("x") for t in prefix.split())


def executed_flops(factors, ks, E, flags, n, subset=None):
    """Algorithmic flops F = n (E + k^2 + 5k + 3) (the reference's
    tune_constants + fitness, rils_rols_cpp.cpp:445-541: E operator nodes
    over the factors, k columns, each operator counting 1) over the candidates a
    launch evaluated in full; a candidate the factor-0 pre-pass settled
    (RR_F_EARLY) is credited only factor 0's operators on the pre-pass's 64
    rows."""
    total = 0.0
    for c in (range(len(factors)) if subset is None else subset):
        if int(flags[c]) & RR_F_EARLY:
            total += 64.0 * (ops_in(factors[c][0]) if factors[c] else 0)
        else:
            total += float(n) * (E[c] + ks[c] * ks[c] + 5 * ks[c] + 3)
    return total


def eval_kernel_flops(factors, flags, n):
    """Operators the FP64 evaluation kernel runs: every distinct non-bare factor
    of a candidate the pre-pass did not settle, once over all n rows (equal
    factors share one evaluation)."""
    uniq = set()
    for c, fs in enumerate(factors):
        if not (int(flags[c]) & RR_F_EARLY):
            uniq.update(f for f in fs if not BARE.match(f))
    return float(n) * sum(ops_in(f) for f in uniq), len(uniq)


def batched_main(args):
    import rils_rols_cpp
    t0 = time.perf_counter()
    X, y, target, trees, factors = workload()
    t_gen = time.perf_counter() - t0
    runner = rils_rols_cpp.BatchRunner(np.ascontiguousarray(X.T), y, 0)
    if args.fp64_eval_rows:
        runner.set_option("fp64_eval_rows", args.fp64_eval_rows)
    if args.fp64_eval_form is not None:
        runner.set_option("fp64_eval_form", args.fp64_eval_form)
    if args.fp32_rows:
        runner.set_option("fp32_rows", args.fp32_rows)
    if args.fp32_chunk:
        runner.set_option("fp32_chunk_rows", args.fp32_chunk)
    if args.fp32_streams is not None:
        runner.set_option("fp32_streams", args.fp32_streams)
    if args.slab_budget_mb:
        runner.set_option("slab_budget_mb", args.slab_budget_mb)
    if args.host_threads:
        runner.set_option("host_threads", args.host_threads)
    t0 = time.perf_counter()
    nf = runner.prepare_ols(trees)
    t_prep = time.perf_counter() - t0
    ks = np.array(nf) + 1
    E = np.array([sum(ops_in(f) for f in fs) for fs in factors])
    algo_bytes = 2.0 * N_ROWS * (N_FEAT + 1) * 8   # passes x n x (d_used + 1) x 8 B

    res = {}
    tuned = {}
    for prec in (["fp64", "fp32"] if args.precision == "both" else [args.precision]):
        form = "fp32c" if prec == "fp32" and args.fp32_form == "checked" else prec
        if not args.no_autotune:
            # the schedule (rows per thread, FP32 chunk size, bucket streams) is picked by the
            # library's autotuner on this batch, outside the timed region
            tuned[prec] = runner.autotune(form, 2)
            print("[bench] autotune %s: %s" % (form, tuned[prec]), file=sys.stderr, flush=True)
        runner.set_option("phase_timing", 1)
        total, lat, kms, coefs, fit, n64, f32, phases = timed_steps(runner, runner, form, args.warmup, args.steps)
        runner.set_option("phase_timing", 0)
        if form == "fp32c" and not args.timed_only:
            # the FP32 kernels' own flags and bounds, from one untimed call of the unchecked form
            f32["fit"] = runner.run("fp32")[1]
        elif form == "fp32c":
            f32["fit"] = fit
        res[prec] = dict(total=total, lat=lat, kms=kms, coefs=coefs, fit=fit, n64=n64, f32=f32, phases=phases)
        if form == "fp64" and args.pipeline and not args.timed_only:
            # the same steps pipelined on one context: step i + 1 is checked, planned and launched
            # while step i runs (rr_fitness_submit / rr_fitness_wait)
            if args.warmup:
                runner.run_pipelined(args.warmup)
            _, _, plat, ptot = runner.run_pipelined(args.steps)
            res[prec]["multi"] = (ptot, np.array(plat))
            res[prec]["overlap"] = "pipelined: each step submitted before the previous one is waited for"
            print("[bench] %s pipelined: %.3f ms/step wall, %.0f evals/s" %
                  (prec, 1e3 * ptot / args.steps, N_TREES * args.steps / ptot), file=sys.stderr, flush=True)
        elif args.contexts > 1 and form != "fp32":
            # the same steps over several contexts driven by their own host threads
            others = []
            for _ in range(args.contexts - 1):
                o = rils_rols_cpp.BatchRunner(np.ascontiguousarray(X.T), y, 0)
                for k in ("fp64_eval_rows", "fp64_eval_form", "fp32_rows", "fp32_chunk_rows", "fp32_streams",
                          "slab_budget_mb", "host_threads"):
                    o.set_option(k, runner.get_option(k))
                o.prepare_ols(trees)
                others.append(o)
            res[prec]["multi"] = timed_steps_multi([runner] + others, form, args.warmup, args.steps)
            res[prec]["overlap"] = "%d contexts, one host thread each" % args.contexts
            print("[bench] %s, %d contexts: %.3f ms/step wall, %.0f evals/s" %
                  (prec, args.contexts, 1e3 * res[prec]["multi"][0] / args.steps,
                   N_TREES * args.steps / res[prec]["multi"][0]), file=sys.stderr, flush=True)
            del others
        print("[bench] %s: %.3f ms/step wall, %.3f ms kernel, %.0f evals/s, phases %s" %
              (prec, 1e3 * total / args.steps, kms.mean(), N_TREES * args.steps / total,
               {k: round(v, 4) for k, v in phases.items()}), file=sys.stderr, flush=True)

    out = {}
    if "fp64" in res:
        r = res["fp64"]
        flags = r["fit"][:, 4].astype(np.int64)
        ph = r["phases"]
        eval_flops, n_eval = eval_kernel_flops(factors, flags, N_ROWS)
        step_flops = executed_flops(factors, ks, E, flags, N_ROWS)
        kernel_s = r["kms"].mean() / 1e3
        traffic = pmc_traffic("fp64_step")
        eval_ach = eval_flops / (ph["eval"] / 1e3) / 1e12
        out["fp64"] = {
            "evals_per_s": N_TREES * args.steps / r["total"],
            "ms_per_step": 1e3 * r["total"] / args.steps,
            "kernel_ms": float(r["kms"].mean()),
            "phase_ms": ph,
            "p50_step_ms": float(np.percentile(r["lat"], 50) * 1e3),
            "p99_step_ms": float(np.percentile(r["lat"], 99) * 1e3),
            # the dominant kernel: the factor evaluation (its share of the step in phase_ms)
            "roofline": {"bound": "valu", "achieved": eval_ach, "peak": PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s",
                         "frac": eval_ach / PEAK_FP64_VALU_TFLOPS,
                         "traffic": traffic,
                         "kernel": "k_alive_fp64 + k_eval_fp64<%d, %d>" % (runner.get_option("fp64_eval_rows"),
                                                                          runner.get_option("fp64_eval_form")),
                         "launch_ms": ph["eval"], "algorithmic_flops": eval_flops, "factors_evaluated": n_eval,
                         "counters": pmc_counters("fp64_step")},
            "step_roofline": {"bound": "valu", "achieved": step_flops / kernel_s / 1e12,
                              "peak": PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s",
                              "frac": step_flops / kernel_s / 1e12 / PEAK_FP64_VALU_TFLOPS,
                              "algorithmic_flops": step_flops, "launch_ms": float(r["kms"].mean()),
                              "algorithmic_bytes": algo_bytes, "traffic": traffic,
                              "traffic_over_algorithmic_bytes": traffic / algo_bytes if traffic else None},
            "early_settled": int(np.count_nonzero(flags & RR_F_EARLY)),
            "sentinel_frac": float(np.mean(flags & 1)),
        }
    if "fp32" in res:
        r = res["fp32"]
        own = r["f32"]["fit"]                     # the FP32 kernels' own results, before FP64 re-scores
        fl32 = own[:, 4].astype(int)
        ph = r["phases"]
        in32 = [c for c in range(N_TREES) if len(factors[c]) <= 8 and all(factor_reads_variable(f) for f in factors[c])]
        f32_flops = executed_flops(factors, ks, E, fl32, N_ROWS, in32)
        ach = f32_flops / (ph["fp32"] / 1e3) / 1e12 if ph["fp32"] > 0 else 0.0
        reasons = np.bincount(np.where((fl32 & 2) > 0, -own[:, 3].astype(int), 0), minlength=9).tolist()
        out["fp32"] = {
            "evals_per_s": N_TREES * args.steps / r["total"],
            "ms_per_step": 1e3 * r["total"] / args.steps,
            "kernel_ms": float(r["kms"].mean()),
            "phase_ms": ph,
            "p50_step_ms": float(np.percentile(r["lat"], 50) * 1e3),
            "p99_step_ms": float(np.percentile(r["lat"], 99) * 1e3),
            "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP32_TFLOPS, "traffic": pmc_traffic("fp32_step"),
                         "kernel": "k_acc_fp32 + k_fin_fp32 (all factor-count buckets)", "launch_ms": ph["fp32"],
                         "candidates": len(in32), "algorithmic_flops": f32_flops},
            "needs_fp64_frac": float(np.mean((fl32 & 2) > 0)),
            "needs_fp64_reasons": reasons,
            "form": args.fp32_form,
            "fp64_results_per_step": int(r["n64"]),
        }
        if "fp64" in res:
            a, b = res["fp64"]["fit"], r["fit"]              # final results of the FP32 form's step
            live = (a[:, 4].astype(int) & 1) == 0
            d = np.abs(a[live, 0] - b[live, 0])
            certified = live & ((b[:, 4].astype(int) & 4) == 0)
            out["fp32"]["certified"] = int(certified.sum())
            out["fp32"]["r2_abs_err_max"] = float(d.max()) if d.size else None
            out["fp32"]["r2_abs_err_max_certified"] = \
                float(np.max(np.abs(a[certified, 0] - b[certified, 0]))) if certified.any() else None
            out["fp32"]["sentinels_agree"] = bool(np.array_equal(a[:, 4].astype(int) & 1, b[:, 4].astype(int) & 1))

    for prec, o in out.items():
        if "multi" in res[prec]:
            # the step loop over several contexts: its wall time and per-call latencies are the form's
            # figures, the single-context loop's beside them
            tot, lat = res[prec]["multi"]
            o["one_call_at_a_time"] = {k: o[k] for k in ("evals_per_s", "ms_per_step", "p50_step_ms", "p99_step_ms")}
            o.update(evals_per_s=N_TREES * args.steps / tot, ms_per_step=1e3 * tot / args.steps,
                     p50_step_ms=float(np.percentile(lat, 50) * 1e3),
                     p99_step_ms=float(np.percentile(lat, 99) * 1e3), overlap=res[prec]["overlap"])

    cpu = None
    cpu1 = None
    check = None
    if not args.no_cpu:
        nt = host_threads()
        v, count, secs = cpu_baseline(X, y, trees, args.cpu_seconds, threads=nt)
        cpu = {"value": v, "unit": "fitness evals/s", "cores": nt, "kind": "port",
               "sample": "first %d of the 4096 candidates at full n=1e5, oracle tune_constants+fitness with "
                         "Eigen's QR order on the data, %d threads over candidates, %.1f s" % (count, nt, secs)}
        v1, count1, secs1 = cpu_baseline(X, y, trees, args.cpu_seconds / 2, threads=1)
        cpu1 = {"value": v1, "cores": 1, "sample": "first %d candidates, single thread, %.1f s" % (count1, secs1)}
        if "fp64" in res:
            # the timed FP64 results against the oracle (canonical order) on a sample, bit for bit
            count = min(count, 1024)
            cfit = oracle_fit(X, y, trees[:count], nt)
            g = res["fp64"]["fit"][:count]
            same = [np.float64(g[c, 0]).tobytes() == np.float64(cfit[c, 0]).tobytes() and
                    np.float64(g[c, 1]).tobytes() == np.float64(cfit[c, 1]).tobytes() for c in range(count)]
            check = {"candidates": count, "fitness_bit_identical": int(sum(same))}

    # the headline: the form that delivers trusted results fastest (FP64: bit-exact; the FP32 form:
    # certified |dR2| <= 1e-9, everything else re-scored in FP64)
    head = max(out, key=lambda k: out[k]["evals_per_s"])
    h = out[head]
    line = {
        "metric": "fitness evals/s at n=1e5 (batched FP32+refine vs plain FP64); p99 step latency",
        "value": h["evals_per_s"],
        "unit": "fitness evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64refine" if head == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": "batched candidate eval: 4096 random trees (3-50 nodes, <=8 additive terms), "
                               "n=100000 x 10 features, OLS fit per tree", "n": N_ROWS, "features": N_FEAT,
                   "candidates": N_TREES, "seeds": {"trees": 12345, "data": 23654}, "target": target,
                   "form": head},
        "p50_step_ms": h["p50_step_ms"],
        "p99_step_ms": h["p99_step_ms"],
        "roofline": h["roofline"],
        "cpu_baseline": cpu,
        "cpu_single_thread": cpu1,
        "fp64_vs_oracle": check,
        "ratios": {
            "fp32_over_fp64": (out["fp32"]["evals_per_s"] / out["fp64"]["evals_per_s"])
            if "fp32" in out and "fp64" in out else None,
            "over_cpu_all_cores": h["evals_per_s"] / cpu["value"] if cpu else None,
            "over_cpu_one_core": h["evals_per_s"] / cpu1["value"] if cpu1 else None,
        },
        "forms": out,
        "schedule": {"autotuned": tuned,
                     "fp64_eval_rows": runner.get_option("fp64_eval_rows"),
                     "fp64_eval_form": runner.get_option("fp64_eval_form"),
                     "fp32_rows": runner.get_option("fp32_rows"),
                     "fp32_chunk_rows": runner.get_option("fp32_chunk_rows"),
                     "fp32_streams": runner.get_option("fp32_streams")},
        "host": {"workload_gen_s": t_gen, "host_prepare_s": t_prep, "mean_k": float(ks.mean()),
                 "mean_ops": float(E.mean())},
    }
    print(json.dumps(line))

if __name__ == "__main__":
    main()
