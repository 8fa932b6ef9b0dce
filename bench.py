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
precision form (FP32 evaluation, compensated Gram sums, FP64 solve + one
refinement step).  `value` is the FP32 form's evals/s.

Prints one JSON line.  --gpus N with N > 1 exits: this build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "rils-rols_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

N_ROWS, N_FEAT, N_TREES = 100_000, 10, 4096
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


def fp64_algorithmic_bytes(k, steps, sentinel, n):
    """HBM bytes the FP64 OLS kernel must move for one candidate with k columns
    (factors + free term) whose pivoted QR ran `steps` Householder steps: the
    evaluation pass writes k+1 columns (A and the copy of y) plus a kept copy
    of the k-1 factor columns and reads y;
    QR step p reads the pivot column and the k-p remaining targets (k-p-1
    columns of A plus y) in its dot pass and writes the Householder vector,
    then reads the vector and the k-p targets and writes the targets back;
    the score pass (skipped for sentinels) reads y and the kept factor
    columns.  Features are read from
    cache-resident X and not counted."""
    words = (k + 1) + (k - 1) + 1
    for p in range(steps):
        t = k - p
        words += (1 + t + 1) + (1 + 2 * t)
    if not sentinel:
        words += 1 + (k - 1)
    return 8.0 * n * words


def timed_steps(runner, rescorer, precision, warmup, steps):
    """K steps of one batched call.  In the FP32 form a step also re-scores in
    FP64 every candidate the FP32 kernels could not certify (RR_F_NEEDS64),
    so each step delivers a trusted fitness for all 4096 candidates."""
    def one():
        coefs, fit = runner.run(precision)
        ms = runner.last_kernel_ms()
        n64 = 0
        if precision == "fp32":
            need = np.nonzero(fit[:, 4].astype(np.int64) & 2)[0]
            n64 = len(need)
            if n64:
                rescorer.select(need)
                c2, f2 = rescorer.run_selected("fp64")
                ms += rescorer.last_kernel_ms()
                fit[need] = f2
        return coefs, fit, ms, n64
    for _ in range(warmup):
        one()
    lat, kms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        coefs, fit, ms, n64 = one()
        lat.append(time.perf_counter() - s)
        kms.append(ms)
    total = time.perf_counter() - t0
    return total, np.array(lat), np.array(kms), coefs, fit, n64


def cpu_baseline(X, y, trees, seconds_target=15.0):
    """The oracle (CPU restatement, g++ -O3) on a bounded sample of the same
    workload, single thread: tune_constants + fitness per candidate."""
    import oracle_lib as O
    od = O.Dataset(X, y)
    probe = max(4, 8)
    secs, _ = od.batch_tune(trees[:probe], threads=1)
    per = secs / probe
    count = int(min(len(trees), max(probe, seconds_target / max(per, 1e-6))))
    secs, fit = od.batch_tune(trees[:count], threads=1)
    return count / secs, count, secs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "fp32"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--fp64-rows", type=int, default=None, help="samples per lane per tile, FP64 kernels (1, 2, 4)")
    ap.add_argument("--fp32-rows", type=int, default=None, help="samples per lane per tile, FP32 kernels (1, 2)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: collectives and multi-GPU sharding are not part of this build "
              "(see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)

    import rils_rols_cpp
    t0 = time.perf_counter()
    X, y, target, trees, factors = workload()
    t_gen = time.perf_counter() - t0
    runner = rils_rols_cpp.BatchRunner(np.ascontiguousarray(X.T), y, 0)
    if args.fp64_rows:
        runner.set_option("fp64_rows", args.fp64_rows)
    if args.fp32_rows:
        runner.set_option("fp32_rows", args.fp32_rows)
    t0 = time.perf_counter()
    nf = runner.prepare_ols(trees)
    t_prep = time.perf_counter() - t0
    ks = np.array(nf) + 1
    E = np.array([sum(ops_in(f) for f in fs) for fs in factors])
    flops = float(np.sum(N_ROWS * (E + ks * ks + 5 * ks + 3)))

    res = {}
    for prec in (["fp64", "fp32"] if args.precision == "both" else [args.precision]):
        total, lat, kms, coefs, fit, n64 = timed_steps(runner, runner, prec, args.warmup, args.steps)
        res[prec] = dict(total=total, lat=lat, kms=kms, coefs=coefs, fit=fit, n64=n64)
        print("[bench] %s: %.3f ms/step wall, %.3f ms kernel, %.0f evals/s" %
              (prec, 1e3 * total / args.steps, kms.mean(), N_TREES * args.steps / total), file=sys.stderr, flush=True)

    out = {}
    if "fp64" in res:
        r = res["fp64"]
        flags = r["fit"][:, 4].astype(np.int64)
        steps = flags >> 16
        byts = float(sum(fp64_algorithmic_bytes(int(k), int(st), bool(fl & 1), N_ROWS)
                         for k, st, fl in zip(ks, steps, flags)))
        kernel_s = r["kms"].mean() / 1e3
        out["fp64"] = {
            "evals_per_s": N_TREES * args.steps / r["total"],
            "ms_per_step": 1e3 * r["total"] / args.steps,
            "kernel_ms": float(r["kms"].mean()),
            "p50_step_ms": float(np.percentile(r["lat"], 50) * 1e3),
            "p99_step_ms": float(np.percentile(r["lat"], 99) * 1e3),
            "roofline": {"bound": "hbm", "achieved": byts / kernel_s / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": byts / kernel_s / 1e9 / PEAK_HBM_GBS, "traffic": None,
                         "algorithmic_bytes": byts},
            "valu_tflops": flops / kernel_s / 1e12,
            "sentinel_frac": float(np.mean(r["fit"][:, 4].astype(int) & 1)),
        }
    if "fp32" in res:
        r = res["fp32"]
        kernel_s = r["kms"].mean() / 1e3
        out["fp32"] = {
            "evals_per_s": N_TREES * args.steps / r["total"],
            "ms_per_step": 1e3 * r["total"] / args.steps,
            "kernel_ms": float(r["kms"].mean()),
            "p50_step_ms": float(np.percentile(r["lat"], 50) * 1e3),
            "p99_step_ms": float(np.percentile(r["lat"], 99) * 1e3),
            "roofline": {"bound": "valu", "achieved": flops / kernel_s / 1e12, "peak": PEAK_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": flops / kernel_s / 1e12 / PEAK_FP32_TFLOPS, "traffic": None},
            "needs_fp64_frac": float(np.mean((r["fit"][:, 4].astype(int) & 2) > 0)),
        }
        if "fp64" in res:
            a, b = res["fp64"]["fit"], r["fit"]
            ok = ((a[:, 4].astype(int) & 1) == 0) & ((b[:, 4].astype(int) & 3) == 0)
            r2a, r2b = 1 - a[ok, 0], 1 - b[ok, 0]
            out["fp32"]["r2_abs_err_max"] = float(np.max(np.abs(r2a - r2b))) if ok.any() else None
            out["fp32"]["r2_abs_err_median"] = float(np.median(np.abs(r2a - r2b))) if ok.any() else None

    cpu = None
    if not args.no_cpu:
        v, count, secs = cpu_baseline(X, y, trees, args.cpu_seconds)
        cpu = {"value": v, "unit": "fitness evals/s", "cores": 1, "kind": "port",
               "sample": "first %d of the 4096 candidates at full n=1e5, oracle tune_constants+fitness, "
                         "single thread, %.1f s" % (count, secs)}

    head = "fp32" if "fp32" in out else "fp64"
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
                   "candidates": N_TREES, "seeds": {"trees": 12345, "data": 23654}, "target": target},
        "p50_step_ms": h["p50_step_ms"],
        "p99_step_ms": h["p99_step_ms"],
        "roofline": h["roofline"],
        "cpu_baseline": cpu,
        "forms": out,
        "host": {"workload_gen_s": t_gen, "host_prepare_s": t_prep, "mean_k": float(ks.mean()),
                 "mean_ops": float(E.mean())},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
