"""The benchmark configurations at their real sizes, checked on the GPU.

- batched candidate eval (4096 trees, n = 1e5 x 10): the whole batch goes
  through the same BatchRunner path the benchmark times, and a 512-tree slice
  of it is compared with the oracle bit for bit (tune_constants + fitness,
  rils_rols_cpp.cpp:445-541);
- precision path (n = 1e6 x 20): 2048 trees in the FP32 form against the FP64
  form (every certified result inside its bound), and 64 FP64 candidates
  against the oracle bit for bit;
- full search (n = 1e6 x 5, max_complexity 30): a 1000-call run of the
  product search accepts the oracle's moves (fit_inner,
  rils_rols_cpp.cpp:772-859);
- serving mix: 12 runs of the mix (two with n > 5e5, four classification)
  submitted on the mix's Poisson schedule; every run ends the same alone and
  in the mix, and runs share launches.
"""
import os

import numpy as np
import pytest

import bench
import capi
import oracle_lib as O

pytestmark = pytest.mark.gpu
BIG_TIME = 10 ** 8


def threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 4


def same_bits(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return bool(np.all((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))))


def test_batched_eval_slice_matches_oracle_at_full_size():
    import rils_rols_cpp
    print("case: batched eval 4096 trees n=1e5 x 10, 512-tree slice vs oracle", flush=True)
    X, y, _, trees, factors = bench.workload()
    runner = rils_rols_cpp.BatchRunner(np.ascontiguousarray(X.T), y, 0)
    nf = runner.prepare_ols(trees)
    coefs, fit = runner.run("fp64")
    cs = runner.cand_start()
    sel = list(range(0, len(trees), 8))
    od = O.Dataset(X, y)
    want = od.batch_tune_full([trees[i] for i in sel], threads=threads())
    live = 0
    for i, w in zip(sel, want):
        assert w["factors"] == list(factors[i]), i
        o = cs[i] + i
        c = coefs[o:o + nf[i] + 1]
        assert same_bits(c, w["coefs"]), (i, trees[i], c, w["coefs"])
        assert int(fit[i, 3]) == w["rank"], i
        assert same_bits(fit[i, :2], w["fit"][:2]), (i, trees[i], fit[i], w["fit"])
        live += int((int(fit[i, 4]) & 1) == 0)
    print("slice: %d candidates, %d not sentinels" % (len(sel), live), flush=True)
    assert live >= 50


def test_precision_path_certificates_and_fp64_at_n_1e6():
    import rils_rols_cpp
    print("case: precision path 2048 trees n=1e6 x 20", flush=True)
    X, y, trees = bench.precision_workload(1_000_000, 2048)
    runner = rils_rols_cpp.BatchRunner(np.ascontiguousarray(X.T), y, 0)
    nf = np.array(runner.prepare_ols(trees))
    c64, f64 = runner.run("fp64")
    c32, f32 = runner.run("fp32")
    fl32, fl64 = f32[:, 4].astype(int), f64[:, 4].astype(int)
    cert = ((fl32 & 3) == 0) & ((fl64 & 1) == 0)
    fa = (1 + f64[cert, 0]) * (1 + f64[cert, 1])
    fb = (1 + f32[cert, 0]) * (1 + f32[cert, 1])
    rel = np.abs(fa - fb) / fa
    print("certified %d of %d; max rel err %.3g; max |dR2| %.3g" %
          (cert.sum(), len(trees), rel.max() if cert.any() else 0,
           np.abs(f64[cert, 0] - f32[cert, 0]).max() if cert.any() else 0), flush=True)
    reasons = np.bincount(-f32[(fl32 & 2) > 0, 3].astype(int), minlength=9)
    print("uncertified by reason:", reasons.tolist(), flush=True)
    assert cert.sum() >= 1
    assert np.all(rel <= f32[cert, 2])                                    # every certificate holds
    assert np.all(np.abs(f64[cert, 0] - f32[cert, 0]) <= 1e-9)           # R2 within 1e-9 of FP64
    assert np.all(((fl64 & 1) > 0) >= ((fl32 & 1) > 0))                   # an FP32 sentinel is an FP64 sentinel
    # the checked form's final results: FP64's bits where it took FP64, R2 within 1e-9 everywhere
    cc, fc = runner.run("fp32c")
    flc = fc[:, 4].astype(int)
    assert not np.any(flc & 2)
    took = (flc & 4) > 0
    assert same_bits(fc[took, :2], f64[took, :2])
    fine = ~took & ((flc & 1) == 0)
    assert np.all(np.abs(fc[fine, 0] - f64[fine, 0]) <= 1e-9)
    assert np.all((flc[~took] & 1) == (fl64[~took] & 1))
    # FP64 against the oracle on 64 candidates: 48 that are not sentinels, 16 that are
    live = [i for i in range(len(trees)) if not (fl64[i] & 1)][:48]
    dead = [i for i in range(len(trees)) if fl64[i] & 1][:16]
    sel = live + dead
    want = O.Dataset(X, y).batch_tune_full([trees[i] for i in sel], threads=threads())
    cs = runner.cand_start()
    for i, w in zip(sel, want):
        o = cs[i] + i
        assert same_bits(c64[o:o + nf[i] + 1], w["coefs"]), (i, trees[i])
        assert same_bits(f64[i, :2], w["fit"][:2]), (i, trees[i], f64[i], w["fit"])


def test_full_search_at_n_1e6_accepts_the_oracles_moves():
    import rils_rols_cpp
    print("case: full search n=1e6 x 5, 1000 fit calls, max_complexity 30", flush=True)
    X, y, _ = bench.search_workload(1_000_000)
    want = O.fit(X, y, max_fit_calls=1000, max_time=BIG_TIME, random_state=12345, max_complexity=30)
    n, d = X.shape
    rr = rils_rols_cpp.rils_rols(False, 1000, BIG_TIME, 0.001, 30, 1.0, False, 12345)
    rr.fit(np.ascontiguousarray(X).reshape(-1), y, n, d)
    moves = [tuple(m) for m in rr.get_moves()]
    print("moves: %d, model %s" % (len(moves), rr.get_model_string()), flush=True)
    assert len(want["moves"]) >= 3
    assert moves == want["moves"]
    assert rr.get_fit_calls() == want["fit_calls"]
    assert rr.get_model_string() == want["model"]


SERVING_RUNS = list(range(10)) + [33, 44]   # runs 0-9 of the mix, and two with n > 5e5


def test_serving_mix_runs_same_alone_and_in_the_mix():
    import time

    import rils_rols_cpp
    print("case: serving mix, 12 runs on the Poisson schedule", flush=True)
    plan = bench.serving_plan()
    runs = bench.serving_workload(keep=set(SERVING_RUNS))
    assert sum(r["classification"] for r in runs) >= 3 and sum(r["X"].shape[0] >= 500_000 for r in runs) >= 2
    arrivals = [p[1] for p in plan[:len(runs)]]   # the mix's first 12 arrival times
    budget = 100000   # the reference's default max_fit_calls (rils_rols.py:156): runs last long enough to meet

    def solo(r):
        n, d = r["X"].shape
        rr = rils_rols_cpp.rils_rols(r["classification"], budget, BIG_TIME, 0.001, 50, 1.0, False, r["seed"])
        rr.fit(np.ascontiguousarray(r["X"]).reshape(-1), r["y"], n, d)
        return [tuple(m) for m in rr.get_moves()], rr.get_model_string(), rr.get_fit_calls()

    want = [solo(r) for r in runs]
    print("solo runs done", flush=True)
    # a 300 us gather window, as in the serving benchmark: a launch waits that long for the other
    # resident runs' requests
    srv = rils_rols_cpp.Server(0, 0, 64, 1, 0, 300)
    start = time.perf_counter()
    ids = []
    for r, t in zip(runs, arrivals):
        wait = t - (time.perf_counter() - start)
        if wait > 0:
            time.sleep(wait)
        ids.append(srv.submit(r["X"], r["y"], classification=r["classification"], max_fit_calls=budget,
                              max_seconds=BIG_TIME, random_state=r["seed"]))
    got = [srv.wait(i) for i in ids]
    st = srv.stats()
    print("dispatcher:", st, flush=True)
    for g, (moves, model, fc), r in zip(got, want, runs):
        assert g["ok"], g["error"]
        assert [tuple(m) for m in g["moves"]] == moves, r["index"]
        assert g["model"] == model and g["fit_calls"] == fc, r["index"]
    assert st["requests"] > st["launches"]          # on average more than one run per launch
    assert st["max_requests_per_launch"] >= 2
