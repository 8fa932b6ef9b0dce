"""Serving loop on the GPU: many runs resident, their sweeps merged into shared
launches as they arrive.  Each run's result must be the same alone and in the
mix: same moves (fit-call count, fitness bits, model at every move), same
final model and fit calls."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BIG_TIME = 10 ** 8


def problems():
    out = []
    for i, (n, d, budget, cls, prec) in enumerate([(300, 2, 4000, False, "fp64"), (2500, 3, 3000, False, "fp64"),
                                                   (1200, 4, 3000, True, "fp64"), (800, 2, 3000, False, "fp32"),
                                                   (5000, 3, 2500, False, "fp64"), (600, 5, 3000, False, "fp64")]):
        rng = np.random.default_rng(100 + i)
        X = rng.uniform(-3, 3, (n, d))
        y = np.sin(X[:, 0]) * X[:, d - 1] + 0.5 * X[:, 0] ** 2 + rng.normal(0, 0.01, n)
        if cls:
            y = (y > np.median(y)).astype(np.float64)
        out.append(dict(X=X, y=y, classification=cls, max_fit_calls=budget, random_state=7 + i, precision=prec))
    return out


def solo(p):
    import rils_rols_cpp
    n, d = p["X"].shape
    rr = rils_rols_cpp.rils_rols(p["classification"], p["max_fit_calls"], BIG_TIME, 0.001, 50, 1.0, False,
                                 p["random_state"])
    rr.set_precision(p["precision"])
    rr.fit(np.ascontiguousarray(p["X"]).reshape(-1), p["y"], n, d)
    return [tuple(m) for m in rr.get_moves()], rr.get_model_string(), rr.get_fit_calls()


@pytest.mark.parametrize("max_active,budget_mb", [(8, 0), (2, 1)])
def test_each_run_same_alone_and_in_the_mix(max_active, budget_mb):
    import rils_rols_cpp
    print("case: serving mix max_active=%d hbm_budget_mb=%d" % (max_active, budget_mb), flush=True)
    ps = problems()
    want = [solo(p) for p in ps]
    srv = rils_rols_cpp.Server(0, budget_mb, max_active, 1, 0)
    ids = [srv.submit(p["X"], p["y"], classification=p["classification"], max_fit_calls=p["max_fit_calls"],
                      max_seconds=BIG_TIME, random_state=p["random_state"], precision=p["precision"]) for p in ps]
    got = [srv.wait(i) for i in ids]
    st = srv.stats()
    print("dispatcher:", st, flush=True)
    for g, (moves, model, fc), p in zip(got, want, ps):
        assert g["ok"], g["error"]
        assert [tuple(m) for m in g["moves"]] == moves
        assert g["model"] == model and g["fit_calls"] == fc
    if max_active > 1 and budget_mb == 0:
        assert st["max_requests_per_launch"] > 1     # sweeps of different runs did share launches
    assert srv.resident_bytes() == 0


def test_admission_by_footprint_serialises_runs():
    """With a budget below two datasets, runs are admitted one at a time, in
    submission order."""
    import rils_rols_cpp
    ps = problems()[:3]
    fp = [rils_rols_cpp.Server.footprint(p["X"].shape[0], p["X"].shape[1]) for p in ps]
    budget_mb = max(1, (max(fp) + (1 << 20) - 1) >> 20)
    srv = rils_rols_cpp.Server(0, budget_mb, 64, 1, 0)
    ids = [srv.submit(p["X"], p["y"], classification=p["classification"], max_fit_calls=1500, max_seconds=BIG_TIME,
                      random_state=p["random_state"]) for p in ps]
    got = [srv.wait(i) for i in ids]
    for a, b in zip(got, got[1:]):
        if a["footprint"] + b["footprint"] > budget_mb << 20:
            assert b["admitted_s"] >= a["finished_s"] - 1e-6
