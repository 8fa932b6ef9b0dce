"""End-to-end search: the product (rils_rols_cpp on the GPU) vs the oracle.

For a fixed seed the product must accept the same sequence of moves as the
CPU restatement of the reference's search: every local-search acceptance and
every new best solution, with the fit-call count at which it happened, the
exact fitness bits and the model string; and end with the same model and the
same number of fit calls.  The time budget is set so large that only the
fit-call budget ends a run (the wall-clock part of finished(),
rils_rols_cpp.cpp:659-661, is not reproducible).
"""
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIG_TIME = 10 ** 8


def load_config1(which="train"):
    d = np.loadtxt(os.path.join(HERE, "golden", "config1_%s.csv" % which), delimiter=",", skiprows=1)
    return d[:, :2], d[:, 2]


def run_product(X, y, classification=False, max_fit_calls=100000, complexity_penalty=0.001, max_complexity=50,
                sample_size=1.0, random_state=0, precision="fp64", options=()):
    import rils_rols_cpp
    n, d = X.shape
    rr = rils_rols_cpp.rils_rols(classification, max_fit_calls, BIG_TIME, complexity_penalty, max_complexity,
                                 sample_size, False, random_state)
    rr.set_precision(precision)
    for name, value in options:
        rr.set_context_option(name, value)
    rr.fit(np.ascontiguousarray(X, dtype=np.float64).reshape(-1), np.ascontiguousarray(y, dtype=np.float64), n, d)
    return rr


def assert_same_run(rr, want):
    moves = [tuple(m) for m in rr.get_moves()]
    assert len(moves) == len(want["moves"])
    for got, exp in zip(moves, want["moves"]):
        assert got == exp, (got, exp)
    assert rr.get_fit_calls() == want["fit_calls"]
    assert rr.get_model_string() == want["model"]


@pytest.mark.parametrize("max_fit_calls,graph", [(100000, 0), (100000, 1)])
def test_config1_move_sequence_identical(max_fit_calls, graph):
    """The reference's own small case, the whole search; with graph=1 every
    plain FP64 fitness call goes out as a replayed executable graph."""
    print("case: config1 max_fit_calls=%d seed=12345 graph=%d" % (max_fit_calls, graph), flush=True)
    X, y = load_config1()
    want = O.fit(X, y, max_fit_calls=max_fit_calls, max_time=BIG_TIME, random_state=12345)
    rr = run_product(X, y, max_fit_calls=max_fit_calls, random_state=12345, options=[("graph", graph)])
    assert_same_run(rr, want)
    print("config1 model:", rr.get_model_string(), "fit_calls:", rr.get_fit_calls(), "stats:", rr.get_stats(),
          flush=True)
    # predict on the held-out rows equals the oracle's evaluation of the same model
    Xt, yt = load_config1("test")
    od = O.Dataset(Xt, yt)
    got = rr.predict(np.ascontiguousarray(Xt).reshape(-1), Xt.shape[0], Xt.shape[1])
    assert np.array_equal(got, od.eval(rr.get_model_prefix()))


def test_refits_of_one_object_repeat_the_moves():
    """Fitting the same object again starts from the same state: each fit
    makes the oracle's moves and launches the same candidates."""
    print("case: three fits of one object, config1, 20000 fit calls", flush=True)
    import rils_rols_cpp
    X, y = load_config1()
    want = O.fit(X, y, max_fit_calls=20000, max_time=BIG_TIME, random_state=12345)
    n, d = X.shape
    rr = rils_rols_cpp.rils_rols(False, 20000, BIG_TIME, 0.001, 50, 1.0, False, 12345)
    cands = []
    for _ in range(3):
        rr.fit(np.ascontiguousarray(X).reshape(-1), np.ascontiguousarray(y), n, d)
        assert_same_run(rr, want)
        cands.append(rr.get_stats()["batched_cands"])
    assert cands[0] == cands[1] == cands[2]


@pytest.mark.parametrize("seed,n,d,budget", [(1, 300, 3, 8000), (7, 1000, 5, 5000), (23654, 2000, 2, 4000)])
def test_random_problem_move_sequence_identical(seed, n, d, budget):
    print("case: random problem seed=%d n=%d d=%d budget=%d" % (seed, n, d, budget), flush=True)
    rng = np.random.default_rng(seed)
    X = rng.uniform(-3, 3, (n, d))
    y = 1.5 * np.sin(X[:, 0]) * X[:, d - 1] - 0.7 * X[:, 0] ** 2 + rng.normal(0, 0.01, n)
    want = O.fit(X, y, max_fit_calls=budget, max_time=BIG_TIME, random_state=seed, sample_size=0.9)
    rr = run_product(X, y, max_fit_calls=budget, random_state=seed, sample_size=0.9)
    assert_same_run(rr, want)


def test_classification_move_sequence_identical():
    print("case: breast_cancer classification budget=3000", flush=True)
    from sklearn.datasets import load_breast_cancer
    from sklearn.model_selection import train_test_split
    X, y = load_breast_cancer(return_X_y=True)
    X, _, y, _ = train_test_split(X, y, train_size=0.75, random_state=12345)
    want = O.fit(X, y, classification=True, max_fit_calls=3000, max_time=BIG_TIME, max_complexity=20,
                 random_state=12345)
    rr = run_product(X, y, classification=True, max_fit_calls=3000, max_complexity=20, random_state=12345)
    assert_same_run(rr, want)
    p = rr.predict(np.ascontiguousarray(X).reshape(-1), X.shape[0], X.shape[1])
    assert set(np.unique(p)) <= {0.0, 1.0}


def test_python_api_drop_in():
    """The estimator API of rils_rols/rils_rols.py:154-188 on the GPU path."""
    print("case: python API regressor/classifier", flush=True)
    from rils_rols import RILSROLSBinaryClassifier, RILSROLSRegressor
    X, y = load_config1()
    Xt, yt = load_config1("test")
    reg = RILSROLSRegressor(max_fit_calls=5000, sample_size=1, random_state=12345)
    reg.fit(X, y)
    want = O.fit(X, y, max_fit_calls=5000, max_time=100, random_state=12345)
    assert reg.model == want["model"]
    assert reg.fit_calls == want["fit_calls"]
    assert reg.score(Xt, yt) > 0.99
    assert "fitCalls=%d" % reg.fit_calls in reg.fit_report_string()
    assert reg.model_string() is not None
    Xc = X.copy()
    yc = (y > np.median(y)).astype(float)
    clf = RILSROLSBinaryClassifier(max_fit_calls=2000, max_complexity=20, random_state=3)
    clf.fit(Xc, yc)
    pr = clf.predict_proba(Xc)
    assert pr.shape == (len(yc), 2)
    assert 0.0 <= clf.score(Xc, yc) <= 1.0
    with pytest.raises(Exception):
        clf.fit(Xc, y)   # non-binary targets
    with pytest.raises(Exception):
        RILSROLSRegressor().predict(X)   # predict before fit


@pytest.mark.parametrize("case", ["config1", "random"])
def test_fp32_sweeps_accept_the_same_moves(case):
    """Sweeps scored by the certified FP32 form (FP64 only where a decision
    needs it) must give the run of the FP64 form: same moves, fitness bits,
    fit calls and model."""
    print("case: fp32 sweeps %s" % case, flush=True)
    if case == "config1":
        X, y = load_config1()
        kw = dict(max_fit_calls=30000, random_state=12345)
    else:
        rng = np.random.default_rng(5)
        X = rng.uniform(-3, 3, (20000, 4))
        y = np.exp(0.3 * X[:, 0]) * X[:, 1] - np.sqrt(np.abs(X[:, 2])) + rng.normal(0, 0.01, 20000)
        kw = dict(max_fit_calls=6000, random_state=5)
    want = O.fit(X, y, max_time=BIG_TIME, **kw)
    rr = run_product(X, y, precision="fp32", **kw)
    assert_same_run(rr, want)
    st = rr.get_stats()
    print("fp32 sweeps: batched %d, re-scored in FP64 %d" % (st["batched_cands"], st["rescored_fp64"]), flush=True)
    assert st["batched_cands"] > 0


@pytest.mark.parametrize("budget,calls,nmoves,model", [(10 ** 6, 37, 1, "x0"), (8, 8, 1, "x0"), (3, 4, 1, "x0"),
                                                       (2, 2, 0, "2.500000")])
def test_product_local_search_hand_trace(budget, calls, nmoves, model):
    """The hand-traced local search of tests/test_search_pins.py on the GPU path
    (x0 = [1, 2, 3, 4], y = x0, from the tree "1")."""
    import rils_rols_cpp
    from test_search_pins import MOVE, TRACE_X, TRACE_Y
    print("case: local search hand trace budget=%d" % budget, flush=True)
    rr = rils_rols_cpp.rils_rols(False, budget, BIG_TIME, 0.001, 50, 1.0, False, 0)
    rr.local_search(TRACE_X.reshape(-1), TRACE_Y, 4, 1, "c0x1p+0")
    assert rr.get_fit_calls() == calls
    assert [tuple(m) for m in rr.get_moves()] == [MOVE] * nmoves
    assert rr.get_model_string() == model


def test_more_than_200_features_move_sequence_identical():
    """d = 210 > max_feat: the search keeps the 200 features of highest
    R2(x_i, y) (relevant_features, rils_rols_cpp.cpp:753-770) and runs the
    oracle's moves."""
    from test_search_pins import feature_filter_data
    print("case: 210 features", flush=True)
    X, y = feature_filter_data(n=300, noise=True)
    want = O.fit(X, y, max_fit_calls=3000, max_time=BIG_TIME, random_state=4)
    rr = run_product(X, y, max_fit_calls=3000, random_state=4)
    assert_same_run(rr, want)


@pytest.mark.parametrize("rows", [20000, 8000])
def test_python_api_sample_size_auto(rows, capsys):
    """sample_size=0 (rils_rols.py:50-88): above 10000 rows a 10% split is fit
    with 1% of the budget and kept when its R2 and the whole set's R2 differ by
    less than 0.01; the real fit then runs with the rest of the budget.  The
    oracle, driven through the same decision, must end on the same model."""
    from sklearn.metrics import r2_score
    from sklearn.model_selection import train_test_split

    from rils_rols import RILSROLSRegressor
    print("case: sample_size=0 rows=%d" % rows, flush=True)
    rng = np.random.default_rng(rows)
    X = rng.uniform(-3, 3, (rows, 2))
    y = 2.0 * np.sin(X[:, 0]) + X[:, 1] + rng.normal(0, 0.01, rows)
    reg = RILSROLSRegressor(max_fit_calls=4000, max_time=10 ** 6, sample_size=0, random_state=9)
    reg.fit(X, y)
    out = capsys.readouterr().out
    assert "Automatically tuning sample size:" in out
    if rows <= 10000:
        assert "Training size is smaller than 10000 so setting it to full sample 1 (100%)." in out
        assert reg.sample_size == 1 and reg.max_fit_calls == 4000
        want = O.fit(X, y, max_fit_calls=4000, max_time=BIG_TIME, random_state=9)
    else:
        Xp, _, yp, _ = train_test_split(X, y, train_size=0.1, random_state=9)
        trial = O.fit(Xp, yp, max_fit_calls=40, max_time=BIG_TIME, random_state=9)
        od_p, od_a = O.Dataset(Xp, yp), O.Dataset(X, y)
        gap = abs(r2_score(y, od_a.eval(trial["prefix"])) - r2_score(yp, od_p.eval(trial["prefix"])))
        share = 0.1 if gap < 0.01 else 1
        assert "Setting sample_size=%s" % share in out
        assert reg.sample_size == share and reg.max_fit_calls == 4000 - 40.0
        want = O.fit(X, y, max_fit_calls=3960, max_time=BIG_TIME, random_state=9, sample_size=share)
    assert reg.model == want["model"]
    assert reg.fit_calls == want["fit_calls"]
    assert "maxFitCalls=%s" % reg.max_fit_calls in reg.fit_report_string()
