"""Exact recovery on the reference's own noise-free datasets
(paper_resources/random_12345_data; fixtures written by
tests/golden/make_recovery.py: the first 1000 rows of each).  The reference
reports R2 = 1 and the generating formula on every one of them
(random_12345_results/random_noise_0.txt); the search here must end on the
same formula, printed in the reference's to_string format (node.h:255-301),
with zero error on the data.  The oracle runs on the CPU; the product runs
the same searches on the GPU and must also agree with the oracle move for
move."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
BUDGET = 20000

# the generating formula (random_12345_truth/all.f) as the reference's to_string prints it; a
# search may end on it with its terms in another order (the order of its moves at the noise
# floor of an exact fit), so a model is compared with the truth as a formula
EXPECTED = {
    "random_04_02_0010000_02": "(ln(x0)+x1)",
    "random_04_01_0010000_00": "(ln(x0)+1)",
    "random_05_02_0010000_04": "(exp(x0)+sin(x1))",
    "random_04_02_0010000_04": "(x0*exp(x1))",
    "random_05_02_0010000_02": "(ln(x0)+cos(x1))",
    "random_06_02_0010000_04": "((((x0)**2)+x1)+-1)",
    "random_07_03_0010000_01": "(((x0+x1)+(-1*x2))+-1)",
}


def same_formula(model, name):
    """The model (reference print format) equals the generating formula."""
    import sympy
    truth = json.load(open(os.path.join(HERE, "golden", "recovery.json")))[name]["truth"]
    m = sympy.sympify(model.replace("ln(", "log("))
    return sympy.simplify(m - sympy.sympify(truth)) == 0 and \
        sympy.simplify(m - sympy.sympify(EXPECTED[name].replace("ln(", "log("))) == 0


def load(name):
    d = np.loadtxt(os.path.join(HERE, "golden", "recovery_%s.csv" % name), delimiter=",", skiprows=1)
    return d[:, :-1], d[:, -1]


def test_fixture_metadata_matches_expected_formulas():
    meta = json.load(open(os.path.join(HERE, "golden", "recovery.json")))
    assert set(meta) == set(EXPECTED)
    for name, m in meta.items():
        assert m["reference_result"]["R2"] == "1.0000000"


@pytest.mark.parametrize("name", ["random_04_02_0010000_02", "random_05_02_0010000_02", "random_07_03_0010000_01"])
def test_oracle_recovers_generating_formula(name):
    X, y = load(name)
    r = O.fit(X, y, max_fit_calls=BUDGET, max_time=10 ** 8, random_state=23654)
    assert same_formula(r["model"], name), r["model"]
    pred = O.Dataset(X, y).eval(r["prefix"])
    assert np.max(np.abs(pred - y)) <= 1e-9 * max(1.0, np.max(np.abs(y)))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_product_recovers_generating_formula(name):
    import rils_rols_cpp
    print("case: recovery %s" % name, flush=True)
    X, y = load(name)
    n, d = X.shape
    rr = rils_rols_cpp.rils_rols(False, BUDGET, 10 ** 8, 0.001, 50, 1.0, False, 23654)
    rr.fit(np.ascontiguousarray(X).reshape(-1), y, n, d)
    assert same_formula(rr.get_model_string(), name), rr.get_model_string()
    pred = rr.predict(np.ascontiguousarray(X).reshape(-1), n, d)
    assert np.max(np.abs(pred - y)) <= 1e-9 * max(1.0, np.max(np.abs(y)))
    want = O.fit(X, y, max_fit_calls=BUDGET, max_time=10 ** 8, random_state=23654)
    assert [tuple(m) for m in rr.get_moves()] == want["moves"]
