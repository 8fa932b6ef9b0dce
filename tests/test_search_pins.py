"""Known answers, derived by hand from the reference, for the candidate lists
of the local search and for one local search end to end.

all_candidates (rils_rols_cpp.cpp:348-443) in local-search mode: the whole
ordered list for two trees, (x0+2) and ln(sin(x0)), with one feature
(tests/golden/all_candidates_ls_*.txt).  The lists fix the generator order of
change_candidates (332-346), the BFS order over subtrees (372-410), the
string dedup at insertion (363-370), the expand/normalize/simplify rounds and
their early stop on an unchanged size (413-425), and the dedup after them
(427-439).  Both the oracle and the product's host engine must give them.

local_search (601-643), traced by hand on x0 = [1, 2, 3, 4], y = x0, from
the tree "1":
  fit call 1   tune_constants(1) -> the constant 2.5 (the mean), fitness
               (1, sqrt(1.25), 1);
  sweep 1      all_candidates(2.5): 13 candidates (L1 below); j = 0 "1"
               tunes to the same constant, equal fitness, not accepted (fit
               call 2); j = 1 "x0" tunes to the bare x0 (coefficient within
               1e-12 of 1, free term within 1e-12 of 0, rils_rols_cpp.cpp:
               488-517) whose residual is exactly zero: fitness (0, 0, 1),
               accepted (fit call 3), its rewrite rounds change nothing and
               the extra fitness call (631) is fit call 4, logged as the
               move; pareto = {(0, 0, 1)} dominates every later candidate
               (e0, e1 >= 0, size >= 1), so j = 2..12 are fit calls 5..15;
  sweep 2      improved, so all_candidates(x0): 22 candidates (L2 below),
               none accepted: fit calls 16..37; the loop ends.
Budget cuts (finished() before each candidate, 613): 8 calls end at 8 with
the same move; 3 calls end at 4 (the acceptance's extra call runs past the
budget); 2 calls end at 2 with the constant and no move.

The pareto check (619) never decides an acceptance on its own: a fitness
dominated by a pareto member P has fitness_value >= P's, and P was accepted
no later than the incumbent, whose fitness_value is lower still, unless the
size rule of compare_fitness (552-553) decided, which only ever lowers a size
above max_complexity; so no trace can show it rejecting a candidate that
compare_fitness accepts, and none is asserted.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))

L1 = ["1", "x0", "sin(2.500000)", "cos(2.500000)", "ln(2.500000)", "exp(2.500000)", "sqrt(2.500000)",
      "((2.500000)**2)", "(1+x0)", "(1-x0)", "(x0-1)", "(2.500000/x0)", "(x0/1)"]
L2 = ["1", "sin(x0)", "cos(x0)", "ln(x0)", "exp(x0)", "sqrt(x0)", "((x0)**2)", "(x0+x0)", "(x0+1)", "(x0-x0)",
      "(x0-1)", "(1-x0)", "(x0*x0)", "x0", "(x0/x0)", "(x0/1)", "(-1/x0)", "(0.500000/x0)", "(1/x0)", "(2/x0)",
      "(3.141593/x0)", "(10/x0)"]

TREES = {"all_candidates_ls_x0_plus_2.txt": "+ x0 c0x1p+1", "all_candidates_ls_ln_sin_x0.txt": "ln sin x0"}


def golden(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return [ln.strip() for ln in f if ln.strip() and not ln.startswith("#")]


@pytest.mark.parametrize("name", sorted(TREES))
def test_oracle_local_search_candidates_known_answer(name):
    got = [O.to_string(c) for c in O.all_candidates(TREES[name], 1, local_search=True)]
    assert got == golden(name)


@pytest.mark.parametrize("name", sorted(TREES))
def test_product_local_search_candidates_known_answer(name):
    import rils_rols_cpp
    got = [rils_rols_cpp._rewrite(c, "to_string") for c in rils_rols_cpp._all_candidates(TREES[name], 1)]
    assert got == golden(name)


def test_trace_candidate_lists():
    assert [O.to_string(c) for c in O.all_candidates("c0x1.4p+1", 1)] == L1
    assert [O.to_string(c) for c in O.all_candidates("x0", 1)] == L2


TRACE_X = np.array([[1.0], [2.0], [3.0], [4.0]])
TRACE_Y = TRACE_X[:, 0].copy()
MOVE = (0, 4, 0.0, 0.0, 1, "x0")


@pytest.mark.parametrize("budget,calls,moves,model", [
    (10 ** 6, 1 + len(L1) + 1 + len(L2), [MOVE], "x0"),
    (8, 8, [MOVE], "x0"),
    (3, 4, [MOVE], "x0"),
    (2, 2, [], "2.500000"),
])
def test_oracle_local_search_hand_trace(budget, calls, moves, model):
    r = O.local_search(TRACE_X, TRACE_Y, "c0x1p+0", max_fit_calls=budget)
    assert r["fit_calls"] == calls
    assert r["moves"] == moves
    assert r["model"] == model
    if model == "x0":
        assert r["final"] == (0.0, 0.0, 1)
    else:   # the mean: R2 = 0 exactly, RMSE = sqrt(1.25)
        assert r["final"] == (1.0, float(np.sqrt(1.25)), 1)


def feature_filter_data(n=50, noise=False):
    """210 features x_i = c_i u with y = u (u = sin(row)): R2(x_i, y) of
    relevant_features (rils_rols_cpp.cpp:753-770, R2 with x_i in the role of
    y) is 1 - ((c_i - 1) / c_i)^2, decreasing in c_i >= 1.  c_i = 1 + 0.01 (209
    - i), except x_5, a copy of x_150 (a tie in R2)."""
    u = np.sin(np.arange(n, dtype=np.float64))
    c = 1.0 + 0.01 * (209 - np.arange(210))
    X = np.outer(u, c)
    X[:, 5] = X[:, 150]
    y = u.copy()
    if noise:   # for a search: a target that is not one of the features
        y = u + 0.3 * np.cos(3.0 * np.arange(n)) + 0.1 * u * u
    return X, y


# Sorted by (R2, index) descending (std::greater on the tuple, 767): 209 down
# to 150, then the tie x_5 (same R2 as x_150, smaller index), then 149 down to
# 11; x_10 and lower fall outside the 200.
FEATURES_KEPT = list(range(209, 149, -1)) + [5] + list(range(149, 10, -1))


def test_relevant_features_known_answer():
    assert len(FEATURES_KEPT) == 200
    X, y = feature_filter_data()
    assert O.relevant_features(X, y) == FEATURES_KEPT
    import rils_rols_cpp
    assert rils_rols_cpp._relevant_features(X, y) == FEATURES_KEPT
    # 200 features or fewer: all, in order
    assert O.relevant_features(X[:, :200], y) == list(range(200))
    assert rils_rols_cpp._relevant_features(X[:, :200], y) == list(range(200))
