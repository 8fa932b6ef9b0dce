"""Product host engine (rils_rols_cpp hooks, C++) vs the oracle (CPU): exact.

Tree edits, strings, candidate generation and ordering must be bit-exact
(rils_rols_cpp.cpp:348-443; node.cpp).  Trees are compared in the prefix
exchange format, whose constants are exact hexadecimal doubles.
"""
import ctypes

import numpy as np
import pytest

import capi
import oracle_lib as O
import treegen

rrc = pytest.importorskip("rils_rols_cpp")


def corpus(seed, count, lo, hi, nvars, ops=treegen.OPS):
    trees = treegen.random_trees(seed, count, lo, hi, nvars, ops)
    # shapes that exercise every rewrite rule of node.cpp:152-374
    trees += [
        "* + x0 x1 x2", "* x0 - x1 x2", "* + x0 x1 - x2 x3", "* - x0 c0x1p+1 + x1 c0x1.8p+1",
        "+ c0x1p+1 - c0x1.8p+1 x0", "- c0x1p+1 - c0x1.8p+1 x0", "+ c0x1p+1 + x0 c0x1.8p+1",
        "- c0x1p+1 - x0 c0x1.8p+1", "- + x0 c0x1p+1 c0x1.8p+1", "- - x0 c0x1p+1 c0x1.8p+1",
        "+ + c0x1p+1 x0 c0x1.8p+1", "- + c0x1p+1 x0 c0x1.8p+1", "* c0x1p+1 * c0x1.8p+1 x0",
        "* c0x1p+1 * x0 c0x1.8p+1", "* * c0x1p+1 x0 c0x1.8p+1", "* * x0 c0x1p+1 c0x1.8p+1",
        "+ c0x0p+0 x0", "- c0x0p+0 x0", "* c0x0p+0 x0", "* c0x1p+0 x0", "/ c0x0p+0 x0", "+ x0 c0x0p+0",
        "* x0 c0x1p+0", "* x0 c0x0p+0", "/ x0 c0x1p+1", "+ * c0x1.8p+1 x0 c0x1.4p+2", "sqr + x0 x1",
        "* sqr + x0 x1 + x0 x1", "+ c0x1.999999999999ap-4 c0x1.3333333333333p-3", "sin c0x1p+1",
        "exp * x0 + x1 c0x1p+0", "- x0 x0", "/ + x0 x1 c0x1p+2", "c0x1.6p+3", "x1", "* c0x1.2a05f2p+31 x0",
        "+ c-0x0p+0 c0x1.5p-1", "< x0 x1", "> x0 x1", "== x0 x1", "!= x0 x1", "min x0 c0x1p+0", "max x0 x1",
    ]
    return trees


@pytest.mark.parametrize("op", ["to_string", "expand", "simplify", "normalize", "candidate_rounds",
                                "accept_rounds", "size"])
def test_rewrites_match_oracle(op):
    for t in corpus(101, 400, 1, 40, 4):
        got = rrc._rewrite(t, op)
        want = O.rewrite(t, op)
        assert str(got) == want, (op, t)


def test_ols_factors_match_oracle():
    X = np.random.default_rng(0).uniform(-5, 5, (50, 4))
    od = O.Dataset(X, X[:, 0])
    for t in corpus(102, 400, 1, 40, 4):
        assert rrc._ols_factors(t) == od.tune(t)["factors"], t


def test_ols_tree_matches_oracle():
    rng = np.random.default_rng(5)
    X = rng.uniform(-5, 5, (60, 4))
    y = 2.0 * X[:, 0] - np.sin(X[:, 1]) + 0.5
    od = O.Dataset(X, y)
    for t in corpus(103, 200, 3, 30, 3, treegen.BI + ["sin", "sqr"]):
        tu = od.tune(t)
        assert rrc._ols_tree(tu["factors"], list(tu["coefs"])) == tu["ols"], t
    # threshold cases of rils_rols_cpp.cpp:492-505 (|c| < 1e-12 dropped, |c-1| < 1e-12 bare)
    f = ["x0", "sin x1"]
    for coefs in ([1e-12, 1.0, 0.0], [9e-13, 1 + 9e-13, 2.5], [-9e-13, 1 - 1e-12, 1e-12], [0.0, 0.0, 0.0],
                  [1.0, 1.0, 1.0], [float("nan"), 2.0, 0.0]):
        got = rrc._ols_tree(f, coefs)
        assert got is not None


@pytest.mark.parametrize("classification", [False, True])
@pytest.mark.parametrize("local_search", [True, False])
def test_all_candidates_match_oracle(classification, local_search):
    ops = treegen.OPS + (treegen.CLS_BI if classification else [])
    trees = treegen.random_trees(104 + classification, 25, 1, 14, 3, ops)
    trees += ["c0x0p+0", "x0", "+ x0 c0x1.8p+1", "sin x1", "* + x0 x1 x2", "- c0x1p+1 * c0x1.8p+1 x0"]
    for t in trees:
        got = rrc._all_candidates(t, 3, classification, local_search)
        want = O.all_candidates(t, 3, classification, local_search)
        assert got == want, t


def test_compile_matches_test_compiler():
    for t in corpus(105, 300, 1, 60, 5, treegen.OPS + treegen.CLS_BI):
        code, consts, depth = rrc._compile(t)
        arrs = capi.build_batch([[t]])
        assert list(code) == list(arrs["code"]), t
        assert np.array_equal(np.array(consts), arrs["consts"][:arrs["n_consts"]]), t


def test_stack_depth_is_logarithmic():
    # a perfectly balanced tree of 2^8 leaves needs 7 pushes with Sethi-Ullman ordering
    def bal(d):
        return "sin x0" if d == 0 else "+ " + bal(d - 1) + " " + bal(d - 1)
    _, _, depth = rrc._compile(bal(7))
    assert depth == 7
    with pytest.raises(Exception):
        rrc._compile(bal(9))


def test_crand_is_glibc_rand():
    libc = ctypes.CDLL("libc.so.6")
    for seed in (0, 1, 12345, 23654, 2 ** 31 - 1):
        libc.srand(seed)
        want = [libc.rand() for _ in range(200)]
        assert rrc._crand(seed, 200) == want, seed


def test_constant_strings_are_printf_f():
    """A constant's text in the reference's strings (node.h:255-301:
    std::to_string, i.e. printf "%f", for a non-integer; the integer as
    "%d" when it is within 1e-12 of one): the product's formatter against
    Python's "%f" (correctly rounded, as C's) over random magnitudes, ties of
    the sixth decimal, huge, tiny and signed-zero values."""
    rng = np.random.default_rng(7)
    vals = list(rng.normal(0, 1, 3000)) + list(rng.normal(0, 1, 3000) * 10.0 ** rng.integers(-12, 40, 3000))
    vals += [(k + 0.5) * 1e-6 for k in range(-50, 50)] + [0.0000015, 2.5e-7, 1e300, -1e300, 5e-324,
                                                          123456.0000005, 0.1, -0.1, 1.0 / 3.0]
    for v in vals:
        v = float(v)
        r = round(v)
        want = ("%d" % (int(r) if -2147483648.0 <= r < 2147483648.0 else -2147483648)
                if abs(r - v) < 1e-12 else "%f" % v)
        got = rrc._rewrite("c" + v.hex(), "to_string")
        assert got == want, (v.hex(), got, want)
