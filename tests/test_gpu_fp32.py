"""FP32 precision form vs the FP64 form.

Every candidate the FP32 form certifies (flags without RR_F_NEEDS64) must
satisfy |fitness_value(FP32) - fitness_value(FP64)| <= err * fitness_value
with err the bound it reported, and certified sentinels must be FP64
sentinels.  Checked over the random-tree workload at several sizes and over
inputs chosen to hit overflow, domain edges and collinear factors.
"""
import numpy as np
import pytest

import capi
import oracle_lib as O
import treegen

pytestmark = pytest.mark.gpu


def fv(e0, e1):
    return (1 + e0) * (1 + e1)


def check_certificates(ctx, X, y, factor_lists):
    h = ctx.upload(X, y)
    arrs = capi.build_batch(factor_lists)
    c64, f64 = ctx.fitness(h, arrs, mode="ols", precision="fp64")
    c32, f32 = ctx.fitness(h, arrs, mode="ols", precision="fp32")
    ctx.free(h)
    certified = 0
    worst = 0.0
    for i in range(len(factor_lists)):
        a, b = f64[i], f32[i]
        if b["flags"] & 2:
            continue
        certified += 1
        if b["flags"] & 1:
            assert a["flags"] & 1, ("FP32 claims a sentinel, FP64 disagrees", i, factor_lists[i])
            continue
        assert not (a["flags"] & 1), ("FP32 certified a finite fitness, FP64 is a sentinel", i, factor_lists[i])
        x, z = fv(a["e0"], a["e1"]), fv(b["e0"], b["e1"])
        rel = abs(x - z) / x
        assert rel <= b["err"], (i, factor_lists[i], rel, b["err"], a, b)
        worst = max(worst, rel / max(b["err"], 1e-300))
        # kept / bare coefficient pattern of the OLS tree must agree
        for u, v in zip(c64[i], c32[i]):
            assert (abs(u) < 1e-12) == (abs(v) < 1e-12)
            assert (abs(u - 1) < 1e-12) == (abs(v - 1) < 1e-12)
    return certified, worst


@pytest.mark.parametrize("n,d,seed", [(1000, 4, 1), (20000, 10, 2), (100000, 10, 3)])
def test_fp32_certificates_hold_on_random_trees(ctx, n, d, seed):
    print("case: fp32 certificates n=%d d=%d seed=%d" % (n, d, seed), flush=True)
    rng = np.random.default_rng(seed)
    X = rng.uniform(-5, 5, (n, d))
    y = np.sin(X[:, 0]) * X[:, 1] - 0.5 * X[:, 2] ** 2 + np.exp(0.2 * X[:, 3]) + rng.normal(0, 0.01, n)
    import rils_rols_cpp
    trees = treegen.random_trees(seed, 600, 3, 40, d)
    facs = [rils_rols_cpp._ols_factors(t) for t in trees]
    facs = [f for f in facs if len(f) <= 8]
    # near-exact and collinear candidates
    facs += [["* x0 sin x0", "x1"], ["x0", "* x0 c0x1p+1"], ["sin c0x1p+0"], ["x0", "sqr x2", "exp * x3 c0x1.999999999999ap-3"],
             ["* sin x0 x1", "sqr x2", "exp * x3 c0x1.999999999999ap-3"], ["- x0 x0"], ["exp * x0 c0x1.4p+4"],
             ["sqrt - x0 x0"], ["ln x0"], ["/ x0 - x1 x1"]]
    certified, worst = check_certificates(ctx, X, y, facs)
    print("certified %d of %d, worst error/bound %.3g" % (certified, len(facs), worst), flush=True)
    assert certified > len(facs) // 2


@pytest.mark.parametrize("n,d,seed", [(3000, 4, 4), (100000, 10, 5)])
def test_fp32_checked_form_is_final(ctx, n, d, seed):
    """RR_FP32_CHECKED: no result is left uncertified; the results it takes
    from the FP64 form (RR_F_FP64) are bit-for-bit those of RR_FP64, and the
    others are exactly the RR_FP32_REFINED results, certified."""
    print("case: fp32 checked form n=%d d=%d seed=%d" % (n, d, seed), flush=True)
    rng = np.random.default_rng(seed)
    X = rng.uniform(-3, 3, (n, d))
    y = X[:, 0] * X[:, 1] + np.cos(X[:, 2]) + rng.normal(0, 0.05, n)
    import rils_rols_cpp
    trees = treegen.random_trees(seed, 500, 3, 50, d)
    facs = [rils_rols_cpp._ols_factors(t) for t in trees]
    facs = [f for f in facs if 1 <= len(f) <= 12]
    facs += [["c0x1p+1"], ["x0", "c0x1.8p+1"], ["x0", "sin c0x1p+0", "x1"], ["- x0 x0", "x1"],
             ["x0", "x1", "x2", "x3", "* x0 x1", "* x1 x2", "* x2 x3", "* x0 x3", "sqr x0"]]
    h = ctx.upload(X, y)
    arrs = capi.build_batch(facs)
    c64, f64 = ctx.fitness(h, arrs, mode="ols", precision="fp64")
    c32, f32 = ctx.fitness(h, arrs, mode="ols", precision="fp32")
    cc, fc = ctx.fitness(h, arrs, mode="ols", precision="fp32c")
    ctx.free(h)
    n_fp64 = 0
    for i in range(len(facs)):
        a, b, c = f64[i], f32[i], fc[i]
        assert not (c["flags"] & 2), (i, facs[i])
        if c["flags"] & 4:
            n_fp64 += 1
            assert c["e0"].tobytes() == a["e0"].tobytes() and c["e1"].tobytes() == a["e1"].tobytes(), (i, facs[i])
            assert c["rank"] == a["rank"] and (c["flags"] & ~4) == a["flags"], (i, facs[i])
            assert cc[i].tobytes() == c64[i].tobytes(), (i, facs[i])
        else:
            assert not (b["flags"] & 2), (i, facs[i])
            assert c.tobytes() == b.tobytes(), (i, facs[i])
            assert cc[i].tobytes() == c32[i].tobytes(), (i, facs[i])
    # the constant-factor and 9-factor candidates went to the FP64 form
    for i in range(len(facs) - 5, len(facs)):
        if i != len(facs) - 2:
            assert fc[i]["flags"] & 4, (i, facs[i])
    print("fp64 results %d of %d" % (n_fp64, len(facs)), flush=True)


def _check_samples(v64, v32, e32, what):
    """Per-sample contract of the FP32 evaluator: a certain NaN is NaN in
    FP64; a finite bound covers the distance to the FP64 value."""
    v64 = np.asarray(v64, dtype=np.float64)
    v = v32.astype(np.float64)
    e = e32.astype(np.float64)
    cnan = np.isnan(v32) & (e32 == 0)
    assert np.all(np.isnan(v64[cnan])), (what, "certain NaN where FP64 is not NaN")
    cert = np.isfinite(e) & ~cnan
    d = np.abs(v[cert] - v64[cert])
    bad = ~(d <= e[cert])
    assert not bad.any(), (what, v[cert][bad][:5], v64[cert][bad][:5], e[cert][bad][:5])
    ratio = float(np.max(d / np.maximum(e[cert], 1e-300))) if cert.any() else 0.0
    return int(cert.sum()), ratio


def test_fp32_device_math_within_bounds(ctx):
    """Hardware sin/cos/exp/log/sqrt/rcp of the FP32 evaluator against the
    canonical FP64 functions, sample by sample, over their whole usable
    ranges (including arguments near the domain edges and near 1 for ln)."""
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.uniform(-1600, 1600, 40000), rng.uniform(-20, 20, 40000), rng.uniform(-88, 88, 20000),
                        np.exp(rng.uniform(-60, 60, 20000)), 1 + rng.normal(0, 1e-3, 20000),
                        rng.normal(0, 1e-6, 10000), np.array([0.0, -0.0, 1.0, -1.0, 1e-30, 3e38])])
    X = x[:, None]
    od = O.Dataset(X, np.zeros_like(x))
    h = ctx.upload(X, np.zeros_like(x))
    report = {}
    for op in ["x0", "sin x0", "cos x0", "exp x0", "ln x0", "sqrt x0", "sqr x0", "/ c0x1p+0 x0", "* x0 x0",
               "- x0 c0x1.999999999999ap-3", "sin * x0 c0x1.4p+4", "exp sin x0", "ln sqr x0", "sqrt ln x0",
               "min x0 c0x1p+0", "max sin x0 cos x0", "< x0 c0x1p+0", "== sqr x0 x0"]:
        try:
            v64 = od.eval(op)
        except Exception:
            continue
        v32, e32 = ctx.predict_fp32(h, op, len(x))
        report[op] = _check_samples(v64, v32, e32, op)
    ctx.free(h)
    print("certified samples, max |v32-v64|/e:", report, flush=True)
    assert report["sin x0"][0] > 0.9 * len(x) - 40000   # beyond |x| = 1600 sin/cos are left at [-1, 1]
    assert report["exp x0"][0] > 0.5 * len(x)


def test_fp32_random_programs_within_bounds(ctx):
    """Random candidate programs (all operators) over data with signs, zeros
    and a wide range: every certified sample within its bound."""
    import rils_rols_cpp
    rng = np.random.default_rng(11)
    n, d = 4096, 4
    X = np.column_stack([rng.uniform(-5, 5, n), rng.normal(0, 100, n), np.round(rng.uniform(-3, 3, n)),
                         np.exp(rng.uniform(-20, 20, n))])
    od = O.Dataset(X, np.zeros(n))
    h = ctx.upload(X, np.zeros(n))
    trees = treegen.random_trees(11, 300, 3, 30, d)
    total = 0
    worst = 0.0
    for t in trees:
        for f in rils_rols_cpp._ols_factors(t):
            v64 = od.eval(f)
            v32, e32 = ctx.predict_fp32(h, f, n)
            c, r = _check_samples(v64, v32, e32, f)
            total += c
            worst = max(worst, r)
    ctx.free(h)
    print("certified samples %d, worst |v32-v64|/e %.3g" % (total, worst), flush=True)
    assert total > 0
