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


def test_fp32_device_math_within_assumed_bounds(ctx):
    """sinf/cosf/expf/logf/sqrtf of the device against the FP64 functions, on
    raw single-op programs: certified |v32 - v64| <= e must hold per sample,
    checked through the raw FP32 scoring of y = op(x0) (SSR = 0 exactly in
    FP64, so the FP32 result must be within its own bound of 0)."""
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.uniform(-20, 20, 20000), rng.uniform(0.001, 50, 20000), rng.normal(0, 3, 20000)])
    X = x[:, None]
    for op in ["sin x0", "cos x0", "exp x0", "ln x0", "sqrt x0", "sqr x0", "/ c0x1p+0 x0"]:
        od = O.Dataset(X, np.zeros_like(x))
        yv = od.eval(op)
        ok = np.isfinite(yv)
        Xo, yo = X[ok], yv[ok]
        h = ctx.upload(Xo, yo)
        _, f = ctx.fitness(h, capi.build_batch([[op]]), mode="raw", precision="fp32")
        ctx.free(h)
        f = f[0]
        if f["flags"] & 2:
            continue
        # FP64 fitness of the exact tree against its own values: e0 = 0, e1 = 0
        assert fv(f["e0"], f["e1"]) - 1.0 <= f["err"] * fv(f["e0"], f["e1"]) + 1e-15, (op, f)
