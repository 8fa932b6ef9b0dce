"""rr_fitness_batch_noalloc against rr_fitness_batch(RR_FP64), bit for bit.

The entry point carves every work buffer from the library's fixed scratch
array instead of allocating: two calls in a row (the second smaller), a
sample count whose work columns only fit in slices of the array, a batch too
large for the array (a clean RR_E_LIMIT, the context still usable), and
trees scored as given (RAW mode).
"""
import numpy as np
import pytest

import capi
import oracle_lib as O
import treegen

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return bool(np.all((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))))


def data(seed, n, d):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-4, 4, (n, d))
    y = np.sin(X[:, 0]) * X[:, d - 1] + 0.2 * X[:, 1] ** 2 + rng.normal(0, 0.01, n)
    return X, y


def assert_same(got, want):
    (cg, fg), (cw, fw) = got, want
    assert len(cg) == len(cw)
    for i, (a, b) in enumerate(zip(cg, cw)):
        assert same_bits(a, b), i
    assert fg.tobytes() == fw.tobytes()


def test_noalloc_matches_fitness_batch_two_calls():
    print("case: noalloc two calls n=20000 d=4, 300 then 80 trees", flush=True)
    ctx = capi.Context(0)
    X, y = data(91, 20000, 4)
    od = O.Dataset(X, y)
    trees = treegen.random_trees(91, 300, 3, 40, 4)
    facs = [od.tune(t)["factors"] for t in trees]
    h = ctx.upload(X, y)
    big = capi.build_batch(facs)
    small = capi.build_batch(facs[::4][:80])
    for arrs in (big, small):
        assert_same(ctx.fitness_noalloc(h, arrs), ctx.fitness(h, arrs))
    raw = capi.build_batch([[t] for t in trees[:120]])
    assert_same(ctx.fitness_noalloc(h, raw, mode="raw"), ctx.fitness(h, raw, mode="raw"))
    ctx.free(h)
    ctx.close()


def test_noalloc_slices_columns_to_fit_and_rejects_what_does_not():
    ctx = capi.Context(0)
    print("case: noalloc n=1e6 d=8, 52 candidates in slices of the scratch array", flush=True)
    X, y = data(92, 1_000_000, 8)
    h = ctx.upload(X, y)
    rng = np.random.default_rng(92)
    pool = ["sin x%d" % i for i in range(8)] + ["* x%d x%d" % (i, (i + 1) % 8) for i in range(8)] + \
           ["exp * x%d c0x1p-2" % i for i in range(8)]
    cands = [[pool[j] for j in rng.choice(len(pool), size=int(rng.integers(1, 6)), replace=False)] for _ in range(48)]
    # and four of more than 8 factors: the step-synchronous QR with its working columns
    cands += [[pool[j] for j in rng.choice(len(pool), size=int(rng.integers(9, 12)), replace=False)] for _ in range(4)]
    arrs = capi.build_batch(cands)
    assert_same(ctx.fitness_noalloc(h, arrs), ctx.fitness(h, arrs))
    ctx.free(h)
    print("case: noalloc n=6e6 d=8, one candidate of 8 factors: more than the array holds", flush=True)
    X, y = data(93, 6_000_000, 8)
    h = ctx.upload(X, y)
    wide = capi.build_batch([["sin x%d" % i for i in range(8)]])
    with pytest.raises(RuntimeError, match="scratch array"):
        ctx.fitness_noalloc(h, wide)
    c1, f1 = ctx.fitness(h, wide)           # the context's own buffers are intact
    assert np.isfinite(f1["e0"][0])
    small = capi.build_batch([["sin x0"], ["cos x1", "x2"]])
    assert_same(ctx.fitness_noalloc(h, small), ctx.fitness(h, small))
    ctx.free(h)
    ctx.close()
