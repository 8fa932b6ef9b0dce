"""FP64 GPU path vs the CPU oracle: bit-for-bit.

The GPU kernels and the oracle implement the same canonical arithmetic
(rr_math functions, 64-lane canonical sums, Eigen's ColPivHouseholderQR
operation order), so coefficients, 1-R2 and RMSE must be identical bits; the
only allowance is the sign of a NaN produced by a hardware operation (x86 and
AMDGPU default NaNs differ), which is compared as "both NaN".
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
    both_nan = np.isnan(a) & np.isnan(b)
    eq = a.view(np.uint64) == b.view(np.uint64)
    return bool(np.all(eq | both_nan))


def make_data(seed, n, d, lo=-5.0, hi=5.0):
    rng = np.random.default_rng(seed)
    X = rng.uniform(lo, hi, size=(n, d))
    y = np.sin(X[:, 0]) * 3.0 + X[:, min(1, d - 1)] ** 2 - 0.5 * X[:, 0] + rng.normal(0, 0.01, n)
    return X, y


SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 2.0, np.pi, 1e-300, 5e-324, 1e300, -1e300, np.inf, -np.inf, np.nan,
                    709.78, 709.79, -745.1, -745.2, 1e8, 2.0 ** 28, 2.0 ** 28 + 1, 1e22, 6.38e189, 1e308,
                    3.0 * np.pi / 2, 1e-8, 0.7853981633974483, 710.0, 1e16])


def test_elementwise_ops_bitexact(ctx):
    print("case: elementwise ops over special + random inputs", flush=True)
    rng = np.random.default_rng(7)
    x = np.concatenate([SPECIAL, -SPECIAL, rng.uniform(-10, 10, 3000), rng.uniform(-1e6, 1e6, 500),
                        10.0 ** rng.uniform(-300, 300, 500), rng.normal(0, 1, 1000)])
    y2 = rng.normal(0, 2, x.size)
    X = np.stack([x, y2], axis=1)
    h = ctx.upload(X, np.zeros(x.size))
    od = O.Dataset(X, np.zeros(x.size))
    for pre in ["sin x0", "cos x0", "ln x0", "exp x0", "sqrt x0", "sqr x0", "+ x0 x1", "- x0 x1", "* x0 x1",
                "/ x0 x1", "< x0 x1", "> x0 x1", "== x0 x0", "!= x0 x1", "min x0 x1", "max x0 x1",
                "/ x0 c0x0p+0", "- c0x1.8p+1 x0"]:
        g = ctx.predict(h, pre, x.size)
        o = od.eval(pre)
        assert same_bits(g, o), pre
    ctx.free(h)


def test_dataset_stats_match(ctx):
    X, y = make_data(3, 1000, 3)
    h = ctx.upload(X, y)
    assert ctx.stats(h) == O.Dataset(X, y).stats()
    ctx.free(h)


@pytest.mark.parametrize("n,d,seed", [(150, 2, 1), (1000, 5, 2), (4099, 10, 3)])
def test_raw_fitness_bitexact(ctx, n, d, seed):
    print("case: raw n=%d d=%d seed=%d" % (n, d, seed), flush=True)
    X, y = make_data(seed, n, d)
    trees = treegen.random_trees(seed, 300, 1, 40, d)
    h = ctx.upload(X, y)
    od = O.Dataset(X, y)
    arrs = capi.build_batch([[t] for t in trees])
    _, fit = ctx.fitness(h, arrs, mode="raw")
    for t, f in zip(trees, fit):
        e0, e1, _ = od.score(t)
        assert same_bits([f["e0"], f["e1"]], [e0, e1]), t
    ctx.free(h)


# 70001 and 135000 rows: 18 and 33 chunks, folded in two and three groups of
# 16 chunks (the last group of one chunk or a few)
@pytest.mark.parametrize("n,d,seed", [(150, 2, 11), (777, 4, 12), (5000, 10, 13), (64, 3, 14), (7, 3, 15),
                                      (70001, 6, 16), (135000, 3, 17)])
def test_ols_bitexact(ctx, n, d, seed):
    print("case: ols n=%d d=%d seed=%d" % (n, d, seed), flush=True)
    X, y = make_data(seed, n, d)
    trees = treegen.random_trees(seed, 250, 3, 40, d)
    trees += ["+ + * c0x1p+1 x0 sin x1 c0x1.8p+1", "+ sin c0x1p+0 x0", "+ x0 x0", "- * x0 c0x1p+1 + x0 x0",
              "+ + x0 sqr x1 * x0 x1", "ln - x0 x0", "+ x0 ln - x0 x0", "sin c0x1p+0"]
    od = O.Dataset(X, y)
    tuned = [od.tune(t) for t in trees]
    arrs = capi.build_batch([tu["factors"] for tu in tuned])
    h = ctx.upload(X, y)
    coefs, fit = ctx.fitness(h, arrs, mode="ols")
    for t, tu, c, f in zip(trees, tuned, coefs, fit):
        assert len(c) == len(tu["coefs"]), t
        assert same_bits(c, tu["coefs"]), (t, c, tu["coefs"])
        assert f["rank"] == tu["rank"], t
        assert same_bits([f["e0"], f["e1"]], [tu["fit"][0], tu["fit"][1]]), (t, f, tu["fit"])
    ctx.free(h)


def _full_tree(rng, levels, d):
    """A full binary tree: its program needs levels - 1 stack slots even in
    the best evaluation order."""
    if levels == 0:
        return "x%d" % rng.integers(d)
    op = rng.choice(["*", "/", "-", "+", "max"]) if levels > 1 else rng.choice(["*", "-", "+"])
    return "%s %s %s" % (op, _full_tree(rng, levels - 1, d), _full_tree(rng, levels - 1, d))


@pytest.mark.parametrize("form,rows", [(1, 1), (1, 4), (0, 1), (0, 2), (0, 4)])
def test_eval_forms_bitexact(ctx, form, rows):
    """Both factor interpreters (fp64_eval_form 0: operand stack in registers,
    transcendentals called; 1: stack in LDS, inlined) at every row count
    besides the default (form 1, 2 rows, which every other test runs): the
    same bits as the oracle, over a ragged last chunk, with programs of stack
    depth up to 4."""
    print("case: eval form %d rows=%d" % (form, rows), flush=True)
    X, y = make_data(61, 9001, 4)
    rng = np.random.default_rng(61)
    trees = treegen.random_trees(61, 200, 3, 40, 4)
    shallow = trees[:]
    deep = ["* sin %s cos %s" % (_full_tree(rng, L, 4), _full_tree(rng, L, 4)) for L in (2, 3, 4, 5) for _ in range(3)]
    od = O.Dataset(X, y)
    h = ctx.upload(X, y)
    ctx.set_option("fp64_eval_rows", rows)
    ctx.set_option("fp64_eval_form", form)
    try:
        for batch in (shallow, shallow + deep):
            tuned = [od.tune(t) for t in batch]
            coefs, fit = ctx.fitness(h, capi.build_batch([tu["factors"] for tu in tuned]), mode="ols")
            for t, tu, c, f in zip(batch, tuned, coefs, fit):
                assert same_bits(c, tu["coefs"]), (t, c, tu["coefs"])
                assert f["rank"] == tu["rank"], t
                assert same_bits([f["e0"], f["e1"]], [tu["fit"][0], tu["fit"][1]]), (t, f, tu["fit"])
    finally:
        ctx.set_option("fp64_eval_rows", 2)
        ctx.set_option("fp64_eval_form", 1)
    ctx.free(h)


def test_ols_chunked_slab_same_bits(ctx):
    """A tiny slab budget forces many launches; results must not change."""
    X, y = make_data(21, 3000, 4)
    trees = treegen.random_trees(21, 120, 3, 30, 4)
    od = O.Dataset(X, y)
    arrs = capi.build_batch([od.tune(t)["factors"] for t in trees])
    h = ctx.upload(X, y)
    c1, f1 = ctx.fitness(h, arrs)
    ctx.set_option("slab_budget_mb", 1)
    c2, f2 = ctx.fitness(h, arrs)
    ctx.set_option("slab_budget_mb", 32768)
    assert all(same_bits(a, b) for a, b in zip(c1, c2))
    assert same_bits(f1["e0"], f2["e0"]) and same_bits(f1["e1"], f2["e1"])
    ctx.free(h)


def test_bad_program_rejected_before_launch(ctx):
    X, y = make_data(1, 100, 2)
    h = ctx.upload(X, y)
    arrs = capi.build_batch([["x5"]])          # variable index out of range
    with pytest.raises(RuntimeError, match="out of range"):
        ctx.fitness(h, arrs, mode="raw")
    arrs = capi.build_batch([["x0"]])
    arrs["code"] = np.array([capi.D["RR_OP_BIN"]], dtype=np.uint32)  # binary op on empty accumulator
    with pytest.raises(RuntimeError):
        ctx.fitness(h, arrs, mode="raw")
    ctx.free(h)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_multi_dataset_batch_same_bits_as_solo(ctx, precision):
    """Candidates of three datasets (different n, d, y) in one rr_fitness_multi
    launch, interleaved, give the same bits as each dataset's own batch."""
    print("case: multi-dataset batch %s" % precision, flush=True)
    shapes = [(1500, 3, 31), (7000, 2, 32), (333, 4, 33)]
    hs, per_ds = [], []
    for n, d, seed in shapes:
        X, y = make_data(seed, n, d)
        od = O.Dataset(X, y)
        trees = treegen.random_trees(seed, 40, 3, 30, d)
        per_ds.append([od.tune(t)["factors"] for t in trees])
        hs.append(ctx.upload(X, y))
    for mode in ["ols", "raw"]:
        lists = per_ds if mode == "ols" else [[[f[0]] for f in fl if f] for fl in per_ds]
        solo = [ctx.fitness(h, capi.build_batch(fl), mode=mode, precision=precision) for h, fl in zip(hs, lists)]
        mixed, owner = [], []
        for i in range(max(len(fl) for fl in lists)):
            for s, fl in enumerate(lists):
                if i < len(fl):
                    mixed.append(fl[i])
                    owner.append((s, i))
        cm, fm = ctx.fitness(None, capi.build_batch(mixed), mode=mode, precision=precision,
                             handles=[hs[s] for s, _ in owner])
        for q, (s, i) in enumerate(owner):
            cs, fs = solo[s]
            assert same_bits([fm[q]["e0"], fm[q]["e1"], fm[q]["err"]], [fs[i]["e0"], fs[i]["e1"], fs[i]["err"]]), \
                (mode, s, i)
            assert fm[q]["flags"] == fs[i]["flags"] and fm[q]["rank"] == fs[i]["rank"]
            if mode == "ols":
                assert same_bits(cm[q], cs[i]), (s, i)
    for h in hs:
        ctx.free(h)


def test_autotune_keeps_fp64_bits(ctx):
    """Whatever schedule the autotuner picks, FP64 results are the same bits
    (the summation order does not depend on rows per thread)."""
    print("case: autotune fp64", flush=True)
    X, y = make_data(41, 20000, 4)
    trees = treegen.random_trees(41, 150, 3, 30, 4)
    od = O.Dataset(X, y)
    arrs = capi.build_batch([od.tune(t)["factors"] for t in trees])
    h = ctx.upload(X, y)
    ctx.set_option("fp64_eval_rows", 1)
    c1, f1 = ctx.fitness(h, arrs)
    ms = ctx.autotune(h, arrs, precision="fp64")
    assert ms > 0 and ctx.get_option("fp64_eval_rows") in (1, 2, 4) and ctx.get_option("fp64_eval_form") in (0, 1)
    c2, f2 = ctx.fitness(h, arrs)
    assert all(same_bits(a, b) for a, b in zip(c1, c2))
    assert same_bits(f1["e0"], f2["e0"]) and same_bits(f1["e1"], f2["e1"])
    ms32 = ctx.autotune(h, arrs, precision="fp32")
    assert ms32 > 0 and ctx.get_option("fp32_chunk_rows") in (2048, 4096, 8192, 16384)
    ctx.set_option("fp64_eval_rows", 2)
    ctx.set_option("fp64_eval_form", 1)
    ctx.set_option("fp32_rows", 2)
    ctx.set_option("fp32_chunk_rows", 4096)
    ctx.set_option("fp32_streams", 0)
    ctx.free(h)


def test_shared_factors_same_bits(ctx):
    """Candidates that repeat the same factors (equal programs, constants at
    different pool indexes), inside one candidate too, over two datasets in one
    launch, with candidates settled by the factor-0 pre-pass (ln x0 first over
    x0 < 0) sharing factors with live ones, and factors only such candidates
    use: every result is the same bits as the candidate scored alone, and the
    coefficients of candidates with finite columns are the oracle's QR
    solution of [factors, 1] against y."""
    print("case: shared factors, two datasets", flush=True)
    pool = ["sin x0", "* x0 c0x1p+1", "sqr + x1 c0x1.8p+1", "cos * x0 x1", "exp * x1 c0x1p-2",
            "* x0 c0x1.0000000000001p+1", "+ x0 c0x0p+0", "+ x0 c-0x0p+0", "/ x1 + sqr x0 c0x1p+0"]
    only_dead = ["sqrt x1", "sqr sin x1"]
    rng = np.random.default_rng(51)
    cands = []
    for i in range(160):
        f = [pool[j] for j in rng.choice(len(pool), size=int(rng.integers(1, 5)), replace=True)]
        if i % 7 == 0:
            f.append("x%d" % (i % 2))
        if i % 5 == 0:
            f = ["ln x0"] + f + [only_dead[i % 2]]
        cands.append(f)
    cands.append(["sin x0", "sin x0"])
    datasets = [make_data(52, 4100, 2), make_data(53, 9000, 2)]
    hs = [ctx.upload(X, y) for X, y in datasets]
    owner = [i % 2 for i in range(len(cands))]
    cm, fm = ctx.fitness(None, capi.build_batch(cands), handles=[hs[s] for s in owner])
    ods = [O.Dataset(X, y) for X, y in datasets]
    for q, (f, s) in enumerate(zip(cands, owner)):
        cs, fs = ctx.fitness(hs[s], capi.build_batch([f]))
        assert same_bits([fm[q]["e0"], fm[q]["e1"]], [fs[0]["e0"], fs[0]["e1"]]), (q, f)
        assert fm[q]["rank"] == fs[0]["rank"] and fm[q]["flags"] == fs[0]["flags"], (q, f)
        assert same_bits(cm[q], cs[0]), (q, f)
        cols = [ods[s].eval(p) for p in f] + [np.ones(datasets[s][1].size)]
        A = np.stack(cols, axis=1)
        if np.all(np.isfinite(A)):
            c, rank = O.qr_solve(A, datasets[s][1])
            assert same_bits(cm[q], c) and fm[q]["rank"] == rank, (q, f)
    assert any(np.isnan(c).all() for c in cm)   # the pre-pass settled some
    for h in hs:
        ctx.free(h)


def _shared_factor_batch(seed, count):
    """Candidates drawn from a small pool of factors (equal programs with
    constants at different pool indexes), some with a constant factor."""
    pool = ["sin x0", "* x0 c0x1p+1", "sqr + x1 c0x1.8p+1", "cos * x0 x1", "exp * x1 c0x1p-2", "sqrt x1",
            "/ x1 + sqr x0 c0x1p+0", "* sin x0 cos x1", "c0x1.8p+1"]
    rng = np.random.default_rng(seed)
    return [[pool[j] for j in rng.choice(len(pool), size=int(rng.integers(1, 6)), replace=False)]
            for _ in range(count)]


def test_shared_factors_across_slab_chunks_same_bits(ctx):
    """Kept columns are shared within a launch chunk only: with a 1 MB slab
    budget (a few candidates per chunk) the results are the same bits as in
    one chunk."""
    print("case: shared factors, many slab chunks", flush=True)
    X, y = make_data(61, 6000, 2)
    arrs = capi.build_batch(_shared_factor_batch(61, 300))
    h = ctx.upload(X, y)
    c1, f1 = ctx.fitness(h, arrs)
    ctx.set_option("slab_budget_mb", 1)
    c2, f2 = ctx.fitness(h, arrs)
    ctx.set_option("slab_budget_mb", 32768)
    ctx.free(h)
    assert all(same_bits(a, b) for a, b in zip(c1, c2))
    assert f1.tobytes() == f2.tobytes()


def test_shared_factors_checked_form_takes_fp64_bits(ctx):
    """RR_FP32_CHECKED plans FP64 launches over subsets of the batch (the
    candidates sent to FP64 up front, then the ones the FP32 form could not
    certify); on a batch that repeats its factors, every result it takes
    from the FP64 form is the plain FP64 form's, bit for bit."""
    print("case: shared factors, checked form", flush=True)
    X, y = make_data(62, 5000, 2)
    arrs = capi.build_batch(_shared_factor_batch(62, 300))
    h = ctx.upload(X, y)
    c64, f64 = ctx.fitness(h, arrs, precision="fp64")
    cc, fc = ctx.fitness(h, arrs, precision="fp32c")
    ctx.free(h)
    n64 = 0
    for i in range(len(fc)):
        if fc[i]["flags"] & 4:
            n64 += 1
            assert fc[i]["e0"].tobytes() == f64[i]["e0"].tobytes(), i
            assert fc[i]["e1"].tobytes() == f64[i]["e1"].tobytes(), i
            assert cc[i].tobytes() == c64[i].tobytes(), i
    assert n64 > 0   # the constant-factor candidates at least


@pytest.mark.parametrize("n", [2, 3, 4, 5, 9, 300, 4097])
def test_sparse_columns_and_tiny_n_bitexact(ctx, n):
    """The QR's last two steps on the edge cases they branch on: columns that
    are zero below row 0 or 1 (an identity reflector, tau = 0, at the step
    before the last), a single row left (n - p == 1), and n at or just above
    the column count."""
    print("case: sparse columns n=%d" % n, flush=True)
    rng = np.random.default_rng(70 + n)
    X = np.zeros((n, 4))
    X[:, 0] = rng.uniform(-3, 3, n)
    X[0, 1] = 2.5                      # zero below row 0
    X[:2, 2] = [1.5, -0.75]            # zero below row 1
    X[:, 3] = rng.uniform(0.5, 2.0, n)
    y = 1.5 * X[:, 0] - 0.5 * X[:, 3] + rng.normal(0, 0.01, n)
    trees = ["x1", "x2", "+ x1 x0", "+ x2 x0", "+ x1 x2", "+ + x0 x1 x2", "+ + x1 x2 x3", "* x1 c0x1p+1",
             "+ x0 x3", "+ + x0 x3 sqr x1", "+ + + x0 x1 x2 x3", "+ sin x1 x0", "+ sqr x2 x1"]
    od = O.Dataset(X, y)
    tuned = [od.tune(t) for t in trees]
    arrs = capi.build_batch([tu["factors"] for tu in tuned])
    h = ctx.upload(X, y)
    coefs, fit = ctx.fitness(h, arrs, mode="ols")
    ctx.free(h)
    for t, tu, c, f in zip(trees, tuned, coefs, fit):
        assert same_bits(c, tu["coefs"]), (t, c, tu["coefs"])
        assert f["rank"] == tu["rank"], t
        assert same_bits([f["e0"], f["e1"]], [tu["fit"][0], tu["fit"][1]]), (t, f, tu["fit"])


def wide_trees(rng, d, count, lo=10, hi=20):
    """Sums of lo..hi-1 distinct terms over d variables: candidates of more
    than ten columns (with the free term), which take the step-synchronous QR."""
    pool = []
    for i in range(d):
        pool += ["x%d" % i, "sin x%d" % i, "cos x%d" % i, "sqr x%d" % i, "exp * c0x1p-2 x%d" % i]
        pool += ["* x%d x%d" % (i, j) for j in range(i + 1, d)]
    trees = []
    for _ in range(count):
        m = int(rng.integers(lo, hi))
        terms = [pool[i] for i in rng.choice(len(pool), m, replace=False)]
        trees.append("+ " * (m - 1) + " ".join(terms))
    return trees


@pytest.mark.parametrize("n", [7, 12, 150, 4096, 4097, 9000])
def test_wide_candidates_bitexact(ctx, n):
    """Candidates of 11 to 20 columns: the whole QR in one launch when every
    candidate's rows fit one chunk (n <= 4096), launches per step beyond;
    n below the column count stops after n steps."""
    print("case: wide candidates n=%d" % n, flush=True)
    rng = np.random.default_rng(90 + n)
    X, y = make_data(90 + n, n, 4)
    trees = wide_trees(rng, 4, 40)
    od = O.Dataset(X, y)
    tuned = [od.tune(t) for t in trees]
    assert all(len(tu["factors"]) >= 10 for tu in tuned)
    arrs = capi.build_batch([tu["factors"] for tu in tuned])
    h = ctx.upload(X, y)
    coefs, fit = ctx.fitness(h, arrs, mode="ols")
    ctx.free(h)
    for t, tu, c, f in zip(trees, tuned, coefs, fit):
        assert same_bits(c, tu["coefs"]), (t, c, tu["coefs"])
        assert f["rank"] == tu["rank"], t
        assert same_bits([f["e0"], f["e1"]], [tu["fit"][0], tu["fit"][1]]), (t, f, tu["fit"])


@pytest.mark.parametrize("sizes", [(150, 3000), (150, 5000)])
def test_wide_candidates_of_several_datasets_bitexact(ctx, sizes):
    """Wide candidates of two datasets interleaved in one launch: each its own
    dataset's rows, the whole QR in one launch when both fit one chunk, per
    step when one does not."""
    print("case: wide candidates of datasets %s" % (sizes,), flush=True)
    hs, tuned, owner = [], [], []
    for s, n in enumerate(sizes):
        rng = np.random.default_rng(190 + n)
        X, y = make_data(190 + n, n, 3)
        od = O.Dataset(X, y)
        tuned.append([od.tune(t) for t in wide_trees(rng, 3, 20, 10, 16)])
        hs.append(ctx.upload(X, y))
    mixed = []
    for i in range(20):
        for s in range(len(sizes)):
            mixed.append(tuned[s][i]["factors"])
            owner.append((s, i))
    assert all(len(f) >= 10 for f in mixed)
    coefs, fit = ctx.fitness(None, capi.build_batch(mixed), mode="ols", handles=[hs[s] for s, _ in owner])
    for h in hs:
        ctx.free(h)
    for q, (s, i) in enumerate(owner):
        tu = tuned[s][i]
        assert same_bits(coefs[q], tu["coefs"]), (s, i)
        assert fit[q]["rank"] == tu["rank"], (s, i)
        assert same_bits([fit[q]["e0"], fit[q]["e1"]], [tu["fit"][0], tu["fit"][1]]), (s, i)


def test_phase_timing_and_early_flag(ctx):
    """Option phase_timing splits the kernel time of a call by phase; the
    phases of the plain FP64 form (one stream) add up to no more than the
    call's kernel time.  Candidates the factor-0 pre-pass settles carry
    RR_F_EARLY and are sentinels; no other candidate carries it."""
    print("case: phase timing", flush=True)
    X, y = make_data(81, 20000, 3)
    trees = treegen.random_trees(81, 200, 3, 30, 3) + ["ln x0", "+ sqrt x1 x2"]
    od = O.Dataset(X, y)
    tuned = [od.tune(t) for t in trees]
    arrs = capi.build_batch([tu["factors"] for tu in tuned])
    h = ctx.upload(X, y)
    ctx.set_option("phase_timing", 1)
    try:
        coefs, fit = ctx.fitness(h, arrs)
        ph = ctx.phase_ms()
        total = ctx.kernel_ms()
        _, fit32 = ctx.fitness(h, arrs, precision="fp32")
        ph32 = ctx.phase_ms()
    finally:
        ctx.set_option("phase_timing", 0)
    ctx.free(h)
    print("phases:", ph, "total", total, "fp32:", ph32, flush=True)
    for k in ("prepass", "eval", "qr", "score", "host_check", "host_plan"):
        assert ph[k] > 0, k
    assert ph["fp32"] == 0 and ph["host_tail"] >= 0
    assert sum(ph[k] for k in ("prepass", "eval", "qr", "score")) <= total * 1.05 + 0.01
    assert ph32["fp32"] > 0 and ph32["prepass"] > 0
    early = (fit["flags"] & capi.D["RR_F_EARLY"]) != 0
    assert early.any() and np.all(fit["flags"][early] & 1)
    assert np.array_equal(early, (fit32["flags"] & capi.D["RR_F_EARLY"]) != 0)
    ln_x0 = len(trees) - 2
    assert early[ln_x0]     # ln of a negative x0 within the first rows
    for t, tu, c, f in zip(trees, tuned, coefs, fit):   # the flag does not change any result
        assert same_bits(c, tu["coefs"]) and same_bits([f["e0"], f["e1"]], [tu["fit"][0], tu["fit"][1]]), t


def test_graph_replays_same_bits(ctx):
    """Option graph: the plain FP64 form's launches go out as an executable
    graph, kept per launch sequence and updated with each call's arguments.
    Calls of different shapes (OLS batches of different sizes, a RAW batch)
    with eager calls between the replays: every replay gives the eager bits,
    and repeated shapes reuse their executable graph."""
    print("case: graph replays", flush=True)
    X, y = make_data(91, 20000, 4)
    od = O.Dataset(X, y)
    h = ctx.upload(X, y)
    tuned_a = [od.tune(t) for t in treegen.random_trees(91, 300, 3, 40, 4)]
    tuned_b = [od.tune(t) for t in treegen.random_trees(92, 80, 3, 20, 4)]
    arr_a = capi.build_batch([tu["factors"] for tu in tuned_a])
    arr_b = capi.build_batch([tu["factors"] for tu in tuned_b])
    raw = capi.build_batch([[t] for t in treegen.random_trees(93, 200, 3, 30, 4)])
    eager = {"a": ctx.fitness(h, arr_a), "b": ctx.fitness(h, arr_b), "raw": ctx.fitness(h, raw, mode="raw")}
    arrays = {"a": arr_a, "b": arr_b, "raw": raw}
    ctx.set_option("graph", 1)
    try:
        made0, upd0 = ctx.get_option("graph_made"), ctx.get_option("graph_updated")
        for step, name in enumerate(["a", "b", "a", "raw", "a", "b", "raw", "a"]):
            mode = "raw" if name == "raw" else "ols"
            c, f = ctx.fitness(h, arrays[name], mode=mode)
            ce, fe = eager[name]
            assert same_bits(f["e0"], fe["e0"]) and same_bits(f["e1"], fe["e1"]), (step, name)
            if mode == "ols":
                assert all(same_bits(p, q) for p, q in zip(c, ce)), (step, name)
            ctx.set_option("graph", 0)                 # eager work between the replays
            c2, f2 = ctx.fitness(h, arrays[name], mode=mode)
            assert same_bits(f2["e0"], fe["e0"]), (step, name)
            ctx.set_option("graph", 1)
        made, upd = ctx.get_option("graph_made") - made0, ctx.get_option("graph_updated") - upd0
        print("graphs made", made, "updated", upd, flush=True)
        assert made >= 1 and upd >= 1 and made + upd == 8
    finally:
        ctx.set_option("graph", 0)
    for tu, c, f in zip(tuned_a, *eager["a"]):            # and the eager bits are the oracle's
        assert same_bits(c, tu["coefs"]) and same_bits([f["e0"], f["e1"]], [tu["fit"][0], tu["fit"][1]])
    ctx.free(h)


def test_pipelined_calls_same_bits(ctx):
    """rr_fitness_submit / rr_fitness_wait: OLS and RAW batches of different
    sizes, two in flight at a time, give rr_fitness_batch's bits; a third
    submit while two are pending is refused; a synchronous call completes the
    pending ones first; with phase timing on, submit runs the call at once."""
    print("case: pipelined calls n=20000 d=4", flush=True)
    X, y = make_data(91, 20000, 4)
    trees = treegen.random_trees(91, 300, 3, 40, 4)
    od = O.Dataset(X, y)
    facs = [od.tune(t)["factors"] for t in trees]
    h = ctx.upload(X, y)
    big = capi.build_batch(facs)
    small = capi.build_batch(facs[::3][:70])
    raw = capi.build_batch([[t] for t in trees[:100]])
    want = {"big": ctx.fitness(h, big), "small": ctx.fitness(h, small), "raw": ctx.fitness(h, raw, mode="raw")}

    def assert_same(got, exp):
        (cg, fg), (cw, fw) = got, exp
        assert len(cg) == len(cw)
        for a, b in zip(cg, cw):
            assert same_bits(a, b)
        assert fg.tobytes() == fw.tobytes()

    seq = [("big", big, "ols"), ("small", small, "ols"), ("raw", raw, "raw"), ("big", big, "ols"),
           ("small", small, "ols")]
    pending = []
    for name, arrs, mode in seq:
        pending.append((ctx.submit(h, arrs, mode), name))
        if len(pending) == 2:
            ctx.wait()
            r, name0 = pending.pop(0)
            assert_same(r(), want[name0])
    ctx.wait()
    r, name0 = pending.pop(0)
    assert_same(r(), want[name0])
    a = ctx.submit(h, small)
    b = ctx.submit(h, big)
    with pytest.raises(RuntimeError, match="two calls are pending"):
        ctx.submit(h, small)
    assert_same(ctx.fitness(h, small), want["small"])   # completes both pending calls first
    assert_same(a(), want["small"])
    assert_same(b(), want["big"])
    ctx.wait()                                           # each submit is still retired by its wait
    ctx.wait()
    with pytest.raises(RuntimeError, match="no pipelined call"):
        ctx.wait()
    ctx.set_option("phase_timing", 1)
    try:
        r = ctx.submit(h, big)
        assert_same(r(), want["big"])                    # written by the submit itself
        ctx.wait()
    finally:
        ctx.set_option("phase_timing", 0)
    ctx.free(h)


def test_repeated_candidates_take_the_same_results(ctx):
    """Candidates whose factor programs are the same (constants compared by
    value) are launched once and the others take the results: every result is
    still the oracle's, synchronously and pipelined; a constant one bit apart
    is a different factor."""
    print("case: repeated candidates n=9000 d=3", flush=True)
    X, y = make_data(77, 9000, 3)
    trees = ["+ sin x0 * x1 x2", "exp * x0 c0x1p-2", "+ + sin x0 * x1 x2 c0x1.8p+1", "+ sin x0 * x1 x2",
             "exp * x0 c0x1.0000000000001p-2", "* c0x1.4p+1 exp * x0 c0x1p-2", "ln x0", "ln x0",
             "+ sqrt x2 x0", "+ x0 sqrt x2", "+ sqrt x2 x0"]
    od = O.Dataset(X, y)
    tuned = [od.tune(t) for t in trees]
    arrs = capi.build_batch([tu["factors"] for tu in tuned])
    h = ctx.upload(X, y)
    got = [ctx.fitness(h, arrs)]
    r = ctx.submit(h, arrs)
    ctx.wait()
    got.append(r())
    for coefs, fit in got:
        for t, tu, c, f in zip(trees, tuned, coefs, fit):
            assert same_bits(c, tu["coefs"]), t
            assert same_bits([f["e0"], f["e1"]], [tu["fit"][0], tu["fit"][1]]), t
        assert fit[0].tobytes() == fit[2].tobytes() == fit[3].tobytes()
        assert fit[1].tobytes() == fit[5].tobytes()
        assert fit[8].tobytes() == fit[10].tobytes()
    assert tuned[4]["factors"] != tuned[1]["factors"]
    ctx.free(h)
