"""Seeded random expression trees in the prefix exchange format.

Ops and leaf distribution follow the batched candidate-eval workload: internal
ops uniform over {+,-,*,/,sin,cos,ln,exp,sqrt,sqr}, leaves 50/50 a variable or
a constant drawn from N(0,1), exact node counts.
"""
import numpy as np

UN = ["sin", "cos", "ln", "exp", "sqrt", "sqr"]
BI = ["+", "-", "*", "/"]
OPS = BI + UN
CLS_BI = ["<", ">", "==", "!=", "min", "max"]


def const_tok(v):
    return "c" + float(v).hex()


def gen(rng, size, nvars, ops=OPS, const_p=0.5):
    if size <= 1:
        if rng.random() < const_p:
            return [const_tok(rng.normal())]
        return ["x%d" % rng.integers(nvars)]
    if size == 2:
        op = UN[rng.integers(len(UN))]
    else:
        op = ops[rng.integers(len(ops))]
    if op in UN:
        return [op] + gen(rng, size - 1, nvars, ops, const_p)
    left = int(rng.integers(1, size - 1))
    return [op] + gen(rng, left, nvars, ops, const_p) + gen(rng, size - 1 - left, nvars, ops, const_p)


def random_tree(rng, size, nvars, ops=OPS, const_p=0.5):
    return " ".join(gen(rng, size, nvars, ops, const_p))


def random_trees(seed, count, lo, hi, nvars, ops=OPS, const_p=0.5):
    rng = np.random.default_rng(seed)
    return [random_tree(rng, int(rng.integers(lo, hi + 1)), nvars, ops, const_p) for _ in range(count)]
