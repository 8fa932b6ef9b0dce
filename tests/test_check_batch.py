"""rr_check_batch (the validation rr_fitness_batch runs before any launch)
against an independent statement of the bytecode rules of include/rr_fitness.h,
on valid programs compiled from random trees and on random corruptions of
them.  CPU only: the check needs no device."""
import os

import numpy as np
import pytest

import capi
import treegen

D = capi.D
MAX_STACK = D["RR_MAX_STACK"]
VALID_UNARY = {D["RR_OP_SIN"], D["RR_OP_COS"], D["RR_OP_LN"], D["RR_OP_EXP"], D["RR_OP_SQRT"], D["RR_OP_SQR"]}


def rules_ok(code, d, n_consts):
    """The bytecode rules, instruction by instruction."""
    if len(code) == 0:
        return False
    sp, acc = 0, False
    for ins in code:
        op, arg = int(ins) & 0xFF, int(ins) >> 8
        if op >= D["RR_OP_BIN"]:
            bop, mode = (op - D["RR_OP_BIN"]) >> 3, (op - D["RR_OP_BIN"]) & 7
            if bop > D["RR_BOP_MAX"] or mode > D["RR_M_CR"] or not acc:
                return False
            if mode in (D["RR_M_S"], D["RR_M_SR"]):
                if sp == 0:
                    return False
                sp -= 1
            if mode in (D["RR_M_V"], D["RR_M_VR"]) and arg >= d:
                return False
            if mode in (D["RR_M_C"], D["RR_M_CR"]) and arg >= n_consts:
                return False
        elif op == D["RR_OP_LDV"]:
            if arg >= d:
                return False
            acc = True
        elif op == D["RR_OP_LDC"]:
            if arg >= n_consts:
                return False
            acc = True
        elif op == D["RR_OP_PUSH"]:
            if not acc:
                return False
            sp += 1
            if sp > MAX_STACK:
                return False
            acc = False
        elif op in VALID_UNARY:
            if not acc:
                return False
        else:
            return False
    return sp == 0 and acc


def one_program(code, n_consts):
    code = np.asarray(code, dtype=np.uint32)
    return {"cand_start": np.array([0, 1], dtype=np.int32), "prog_start": np.array([0, len(code)], dtype=np.int32),
            "code": code if len(code) else np.zeros(1, dtype=np.uint32),
            "consts": np.zeros(max(n_consts, 1)), "n_consts": n_consts, "n_cands": 1}


def corrupt(code, rng):
    code = [int(c) for c in code]
    kind = rng.integers(6)
    i = int(rng.integers(len(code)))
    if kind == 0:      # another opcode, any byte
        code[i] = (code[i] & ~0xFF) | int(rng.integers(256))
    elif kind == 1:    # another operand
        code[i] = (code[i] & 0xFF) | (int(rng.integers(0, 24)) << 8)
    elif kind == 2:    # drop an instruction
        del code[i]
    elif kind == 3:    # duplicate one
        code.insert(i, code[i])
    elif kind == 4:    # an extra push
        code.insert(i, D["RR_OP_PUSH"])
    else:              # truncate
        code = code[:i]
    return code


def test_valid_programs_accepted():
    trees = treegen.random_trees(7, 400, 1, 60, 6)
    for t in trees:
        arrs = capi.build_batch([[t]])
        rc, msg = capi.check_batch(arrs, "raw", 6)
        assert rc == 0, (t, msg)
        assert rules_ok(arrs["code"], 6, arrs["n_consts"])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_check_matches_the_rules_on_corrupted_programs(seed):
    rng = np.random.default_rng(seed)
    trees = treegen.random_trees(seed, 300, 1, 40, 5)
    agree = accepted = 0
    for t in trees:
        base = capi.build_batch([[t]])
        nk = base["n_consts"]
        code = list(base["code"][: base["prog_start"][1]])
        for _ in range(8):
            bad = corrupt(code, rng)
            d = int(rng.integers(1, 8))
            want = rules_ok(bad, d, nk)
            rc, msg = capi.check_batch(one_program(bad, nk), "raw", d)
            assert (rc == 0) == want, (t, bad, d, nk, rc, msg)
            if rc:
                assert "program 0" in msg and rc in (D["RR_E_ARG"], D["RR_E_LIMIT"])
            agree += 1
            accepted += want
    assert 0 < accepted < agree


def test_structure_errors():
    arrs = capi.build_batch([["x0", "x1"], ["sin x0"]])
    assert capi.check_batch(arrs, "ols", 2)[0] == 0
    assert capi.check_batch(arrs, "raw", 2)[0] == D["RR_E_ARG"]          # RAW takes one program per candidate
    assert capi.check_batch(arrs, "ols", 1)[0] == D["RR_E_ARG"]          # x1 with one feature
    bad = dict(arrs, cand_start=np.array([0, 2, 1], dtype=np.int32))
    assert "monotone" in capi.check_batch(bad, "ols", 2)[1]
    bad = dict(arrs, prog_start=np.array([0, 1, 0, 3], dtype=np.int32))
    assert "monotone" in capi.check_batch(bad, "ols", 2)[1]
    bad = dict(arrs, cand_start=np.array([1, 2, 3], dtype=np.int32))
    assert capi.check_batch(bad, "ols", 2)[0] == D["RR_E_ARG"]
    wide = capi.build_batch([["x0"] * D["RR_MAX_COLS"]])
    assert capi.check_batch(wide, "ols", 1)[0] == D["RR_E_LIMIT"]


def test_large_batch_reports_first_invalid_program():
    """A batch big enough to be checked on several host threads: valid as a
    whole, and with invalid programs planted late and early, the one reported
    is the first in batch order."""
    trees = treegen.random_trees(9, 4000, 10, 60, 4)
    arrs = capi.build_batch([[t, "x0"] for t in trees])
    assert arrs["prog_start"][-1] > 4 * 16384        # several ranges of the split
    assert capi.check_batch(arrs, "ols", 4) == (0, "")
    code = arrs["code"].copy()
    n_progs = len(arrs["prog_start"]) - 1
    late, early = n_progs - 5, 3 * n_progs // 10      # in different ranges
    for p in (late, early):
        code[arrs["prog_start"][p]] = D["RR_OP_PUSH"]  # push on an empty accumulator
    rc, msg = capi.check_batch(dict(arrs, code=code), "ols", 4)
    assert rc == D["RR_E_ARG"] and msg.endswith("(program %d)" % early), msg
    code[arrs["prog_start"][early]] = arrs["code"][arrs["prog_start"][early]]
    rc, msg = capi.check_batch(dict(arrs, code=code), "ols", 4)
    assert rc == D["RR_E_ARG"] and msg.endswith("(program %d)" % late), msg


def test_process_exits_after_the_host_pool_was_used():
    """A large batch check runs on the library's host worker threads, which
    live until the process ends: a process that used them exits promptly."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import capi, treegen\n"
        "arrs = capi.build_batch([[t, 'x0'] for t in treegen.random_trees(9, 4000, 10, 60, 4)])\n"
        "assert capi.check_batch(arrs, 'ols', 4) == (0, '')\n"
        "print('checked', flush=True)\n" % os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "checked" in r.stdout, r.stderr[-2000:]
