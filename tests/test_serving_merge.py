"""Batch merging of the serving loop (CPU): several runs' candidate batches
become one launch; constant operands must be renumbered into the merged pool
so every program still reads its own constants."""
import numpy as np

import capi


def decode(arrs):
    """Per candidate, per program: the list of (opcode, operand) with constant
    operands replaced by the constant's value."""
    D = capi.D
    out = []
    cs, ps, code, K = arrs["cand_start"], arrs["prog_start"], arrs["code"], arrs["consts"]
    for c in range(len(cs) - 1):
        progs = []
        for p in range(cs[c], cs[c + 1]):
            ins = []
            for w in code[ps[p]:ps[p + 1]]:
                w = int(w)
                op, arg = w & 0xff, w >> 8
                const = op == D["RR_OP_LDC"] or (op >= D["RR_OP_BIN"] and ((op - D["RR_OP_BIN"]) & 7) >= D["RR_M_C"])
                ins.append((op, float(K[arg]) if const else arg))
            progs.append(ins)
        out.append(progs)
    return out


def test_merge_keeps_every_program():
    import rils_rols_cpp
    b1 = capi.build_batch([["* x0 c0x1.8p+1", "sin + x1 c-0x1p-2"], ["x0"], []])
    b2 = capi.build_batch([["/ c0x1.4p+2 x1"], ["exp * x0 c0x1.999999999999ap-3", "x1", "sqr - x0 c0x1p+0"]])
    b3 = capi.build_batch([["max c0x1p+3 x0"]])
    parts = []
    for b, h in [(b1, 7), (b2, 3), (b3, 7)]:
        parts.append((list(b["cand_start"]), list(b["prog_start"]), [int(w) for w in b["code"]], list(b["consts"]), h))
    cs, ps, code, consts, hs = rils_rols_cpp._merge_batches(parts)
    merged = {"cand_start": np.array(cs), "prog_start": np.array(ps), "code": np.array(code, dtype=np.uint32),
              "consts": np.array(consts)}
    want = decode(b1) + decode(b2) + decode(b3)
    assert decode(merged) == want
    assert hs == [7] * 3 + [3] * 2 + [7]
    assert len(consts) == len(b1["consts"]) + len(b2["consts"]) + len(b3["consts"])


def test_footprint_is_padded_dataset_size():
    import rils_rols_cpp
    assert rils_rols_cpp.Server.footprint(1, 1) == 4096 * 12 + 4096 * 16
    assert rils_rols_cpp.Server.footprint(100000, 10) == 102400 * 10 * 12 + 102400 * 16
