"""ctypes binding of the C ABI (include/rr_fitness.h, librr_hip.so) and a
test-side compiler from prefix trees to the ABI's bytecode.

The compiler here is written independently of the product's C++ one
(rils-rols_amd/csrc/tree.cpp) so that the C-ABI tests do not depend on it.
"""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "rils-rols_amd", "librr_hip.so")
HEADER = os.path.join(ROOT, "include", "rr_fitness.h")


def header_defines():
    out = {}
    for m in re.finditer(r"#define\s+(RR_\w+)\s+\(?(-?\d+)u?\)?", open(HEADER).read()):
        out[m.group(1)] = int(m.group(2))
    return out


D = header_defines()


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(rr_\w+)\s*\(", text, re.M)))


class Batch(ctypes.Structure):
    _fields_ = [("n_cands", ctypes.c_int32), ("cand_start", ctypes.POINTER(ctypes.c_int32)),
                ("prog_start", ctypes.POINTER(ctypes.c_int32)), ("code", ctypes.POINTER(ctypes.c_uint32)),
                ("consts", ctypes.POINTER(ctypes.c_double)), ("n_consts", ctypes.c_int32)]


class Fit(ctypes.Structure):
    _fields_ = [("e0", ctypes.c_double), ("e1", ctypes.c_double), ("err", ctypes.c_double),
                ("rank", ctypes.c_int32), ("flags", ctypes.c_int32)]


FIT_DTYPE = np.dtype([("e0", "<f8"), ("e1", "<f8"), ("err", "<f8"), ("rank", "<i4"), ("flags", "<i4")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("librr_hip.so not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.rr_last_error.restype = ctypes.c_char_p
        L.rr_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.rr_destroy.argtypes = [vp]
        L.rr_dataset_upload.argtypes = [vp, dp, dp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.rr_dataset_free.argtypes = [vp, ctypes.c_int32]
        L.rr_dataset_stats.argtypes = [vp, ctypes.c_int32, dp, dp]
        L.rr_fitness_batch.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Batch), ctypes.c_int32, ctypes.c_int32, dp,
                                       vp]
        L.rr_fitness_multi.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(Batch), ctypes.c_int32,
                                       ctypes.c_int32, dp, vp]
        L.rr_predict.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32, dp,
                                 ctypes.c_int32, dp]
        fp = ctypes.POINTER(ctypes.c_float)
        L.rr_predict_fp32.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32, dp,
                                      ctypes.c_int32, fp, fp]
        L.rr_autotune.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Batch), ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_int32, dp]
        L.rr_set_option.argtypes = [vp, ctypes.c_char_p, ctypes.c_int64]
        L.rr_get_option.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
        L.rr_last_kernel_ms.argtypes = [vp, dp]
        L.rr_last_phase_ms.argtypes = [vp, ctypes.c_int32, dp]
        L.rr_fitness_batch_noalloc.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Batch), ctypes.c_int32, dp, vp]
        L.rr_check_batch.argtypes = [ctypes.POINTER(Batch), ctypes.c_int32, ctypes.c_int32]
        L.rr_fitness_submit.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Batch), ctypes.c_int32, dp, vp]
        L.rr_fitness_wait.argtypes = [vp]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise RuntimeError("rr error %d: %s" % (rc, lib().rr_last_error().decode()))


# ----------------------------------------------------------------- prefix trees
UNARY = {"sin": D["RR_OP_SIN"], "cos": D["RR_OP_COS"], "ln": D["RR_OP_LN"], "exp": D["RR_OP_EXP"],
         "sqrt": D["RR_OP_SQRT"], "sqr": D["RR_OP_SQR"]}
BINARY = {"+": D["RR_BOP_ADD"], "-": D["RR_BOP_SUB"], "*": D["RR_BOP_MUL"], "/": D["RR_BOP_DIV"],
          "<": D["RR_BOP_LT"], ">": D["RR_BOP_GT"], "==": D["RR_BOP_EQ"], "!=": D["RR_BOP_NEQ"],
          "min": D["RR_BOP_MIN"], "max": D["RR_BOP_MAX"]}


def parse(prefix):
    toks = prefix.split()
    pos = [0]

    def rec():
        t = toks[pos[0]]
        pos[0] += 1
        if t in UNARY:
            return (t, rec())
        if t in BINARY:
            a = rec()
            return (t, a, rec())
        if t.startswith("x"):
            return ("x", int(t[1:]))
        if t.startswith("c"):
            return ("c", float.fromhex(t[1:]) if "0x" in t else float(t[1:]))
        raise ValueError(t)

    tree = rec()
    assert pos[0] == len(toks)
    return tree


def _leaf(t):
    return t[0] in ("x", "c")


def _need(t):
    if _leaf(t):
        return 0
    if t[0] in UNARY:
        return _need(t[1])
    a, b = t[1], t[2]
    if _leaf(b):
        return _need(a)
    if _leaf(a):
        return _need(b)
    na, nb = _need(a), _need(b)
    return na + 1 if na == nb else max(na, nb)


class Compiler:
    """Accumulator-machine code generation with Sethi-Ullman ordering."""

    def __init__(self):
        self.code = []
        self.consts = []

    def _kidx(self, v):
        self.consts.append(v)
        return len(self.consts) - 1

    def _emit(self, op, arg=0):
        self.code.append((arg << 8) | op)

    def _leaf_operand(self, t):
        if t[0] == "x":
            return "V", t[1]
        return "C", self._kidx(t[1])

    def gen(self, t):
        if t[0] == "x":
            self._emit(D["RR_OP_LDV"], t[1])
        elif t[0] == "c":
            self._emit(D["RR_OP_LDC"], self._kidx(t[1]))
        elif t[0] in UNARY:
            self.gen(t[1])
            self._emit(UNARY[t[0]])
        else:
            bop = BINARY[t[0]]
            a, b = t[1], t[2]
            base = D["RR_OP_BIN"] + 8 * bop
            if _leaf(b):
                self.gen(a)
                kind, arg = self._leaf_operand(b)
                self._emit(base + (D["RR_M_V"] if kind == "V" else D["RR_M_C"]), arg)
            elif _leaf(a):
                self.gen(b)
                kind, arg = self._leaf_operand(a)
                self._emit(base + (D["RR_M_VR"] if kind == "V" else D["RR_M_CR"]), arg)
            elif _need(a) >= _need(b):
                self.gen(a)
                self._emit(D["RR_OP_PUSH"])
                self.gen(b)
                self._emit(base + D["RR_M_S"])
            else:
                self.gen(b)
                self._emit(D["RR_OP_PUSH"])
                self.gen(a)
                self._emit(base + D["RR_M_SR"])


def build_batch(cands):
    """cands: list of lists of prefix strings (programs per candidate)."""
    comp = Compiler()
    cand_start = [0]
    prog_start = [0]
    for progs in cands:
        for p in progs:
            comp.gen(parse(p))
            prog_start.append(len(comp.code))
        cand_start.append(cand_start[-1] + len(progs))
    arrs = {
        "cand_start": np.array(cand_start, dtype=np.int32),
        "prog_start": np.array(prog_start, dtype=np.int32),
        "code": np.array(comp.code if comp.code else [0], dtype=np.uint32),
        "consts": np.array(comp.consts if comp.consts else [0.0], dtype=np.float64),
        "n_consts": len(comp.consts),
        "n_cands": len(cands),
    }
    return arrs


def _as_batch(arrs):
    b = Batch()
    b.n_cands = arrs["n_cands"]
    b.cand_start = arrs["cand_start"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    b.prog_start = arrs["prog_start"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    b.code = arrs["code"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    b.consts = arrs["consts"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    b.n_consts = arrs["n_consts"]
    return b


def check_batch(arrs, mode="ols", n_features=0):
    """rr_check_batch: (return code, message); needs no device."""
    m = D["RR_MODE_OLS"] if mode == "ols" else D["RR_MODE_RAW"]
    rc = lib().rr_check_batch(ctypes.byref(_as_batch(arrs)), m, n_features)
    return rc, (lib().rr_last_error().decode() if rc else "")


def precision_code(name):
    """RR_FP64 for "fp64", RR_FP32_REFINED for "fp32", RR_FP32_CHECKED for "fp32c"."""
    return {"fp64": D["RR_FP64"], "fp32": D["RR_FP32_REFINED"], "fp32c": D["RR_FP32_CHECKED"]}[name]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    def __init__(self, device=0):
        self.p = ctypes.c_void_p()
        check(lib().rr_create(device, ctypes.byref(self.p)))

    def close(self):
        if self.p:
            lib().rr_destroy(self.p)
            self.p = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def upload(self, X, y, fm=False):
        X = np.asarray(X, dtype=np.float64)
        Xfm = np.ascontiguousarray(X if fm else X.T)
        y = np.ascontiguousarray(y, dtype=np.float64)
        h = ctypes.c_int32()
        check(lib().rr_dataset_upload(self.p, _dp(Xfm), _dp(y), y.size, Xfm.shape[0], ctypes.byref(h)))
        return h.value

    def free(self, h):
        check(lib().rr_dataset_free(self.p, h))

    def stats(self, h):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().rr_dataset_stats(self.p, h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_option(self, name, value):
        check(lib().rr_set_option(self.p, name.encode(), int(value)))

    def kernel_ms(self):
        v = ctypes.c_double()
        check(lib().rr_last_kernel_ms(self.p, ctypes.byref(v)))
        return v.value

    def phase_ms(self):
        """rr_last_phase_ms for every phase, by name."""
        out = {}
        for name in ("PREPASS", "EVAL", "QR", "SCORE", "FP32", "HOST_CHECK", "HOST_PLAN", "HOST_TAIL"):
            v = ctypes.c_double()
            check(lib().rr_last_phase_ms(self.p, D["RR_PHASE_" + name], ctypes.byref(v)))
            out[name.lower()] = v.value
        return out

    def fitness(self, h, arrs, mode="ols", precision="fp64", handles=None):
        """Score a batch against dataset h, or candidate c against handles[c]
        (rr_fitness_multi) when handles is given."""
        m = D["RR_MODE_OLS"] if mode == "ols" else D["RR_MODE_RAW"]
        pr = precision_code(precision)
        nc = arrs["n_cands"]
        ncoef = int(arrs["cand_start"][-1]) + nc
        coefs = np.zeros(max(ncoef, 1))
        fit = np.zeros(max(nc, 1), dtype=FIT_DTYPE)
        b = _as_batch(arrs)
        if handles is None:
            check(lib().rr_fitness_batch(self.p, h, ctypes.byref(b), m, pr, _dp(coefs),
                                         fit.ctypes.data_as(ctypes.c_void_p)))
        else:
            hs = np.ascontiguousarray(handles, dtype=np.int32)
            check(lib().rr_fitness_multi(self.p, hs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(b), m,
                                         pr, _dp(coefs), fit.ctypes.data_as(ctypes.c_void_p)))
        return self._split(arrs, coefs, fit)

    def fitness_noalloc(self, h, arrs, mode="ols"):
        """rr_fitness_batch_noalloc (plain FP64, no allocation)."""
        m = D["RR_MODE_OLS"] if mode == "ols" else D["RR_MODE_RAW"]
        nc = arrs["n_cands"]
        coefs = np.zeros(max(int(arrs["cand_start"][-1]) + nc, 1))
        fit = np.zeros(max(nc, 1), dtype=FIT_DTYPE)
        check(lib().rr_fitness_batch_noalloc(self.p, h, ctypes.byref(_as_batch(arrs)), m, _dp(coefs),
                                             fit.ctypes.data_as(ctypes.c_void_p)))
        return self._split(arrs, coefs, fit)

    def submit(self, h, arrs, mode="ols"):
        """rr_fitness_submit (plain FP64, pipelined): returns a handle whose
        result() is valid after the matching wait()."""
        m = D["RR_MODE_OLS"] if mode == "ols" else D["RR_MODE_RAW"]
        nc = arrs["n_cands"]
        coefs = np.zeros(max(int(arrs["cand_start"][-1]) + nc, 1))
        fit = np.zeros(max(nc, 1), dtype=FIT_DTYPE)
        check(lib().rr_fitness_submit(self.p, h, ctypes.byref(_as_batch(arrs)), m, _dp(coefs),
                                      fit.ctypes.data_as(ctypes.c_void_p)))
        return lambda: self._split(arrs, coefs, fit)

    def wait(self):
        check(lib().rr_fitness_wait(self.p))

    def _split(self, arrs, coefs, fit):
        nc = arrs["n_cands"]
        per = []
        cs = arrs["cand_start"]
        for c in range(nc):
            o = int(cs[c]) + c
            per.append(coefs[o:o + int(cs[c + 1] - cs[c]) + 1])
        return per, fit[:nc]

    def predict(self, h, prefix, n):
        arrs = build_batch([[prefix]])
        out = np.empty(n)
        code = arrs["code"]
        check(lib().rr_predict(self.p, h, code.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(code),
                               _dp(arrs["consts"]), arrs["n_consts"], _dp(out)))
        return out

    def predict_fp32(self, h, prefix, n):
        """FP32 values and their per-sample bounds (inf: uncertain)."""
        arrs = build_batch([[prefix]])
        v = np.empty(n, dtype=np.float32)
        e = np.empty(n, dtype=np.float32)
        code = arrs["code"]
        fp = ctypes.POINTER(ctypes.c_float)
        check(lib().rr_predict_fp32(self.p, h, code.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(code),
                                    _dp(arrs["consts"]), arrs["n_consts"], v.ctypes.data_as(fp), e.ctypes.data_as(fp)))
        return v, e

    def autotune(self, h, arrs, mode="ols", precision="fp64", reps=2):
        m = D["RR_MODE_OLS"] if mode == "ols" else D["RR_MODE_RAW"]
        pr = precision_code(precision)
        b = _as_batch(arrs)
        best = np.zeros(1)
        check(lib().rr_autotune(self.p, h, ctypes.byref(b), m, pr, reps, _dp(best)))
        return float(best[0])

    def get_option(self, name):
        v = ctypes.c_int64()
        check(lib().rr_get_option(self.p, name.encode(), ctypes.byref(v)))
        return v.value

