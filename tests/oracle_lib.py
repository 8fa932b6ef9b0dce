"""ctypes binding of the CPU oracle (oracle/_build/liborc.so) for the tests.

The oracle is the checker: tests compare the product (the HIP path behind the
C-ABI and the rils_rols_cpp module) against it.  Nothing on the product path
imports this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("RR_ORACLE_LIB") or os.path.join(ROOT, "oracle", "_build", "liborc.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_math.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int64]
        L.orc_canon_sum.argtypes = [dp, ctypes.c_int64, ctypes.c_int64]
        L.orc_canon_sum.restype = ctypes.c_double
        L.orc_chunk_sum.argtypes = [dp, ctypes.c_int64, ctypes.c_int64]
        L.orc_chunk_sum.restype = ctypes.c_double
        L.orc_dataset.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_int]
        L.orc_dataset_free.argtypes = [ctypes.c_int]
        L.orc_dataset_stats.argtypes = [ctypes.c_int, dp, dp]
        L.orc_eval.argtypes = [ctypes.c_int, ctypes.c_char_p, dp]
        L.orc_score.argtypes = [ctypes.c_int, ctypes.c_char_p, dp]
        L.orc_tune.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                               ctypes.POINTER(ctypes.c_void_p), dp, ctypes.c_int, ip, ip, dp]
        L.orc_qr_solve.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_int, dp, ip, ctypes.c_int]
        L.orc_set_qr_direct.argtypes = [ctypes.c_int]
        L.orc_rewrite.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.orc_rewrite.restype = ctypes.c_void_p
        L.orc_all_candidates.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_all_candidates.restype = ctypes.c_void_p
        L.orc_fit.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_fit.restype = ctypes.c_void_p
        L.orc_batch_tune.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, dp]
        L.orc_accept.argtypes = [dp, dp, ctypes.c_double, ctypes.c_int, dp]
        L.orc_batch_tune.restype = ctypes.c_double
        L.orc_local_search.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int]
        L.orc_local_search.restype = ctypes.c_void_p
        L.orc_relevant_features.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ip]
        L.orc_batch_tune_full.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, dp, ctypes.c_int, ip, ip, dp,
                                          ctypes.POINTER(ctypes.c_void_p)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _take_str(p):
    if not p:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())
    s = ctypes.string_at(p).decode()
    lib().orc_free(p)
    return s


MATH_OPS = {"sin": 0, "cos": 1, "exp": 2, "log": 3, "sqrt": 4}


def math(op, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    rc = lib().orc_math(MATH_OPS[op], _dp(x), _dp(out), x.size)
    assert rc == 0
    return out


def canon_sum(v, lo=0, hi=None):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_canon_sum(_dp(v), lo, v.size if hi is None else hi)


def chunk_sum(v, lo=0, hi=None):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_chunk_sum(_dp(v), lo, v.size if hi is None else hi)


class Dataset:
    """Oracle-side dataset; X is (n, d) or feature-major (d, n) with fm=True."""

    def __init__(self, X, y, fm=False):
        X = np.asarray(X, dtype=np.float64)
        Xfm = np.ascontiguousarray(X if fm else X.T)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.n = self.y.size
        self.d = Xfm.shape[0]
        self.h = lib().orc_dataset(_dp(Xfm), _dp(self.y), self.n, self.d)

    def __del__(self):
        try:
            lib().orc_dataset_free(self.h)
        except Exception:
            pass

    def stats(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        lib().orc_dataset_stats(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def eval(self, prefix):
        out = np.empty(self.n)
        if lib().orc_eval(self.h, prefix.encode(), _dp(out)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return out

    def score(self, prefix):
        out = np.empty(3)
        if lib().orc_score(self.h, prefix.encode(), _dp(out)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return out[0], out[1], int(out[2])

    def tune(self, prefix, cap=4096):
        ols = ctypes.c_void_p()
        fac = ctypes.c_void_p()
        coefs = np.empty(cap)
        k = ctypes.c_int()
        rank = ctypes.c_int()
        fit = np.empty(3)
        rc = lib().orc_tune(self.h, prefix.encode(), ctypes.byref(ols), ctypes.byref(fac), _dp(coefs), cap,
                            ctypes.byref(k), ctypes.byref(rank), _dp(fit))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        fstr = _take_str(fac.value) if fac.value else ""
        return {
            "ols": _take_str(ols.value),
            "factors": [f for f in fstr.split("\n") if f],
            "coefs": coefs[:k.value].copy(),
            "rank": rank.value,
            "fit": (fit[0], fit[1], int(fit[2])),
        }

    def batch_tune_full(self, prefixes, threads=1, cap=16):
        """tune() for many trees on `threads` host threads: a list of dicts
        with factors, coefs, rank and fit (no OLS-tree string)."""
        m = len(prefixes)
        coefs = np.empty(cap * max(m, 1))
        k = np.empty(max(m, 1), dtype=np.int32)
        rank = np.empty(max(m, 1), dtype=np.int32)
        fit = np.empty(3 * max(m, 1))
        fac = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int)
        rc = lib().orc_batch_tune_full(self.h, "\n".join(prefixes).encode(), threads, _dp(coefs), cap,
                                       k.ctypes.data_as(ip), rank.ctypes.data_as(ip), _dp(fit), ctypes.byref(fac))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        facs = _take_str(fac.value).split("\t") if fac.value else [""] * m
        return [{"factors": [f for f in facs[i].split("\n") if f], "coefs": coefs[cap * i:cap * i + k[i]].copy(),
                 "rank": int(rank[i]), "fit": (fit[3 * i], fit[3 * i + 1], int(fit[3 * i + 2]))} for i in range(m)]

    def batch_tune(self, prefixes, threads=1):
        fit = np.empty(3 * len(prefixes))
        secs = lib().orc_batch_tune(self.h, "\n".join(prefixes).encode(), threads, _dp(fit))
        return secs, fit.reshape(-1, 3)


def set_qr_direct(on):
    """Timing only: every least-squares solve of the oracle takes Eigen's
    order on the data (the reference's algorithm) instead of the canonical
    chunked form; results are then not the GPU's bits."""
    lib().orc_set_qr_direct(int(bool(on)))


def qr_solve(A, b, direct=False):
    """Least squares of the OLS form: the canonical QR (chunked up to 9
    columns), or with direct=True Eigen's order on the data for any k."""
    A = np.asarray(A, dtype=np.float64)
    n, k = A.shape
    Acm = np.ascontiguousarray(A.T)
    b = np.ascontiguousarray(b, dtype=np.float64)
    coefs = np.empty(k)
    rank = ctypes.c_int()
    lib().orc_qr_solve(_dp(Acm), _dp(b), n, k, _dp(coefs), ctypes.byref(rank), int(direct))
    return coefs, rank.value


REWRITE_OPS = {"to_string": 0, "expand": 1, "simplify": 2, "normalize": 3, "size": 4,
               "candidate_rounds": 5, "accept_rounds": 6}


def rewrite(prefix, op):
    return _take_str(lib().orc_rewrite(prefix.encode(), REWRITE_OPS[op]))


def to_string(prefix):
    return rewrite(prefix, "to_string")


def all_candidates(prefix, nvars, classification=False, local_search=True):
    s = _take_str(lib().orc_all_candidates(prefix.encode(), nvars, int(classification), int(local_search)))
    return [c for c in s.split("\n") if c]


def fit(X, y, classification=False, max_fit_calls=100000, max_time=100000, complexity_penalty=0.001,
        max_complexity=50, sample_size=1.0, random_state=0):
    X = np.ascontiguousarray(X, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n, d = X.shape
    rep = _take_str(lib().orc_fit(_dp(X), _dp(y), n, d, int(classification), int(max_fit_calls), int(max_time),
                                  float(complexity_penalty), int(max_complexity), float(sample_size),
                                  int(random_state)))
    return parse_fit_report(rep)


def local_search(X, y, prefix, max_fit_calls=100000, complexity_penalty=0.001, max_complexity=50):
    """One local_search of the oracle's search from `prefix` (every feature
    allowed): the report of parse_fit_report."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n, d = X.shape
    rep = _take_str(lib().orc_local_search(_dp(X), _dp(y), n, d, prefix.encode(), int(max_fit_calls),
                                           float(complexity_penalty), int(max_complexity)))
    return parse_fit_report(rep)


def relevant_features(X, y):
    """Feature indices the oracle's search keeps (all when d <= 200)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n, d = X.shape
    out = np.empty(max(d, 1), dtype=np.int32)
    k = lib().orc_relevant_features(_dp(X), _dp(y), n, d, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return out[:k].tolist()


def parse_fit_report(rep):
    out = {"moves": []}
    for line in rep.splitlines():
        parts = line.split("\t")
        if parts[0] == "model":
            out["model"] = parts[1]
        elif parts[0] == "prefix":
            out["prefix"] = parts[1]
        elif parts[0] == "fit_calls":
            out["fit_calls"] = int(parts[1])
        elif parts[0] == "final":
            out["final"] = (float.fromhex(parts[1]), float.fromhex(parts[2]), int(parts[3]))
        elif parts[0] == "move":
            out["moves"].append((int(parts[1]), int(parts[2]), float.fromhex(parts[3]), float.fromhex(parts[4]),
                                 int(parts[5]), parts[6]))
    return out


def accept(a, b, complexity_penalty=0.001, max_complexity=50):
    """(fitness_value(a), compare_fitness(a, b), dominates(a, b)) of the
    oracle's search for fitness tuples a, b = (1-R2, RMSE, size)."""
    av = np.asarray(a, dtype=np.float64)
    bv = np.asarray(b, dtype=np.float64)
    out = np.empty(3)
    assert lib().orc_accept(_dp(av), _dp(bv), complexity_penalty, max_complexity, _dp(out)) == 0
    return float(out[0]), int(out[1]), bool(out[2])

