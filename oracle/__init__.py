"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (oracle/oracle.cpp) is a single-threaded C++ restatement of the
reference V:N:M reorder path.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module, and only as the checker
or as the timed CPU baseline; the product library never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
CLI_PATH = os.path.join(_HERE, "_build", "spmm_oracle")
REF_DRIVER = os.path.join(_HERE, "_ref", "ref_driver")

OK, E_IO, E_FORMAT, E_ARG, E_OOM, E_UB_PATH = 0, -1, -2, -3, -5, -6


class Coo(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int32), ("ncols", ctypes.c_int32),
                ("nnz", ctypes.c_int64),
                ("row", ctypes.POINTER(ctypes.c_int32)),
                ("col", ctypes.POINTER(ctypes.c_int32)),
                ("val", ctypes.POINTER(ctypes.c_float))]


class SwapRec(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("round", "primary", "target", "v1", "v2", "lgain", "rgain",
                 "primary_cnt_after", "target_cnt_after")]


class HybridStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("nrows", "nnz", "density", "init_size", "final_size",
                 "init_cnt", "final_cnt", "improve_rate")] + \
               [(n, ctypes.c_int32) for n in
                ("init_mb", "final_mb", "init_vnm", "final_vnm",
                 "nm_rounds", "vnm_rounds", "total_rounds")]

    def as_tuple(self):
        return tuple(getattr(self, n) for n, _ in self._fields_)


def build(quiet=True):
    """Compile the oracle with make (g++)."""
    out = subprocess.run(["make", "-C", _HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        i32p = P(ctypes.c_int32)
        L.orc_mtx_read.argtypes = [ctypes.c_char_p, ctypes.c_int, P(Coo)]
        L.orc_mtx_write.argtypes = [ctypes.c_char_p, P(Coo)]
        L.orc_coo_from_tuples.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                          i32p, i32p, P(ctypes.c_float), ctypes.c_int, P(Coo)]
        L.orc_coo_free.argtypes = [P(Coo)]
        L.orc_coo_free.restype = None
        L.orc_locate.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, i32p, i32p]
        L.orc_swap_score.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
        L.orc_pq_pop_order.argtypes = [i32p, ctypes.c_int, i32p]
        L.orc_nm_reorder.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     i32p, P(SwapRec), ctypes.c_int64, P(ctypes.c_int64)]
        L.orc_venom_eval.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, i32p, i32p]
        L.orc_venom_reorder.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        L.orc_hybrid.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, P(HybridStats)]
        L.orc_find_best_sched.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
        L.orc_retained_nnz.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       P(ctypes.c_int64)]
        L.orc_spmm_vnm_f64.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, P(ctypes.c_float),
                                       ctypes.c_int, P(ctypes.c_double)]
        L.orc_spmm_vnm_f64_mt.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, P(ctypes.c_float),
                                          ctypes.c_int, P(ctypes.c_double), ctypes.c_int]
        L.orc_quant_cpp_write.argtypes = [ctypes.c_float]
        L.orc_quant_cpp_write.restype = ctypes.c_float
        L.orc_quant_py_roundtrip.argtypes = [ctypes.c_float]
        L.orc_quant_py_roundtrip.restype = ctypes.c_float
        L.orc_ub_events.restype = ctypes.c_long
        L.orc_reload.argtypes = [P(Coo), ctypes.c_int]
        L.orc_nm_reorder_sample.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, i32p]
        L.orc_allpairs_reorder.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           i32p, P(ctypes.c_int64)]
        L.orc_swap_score_all.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p]
        f64p = P(ctypes.c_double)
        L.orc_lp_features.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, f64p]
        L.orc_lp_loss_grad.argtypes = [P(Coo), ctypes.c_int, ctypes.c_int, f64p, i32p, f64p,
                                       ctypes.c_int, f64p, f64p, f64p]
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class Matrix:
    """A host COO held in numpy arrays (the reference's Mtx host COO)."""

    def __init__(self, nrows, ncols, row, col, val):
        self.nrows = int(nrows)
        self.ncols = int(ncols)
        self.row = np.ascontiguousarray(row, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float32)

    @property
    def nnz(self):
        return int(self.row.shape[0])

    def copy(self):
        return Matrix(self.nrows, self.ncols, self.row.copy(), self.col.copy(), self.val.copy())

    def c(self):
        return Coo(self.nrows, self.ncols, self.nnz, _ptr(self.row, ctypes.c_int32),
                   _ptr(self.col, ctypes.c_int32), _ptr(self.val, ctypes.c_float))

    def sync(self, c):
        self.nrows, self.ncols = c.nrows, c.ncols


def _from_c(c):
    n = int(c.nnz)
    m = Matrix(c.nrows, c.ncols,
               np.ctypeslib.as_array(c.row, (max(n, 1),))[:n].copy(),
               np.ctypeslib.as_array(c.col, (max(n, 1),))[:n].copy(),
               np.ctypeslib.as_array(c.val, (max(n, 1),))[:n].copy())
    lib().orc_coo_free(ctypes.byref(c))
    return m


def _check(rc, what):
    if rc != OK:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def mtx_read(path, M):
    c = Coo()
    _check(lib().orc_mtx_read(str(path).encode(), M, ctypes.byref(c)), "mtx_read")
    return _from_c(c)


def mtx_write(path, m):
    _check(lib().orc_mtx_write(str(path).encode(), ctypes.byref(m.c())), "mtx_write")


def from_tuples(nrows, ncols, row, col, val, M, symmetric=False):
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float32)
    c = Coo()
    _check(lib().orc_coo_from_tuples(nrows, ncols, int(symmetric), row.shape[0],
                                     _ptr(row, ctypes.c_int32), _ptr(col, ctypes.c_int32),
                                     _ptr(val, ctypes.c_float), M, ctypes.byref(c)), "from_tuples")
    return _from_c(c)


def nseg(m, M):
    return (m.nrows + M - 1) // M


def locate(m, N, M):
    ns = nseg(m, M)
    cnt = np.zeros(ns, np.int32)
    zero = np.zeros(ns, np.int32)
    _check(lib().orc_locate(ctypes.byref(m.c()), N, M, _ptr(cnt, ctypes.c_int32),
                            _ptr(zero, ctypes.c_int32)), "locate")
    return cnt, zero


def swap_score(m, N, M, s1, s2):
    out = np.zeros(2 * M * M, np.int32)
    _check(lib().orc_swap_score(ctypes.byref(m.c()), N, M, s1, s2, _ptr(out, ctypes.c_int32)),
           "swap_score")
    return out


def pq_pop_order(score, M):
    score = np.ascontiguousarray(score, dtype=np.int32)
    out = np.zeros(M * M, np.int32)
    _check(lib().orc_pq_pop_order(_ptr(score, ctypes.c_int32), M, _ptr(out, ctypes.c_int32)),
           "pq_pop_order")
    return out


def nm_reorder(m, N, M, sched, maxiter=10, trace_cap=0):
    """Relabels m in place; returns (stats[5], trace records or None)."""
    st = np.zeros(5, np.int32)
    tr = (SwapRec * max(trace_cap, 1))()
    tl = ctypes.c_int64(0)
    c = m.c()
    _check(lib().orc_nm_reorder(ctypes.byref(c), N, M, sched, maxiter, _ptr(st, ctypes.c_int32),
                                tr, trace_cap, ctypes.byref(tl)), "nm_reorder")
    m.sync(c)
    trace = None
    if trace_cap:
        n = min(tl.value, trace_cap)
        trace = np.array([[getattr(tr[i], f) for f, _ in SwapRec._fields_] for i in range(n)],
                         dtype=np.int32).reshape(-1, 9)
    return st, trace


def venom_eval(m, M, V):
    a, b = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().orc_venom_eval(ctypes.byref(m.c()), M, V, ctypes.byref(a), ctypes.byref(b)),
           "venom_eval")
    return a.value, b.value


def venom_reorder(m, M, V, maxiter=10):
    st = np.zeros(5, np.int32)
    c = m.c()
    _check(lib().orc_venom_reorder(ctypes.byref(c), M, V, maxiter, _ptr(st, ctypes.c_int32)),
           "venom_reorder")
    m.sync(c)
    return st


def set_threads(n):
    """Host threads of the reorder path (1 = sequential; results identical)."""
    lib().orc_set_threads(int(n))


def hybrid(m, N, M, V, rounds, sched):
    """Runs the hybrid in memory; returns (stats tuple, final Matrix)."""
    st = HybridStats()
    work = m.copy()  # the oracle writes the result into these arrays
    c = work.c()
    _check(lib().orc_hybrid(ctypes.byref(c), N, M, V, rounds, sched, ctypes.byref(st)), "hybrid")
    n = int(c.nnz)
    out = Matrix(c.nrows, c.ncols, work.row[:n].copy(), work.col[:n].copy(), work.val[:n].copy())
    return st.as_tuple(), out


def find_best_sched(m, N, M, nsched=4):
    b = ctypes.c_int32()
    _check(lib().orc_find_best_sched(ctypes.byref(m.c()), N, M, nsched, ctypes.byref(b)),
           "find_best_sched")
    return b.value


def retained_nnz(m, V, N, M):
    out = ctypes.c_int64()
    _check(lib().orc_retained_nnz(ctypes.byref(m.c()), V, N, M, ctypes.byref(out)),
           "retained_nnz")
    return out.value


def spmm_vnm_f64(m, V, M, B, threads=1):
    """FP64 V:N:M product; threads > 1 uses OpenMP (same result)."""
    B = np.ascontiguousarray(B, dtype=np.float32)
    C = np.zeros((m.nrows, B.shape[1]), np.float64)
    _check(lib().orc_spmm_vnm_f64_mt(ctypes.byref(m.c()), V, M, _ptr(B, ctypes.c_float), B.shape[1],
                                     _ptr(C, ctypes.c_double), threads), "spmm_vnm_f64")
    return C


def quant_cpp_write(v):
    return lib().orc_quant_cpp_write(v)


def quant_py_roundtrip(v):
    return lib().orc_quant_py_roundtrip(v)


def nm_reorder_sample(m, N, M, sched, maxiter, round_cap):
    """First round_cap rounds of an N:M run (CPU timing sample)."""
    st = np.zeros(5, np.int32)
    _check(lib().orc_nm_reorder_sample(ctypes.byref(m.c()), N, M, sched, maxiter, round_cap,
                                       _ptr(st, ctypes.c_int32)), "nm_reorder_sample")
    return st


def reload(m, M):
    """The loader-side effects of the temp.mtx round trip (in place)."""
    c = m.c()
    _check(lib().orc_reload(ctypes.byref(c), M), "reload")
    n = int(c.nnz)
    m.nrows, m.ncols = c.nrows, c.ncols
    m.row, m.col, m.val = m.row[:n].copy(), m.col[:n].copy(), m.val[:n].copy()
    return m


def ub_events():
    return lib().orc_ub_events()


LP_NPARAMS = 22784
LP_F = 16
LP_HEAD = 64


def lp_features(m, N, M):
    """Proposer input features (nrows x 16, FP64) of the matrix as labelled."""
    X = np.zeros((m.nrows, LP_F), np.float64)
    _check(lib().orc_lp_features(ctypes.byref(m.c()), N, M, _ptr(X, ctypes.c_double)),
           "lp_features")
    return X


def lp_loss_grad(m, N, M, params, pairs, target, want_grad=True):
    """(loss, grad[22784] or None, head[nrows x 64]) of the proposer in FP64."""
    P = np.ascontiguousarray(params, np.float64)
    pr = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    tg = np.ascontiguousarray(target, np.float64)
    loss = ctypes.c_double(0)
    grad = np.zeros(LP_NPARAMS, np.float64) if want_grad else None
    head = np.zeros((m.nrows, LP_HEAD), np.float64)
    _check(lib().orc_lp_loss_grad(ctypes.byref(m.c()), N, M, _ptr(P, ctypes.c_double),
                                  _ptr(pr, ctypes.c_int32), _ptr(tg, ctypes.c_double), len(tg),
                                  ctypes.byref(loss),
                                  _ptr(grad, ctypes.c_double) if want_grad else None,
                                  _ptr(head, ctypes.c_double)), "lp_loss_grad")
    return loss.value, grad, head


def allpairs_reorder(m, N, sched, maxiter=10, clean=False):
    """Allpairs schedule 4..7 (M = 4), in place; returns (stats, overwrites)."""
    st = np.zeros(5, np.int32)
    ow = ctypes.c_int64(0)
    c = m.c()
    _check(lib().orc_allpairs_reorder(ctypes.byref(c), N, sched, maxiter, int(clean),
                                      _ptr(st, ctypes.c_int32), ctypes.byref(ow)), "allpairs_reorder")
    m.sync(c)
    return st, ow.value


def swap_score_all(m, N, s1, s2):
    out = np.zeros(104, np.int32)
    _check(lib().orc_swap_score_all(ctypes.byref(m.c()), N, s1, s2, _ptr(out, ctypes.c_int32)),
           "swap_score_all")
    return out
