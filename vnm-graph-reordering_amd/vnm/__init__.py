"""Python binding of libvnm (ctypes over the C ABI in include/vnm.h).

Mirrors the reference's operator surface for the reorder path:
  Mtx(path, M)                      -> mtx/Mtx.h:21 (loads, pads, uploads)
  Mtx.write_to_file(path)           -> mtx/Mtx.h:412-435
  ReorderUtil(mtx, maxiter).run_*() -> reorder/reorder.h:319-390, 963-2474
  hybrid_reordering(...)            -> main.cu:195-337
  find_best_2toX(...)               -> main.cu:72-114
  eval_venom / VNM_reordering       -> main.cu:116-193 (reordering.py)
Errors raise VnmError carrying the library's status code and message; a
missing or unloadable libvnm.so raises at import of the first call, there is
no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libvnm.so")
CLI_PATH = os.path.join(os.path.dirname(_HERE), "bin", "spmm")

OK, E_IO, E_FORMAT, E_ARG, E_HIP, E_OOM, E_UB_PATH, E_NODEV = 0, -1, -2, -3, -4, -5, -6, -7
DT_F32, DT_BF16, DT_F16, DT_FP8 = 0, 1, 2, 3

EXPORTS = [
    "vnm_last_error", "vnm_mtx_read", "vnm_mtx_write", "vnm_coo_from_tuples", "vnm_coo_free",
    "vnm_gen_er", "vnm_gen_rmat", "vnm_gen_chung_lu", "vnm_gen_normal",
    "vnm_ctx_create", "vnm_ctx_destroy", "vnm_ctx_stream", "vnm_ctx_launches",
    "vnm_ctx_timing", "vnm_ctx_timer_read",
    "vnm_graph_create", "vnm_graph_destroy", "vnm_graph_download", "vnm_graph_perm", "vnm_graph_shape",
    "vnm_locate", "vnm_swap_score", "vnm_pq_pop_order", "vnm_nm_reorder",
    "vnm_venom_eval", "vnm_venom_reorder", "vnm_hybrid", "vnm_find_best_sched",
    "vnm_retained_nnz", "vnm_pack", "vnm_packed_info_get", "vnm_packed_destroy", "vnm_spmm",
    "vnm_nm_trace", "vnm_lp_model_create", "vnm_lp_model_destroy", "vnm_lp_model_get",
    "vnm_lp_features", "vnm_lp_forward", "vnm_lp_loss_grad", "vnm_lp_train_step",
    "vnm_pair_score", "vnm_lp_reorder", "vnm_hybrid_learned", "vnm_graph_clone",
    "vnm_spmm_set_shape", "vnm_spmm_get_shape", "vnm_gcn_layer", "vnm_batch_create",
    "vnm_batch_run", "vnm_batch_info", "vnm_batch_destroy", "vnm_nm_reorder_allpairs",
    "vnm_nm_reorder_batch", "vnm_mtx_read_ex", "vnm_mtx_write_ex", "vnm_lp_reorder_batch",
    "vnm_mtx_read_cached", "vnm_spmm_noalloc", "vnm_sort_pairs", "vnm_pack_ex",
]

LP_NPARAMS = 22784
LP_LEARNED, LP_RANDOM, LP_EXHAUSTIVE = 0, 1, 2


class VnmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Coo(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int32), ("ncols", ctypes.c_int32), ("nnz", ctypes.c_int64),
                ("row", ctypes.POINTER(ctypes.c_int32)), ("col", ctypes.POINTER(ctypes.c_int32)),
                ("val", ctypes.POINTER(ctypes.c_float))]


class NMStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("init_size", "final_size", "init_cnt", "final_cnt", "rounds")]

    def as_list(self):
        return [self.init_size, self.final_size, self.init_cnt, self.final_cnt, self.rounds]


class VenomStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("init_inv", "init_inv_strict", "final_inv", "final_inv_strict", "rounds")]

    def as_list(self):
        return [getattr(self, n) for n, _ in self._fields_]


class HybridStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("nrows", "nnz", "density", "init_size", "final_size", "init_cnt",
                 "final_cnt", "improve_rate")] + \
               [(n, ctypes.c_int32) for n in
                ("init_mb", "final_mb", "init_vnm", "final_vnm", "nm_rounds", "vnm_rounds",
                 "total_rounds")] + [("milliseconds", ctypes.c_float)]

    def as_tuple(self):
        return tuple(getattr(self, n) for n, _ in self._fields_)

    def csv(self):
        # main.cu:331-335, std::cout default float formatting
        def f(x):
            return "%g" % x
        t = self
        return ", ".join([f(t.nrows), f(t.nnz), f(t.density), f(t.init_size), f(t.final_size),
                          f(t.init_cnt), f(t.final_cnt), f(t.improve_rate), str(t.init_mb),
                          str(t.final_mb), str(t.init_vnm), str(t.final_vnm), str(t.nm_rounds),
                          str(t.vnm_rounds), str(t.total_rounds), f(t.milliseconds)])


class LPStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("init_cnt", "final_cnt", "init_size", "final_size", "iters", "swaps",
                 "fallbacks", "gain")] + [("exact_pairs", ctypes.c_int64), ("ms", ctypes.c_float)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class PackedInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("V", "M", "nrows", "ncols", "nblocks", "dtype")] + \
               [("retained_nnz", ctypes.c_int64), ("bytes", ctypes.c_int64), ("spmm_bytes", ctypes.c_int64),
                ("items", ctypes.c_int32), ("slots", ctypes.c_int32), ("dense_blocks", ctypes.c_int32),
                ("dense_chunks", ctypes.c_int32)]


_lib = None


def lib():
    """Load libvnm.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VnmError(E_HIP, f"{LIB_PATH} is not built (run vnm-graph-reordering_amd/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    i32p, vp = P(ctypes.c_int32), ctypes.c_void_p
    L.vnm_last_error.restype = ctypes.c_char_p
    L.vnm_mtx_read.argtypes = [ctypes.c_char_p, ctypes.c_int, P(Coo)]
    L.vnm_mtx_write.argtypes = [ctypes.c_char_p, P(Coo)]
    L.vnm_mtx_read_ex.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P(Coo)]
    L.vnm_mtx_read_cached.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P(ctypes.c_int), P(Coo)]
    L.vnm_mtx_write_ex.argtypes = [ctypes.c_char_p, P(Coo), ctypes.c_int]
    L.vnm_coo_from_tuples.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int, ctypes.c_int64,
                                      i32p, i32p, P(ctypes.c_float), ctypes.c_int, P(Coo)]
    L.vnm_coo_free.argtypes = [P(Coo)]
    L.vnm_coo_free.restype = None
    L.vnm_gen_er.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float,
                             ctypes.c_float, ctypes.c_int, P(Coo)]
    L.vnm_gen_rmat.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_uint64, ctypes.c_int, P(Coo)]
    L.vnm_gen_chung_lu.argtypes = [ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_uint64, ctypes.c_int, P(Coo)]
    L.vnm_gen_normal.argtypes = [ctypes.c_uint64, ctypes.c_int64, P(ctypes.c_float)]
    L.vnm_gen_normal.restype = None
    L.vnm_ctx_create.argtypes = [ctypes.c_int, P(vp)]
    L.vnm_ctx_destroy.argtypes = [vp]
    L.vnm_ctx_destroy.restype = None
    L.vnm_ctx_stream.argtypes = [vp]
    L.vnm_ctx_stream.restype = vp
    L.vnm_ctx_launches.argtypes = [vp]
    L.vnm_ctx_launches.restype = ctypes.c_int64
    L.vnm_ctx_timing.argtypes = [vp, ctypes.c_int]
    L.vnm_ctx_timer_read.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_double), P(ctypes.c_int64)]
    L.vnm_graph_create.argtypes = [vp, P(Coo), ctypes.c_int, P(vp)]
    L.vnm_graph_destroy.argtypes = [vp]
    L.vnm_graph_destroy.restype = None
    L.vnm_graph_download.argtypes = [vp, P(Coo)]
    L.vnm_graph_perm.argtypes = [vp, i32p]
    L.vnm_graph_shape.argtypes = [vp, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int64)]
    L.vnm_locate.argtypes = [vp, ctypes.c_int, i32p, i32p]
    L.vnm_swap_score.argtypes = [vp, ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, i32p]
    L.vnm_pq_pop_order.argtypes = [vp, i32p, ctypes.c_int, ctypes.c_int, i32p]
    L.vnm_nm_reorder.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(NMStats)]
    L.vnm_nm_reorder_batch.argtypes = [P(vp), ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, P(NMStats)]
    L.vnm_venom_eval.argtypes = [vp, ctypes.c_int, i32p, i32p]
    L.vnm_venom_reorder.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(VenomStats)]
    L.vnm_hybrid.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             P(HybridStats)]
    L.vnm_find_best_sched.argtypes = [vp, ctypes.c_int, ctypes.c_int, i32p]
    L.vnm_retained_nnz.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_int64)]
    L.vnm_pack.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(vp)]
    L.vnm_packed_info_get.argtypes = [vp, P(PackedInfo)]
    L.vnm_packed_destroy.argtypes = [vp]
    L.vnm_packed_destroy.restype = None
    L.vnm_spmm.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int]
    L.vnm_spmm_noalloc.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int]
    L.vnm_pack_ex.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(vp)]
    L.vnm_sort_pairs.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int]
    f32p = P(ctypes.c_float)
    L.vnm_nm_trace.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p,
                               ctypes.c_int64, P(ctypes.c_int64), P(NMStats)]
    L.vnm_lp_model_create.argtypes = [vp, f32p, ctypes.c_uint64, P(vp)]
    L.vnm_lp_model_destroy.argtypes = [vp]
    L.vnm_lp_model_destroy.restype = None
    L.vnm_lp_model_get.argtypes = [vp, f32p]
    L.vnm_lp_features.argtypes = [vp, ctypes.c_int, f32p]
    L.vnm_lp_forward.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, f32p]
    L.vnm_lp_loss_grad.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, i32p, f32p, ctypes.c_int,
                                   f32p, f32p]
    L.vnm_lp_train_step.argtypes = [vp, vp, ctypes.c_int, i32p, f32p, ctypes.c_int, ctypes.c_float,
                                    f32p]
    L.vnm_pair_score.argtypes = [vp, ctypes.c_int, i32p, i32p, ctypes.c_int, i32p]
    L.vnm_lp_reorder.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_uint64, P(LPStats)]
    L.vnm_lp_reorder_batch.argtypes = [P(vp), ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_uint64, P(LPStats)]
    L.vnm_hybrid_learned.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, P(HybridStats)]
    L.vnm_graph_clone.argtypes = [vp, P(vp)]
    L.vnm_nm_reorder_allpairs.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          P(NMStats)]
    L.vnm_spmm_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.vnm_gcn_layer.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int]
    L.vnm_batch_create.argtypes = [vp, P(vp), ctypes.c_int, P(vp), i32p, ctypes.c_int, ctypes.c_int, P(vp),
                                   P(vp), i32p, P(vp), ctypes.c_int, P(vp)]
    L.vnm_batch_run.argtypes = [vp]
    L.vnm_batch_info.argtypes = [vp, i32p, i32p, i32p]
    L.vnm_batch_destroy.argtypes = [vp]
    L.vnm_batch_destroy.restype = None
    L.vnm_spmm_get_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, i32p]
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise VnmError(rc, lib().vnm_last_error().decode())


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


# ------------------------------------------------------------------ host COO

class HostCoo:
    """Host COO in the reference's order (numpy arrays)."""

    def __init__(self, nrows, ncols, row, col, val):
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.row = np.ascontiguousarray(row, np.int32)
        self.col = np.ascontiguousarray(col, np.int32)
        self.val = np.ascontiguousarray(val, np.float32)

    @property
    def nnz(self):
        return int(self.row.shape[0])

    def c(self):
        return Coo(self.nrows, self.ncols, self.nnz, _p(self.row, ctypes.c_int32),
                   _p(self.col, ctypes.c_int32), _p(self.val, ctypes.c_float))

    def copy(self):
        return HostCoo(self.nrows, self.ncols, self.row.copy(), self.col.copy(), self.val.copy())


def _take(c):
    n = int(c.nnz)
    out = HostCoo(c.nrows, c.ncols,
                  np.ctypeslib.as_array(c.row, (max(n, 1),))[:n].copy(),
                  np.ctypeslib.as_array(c.col, (max(n, 1),))[:n].copy(),
                  np.ctypeslib.as_array(c.val, (max(n, 1),))[:n].copy())
    lib().vnm_coo_free(ctypes.byref(c))
    return out


def mtx_read(path, M, threads=0):
    """Load a MatrixMarket file as the reference loader does; threads = host
    worker threads (0: the machine's, 1: sequential), same result for any."""
    c = Coo()
    _check(lib().vnm_mtx_read_ex(str(path).encode(), M, threads, ctypes.byref(c)))
    return _take(c)


def mtx_read_cached(path, M, threads=0):
    """mtx_read through the binary cache next to the file; returns (coo,
    whether the cache was used)."""
    c = Coo()
    used = ctypes.c_int()
    _check(lib().vnm_mtx_read_cached(str(path).encode(), M, threads, ctypes.byref(used), ctypes.byref(c)))
    return _take(c), bool(used.value)


def mtx_write(path, coo, threads=0):
    _check(lib().vnm_mtx_write_ex(str(path).encode(), ctypes.byref(coo.c()), threads))


def coo_from_tuples(nrows, ncols, row, col, val, M, symmetric=False):
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float32)
    c = Coo()
    _check(lib().vnm_coo_from_tuples(nrows, ncols, int(symmetric), row.shape[0],
                                     _p(row, ctypes.c_int32), _p(col, ctypes.c_int32),
                                     _p(val, ctypes.c_float), M, ctypes.byref(c)))
    return _take(c)


def gen_er(n, k, seed, lo=-1.0, hi=1.0, M=8):
    c = Coo()
    _check(lib().vnm_gen_er(n, k, seed, lo, hi, M, ctypes.byref(c)))
    return _take(c)


def gen_rmat(scale, undirected_edges, seed, a=0.57, b=0.19, c_=0.19, M=8):
    c = Coo()
    _check(lib().vnm_gen_rmat(scale, undirected_edges, a, b, c_, seed, M, ctypes.byref(c)))
    return _take(c)


def gen_chung_lu(n, gamma, mean_degree, seed, M=8):
    c = Coo()
    _check(lib().vnm_gen_chung_lu(n, gamma, mean_degree, seed, M, ctypes.byref(c)))
    return _take(c)


def gen_normal(seed, shape):
    out = np.empty(int(np.prod(shape)), np.float32)
    lib().vnm_gen_normal(seed, out.shape[0], _p(out, ctypes.c_float))
    return out.reshape(shape)


# ------------------------------------------------------------------ device

class Context:
    """A device context.  close() waits for the graphs, packed matrices and
    models created on it: the library handle is destroyed with the last."""

    def __init__(self, device=0):
        self.h = None
        # live objects created on the context, counted rather than held in a
        # weak set: the cyclic collector clears every weak reference of a
        # garbage cycle before it runs the finalizers, so a weak set can look
        # empty while objects of the cycle still have to destroy their handles
        self._live = 0
        self._closing = False
        h = ctypes.c_void_p()
        _check(lib().vnm_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def _adopt(self, obj):
        self._live += 1

    def _release(self, obj):
        self._live -= 1
        self._maybe_destroy()

    def _maybe_destroy(self):
        if self.h and self._closing and self._live == 0:
            lib().vnm_ctx_destroy(self.h)
            self.h = None

    @property
    def stream(self):
        return lib().vnm_ctx_stream(self.h)

    @property
    def launches(self):
        return lib().vnm_ctx_launches(self.h)

    def close(self):
        self._closing = True
        self._maybe_destroy()

    def __del__(self):
        # objects created on the context hold it, so none is alive here
        # unless one was never closed (then the handle is left to process exit)
        try:
            self._closing = True
            self._maybe_destroy()
        except Exception:
            pass

    def timing(self, enable=True):
        _check(lib().vnm_ctx_timing(self.h, int(enable)))

    def timer(self, kernel):
        """(total milliseconds, launches) of a timed kernel."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _check(lib().vnm_ctx_timer_read(self.h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def sort_pairs(self, keys_in, keys_out, n, end_bit=64, vals_in=None, vals_out=None):
        """Stable radix sort of device arrays (pointers) on the context stream."""
        _check(lib().vnm_sort_pairs(self.h, keys_in, keys_out, vals_in, vals_out, n, end_bit))

    def pq_pop_order(self, scores, M):
        scores = np.ascontiguousarray(scores, np.int32).reshape(-1, 2 * M * M)
        out = np.zeros((scores.shape[0], M * M), np.int32)
        _check(lib().vnm_pq_pop_order(self.h, _p(scores, ctypes.c_int32), scores.shape[0], M,
                                      _p(out, ctypes.c_int32)))
        return out


class Graph:
    """Device-resident matrix (the reference's Mtx<float> device side)."""

    def __init__(self, ctx, coo, M):
        self.ctx = ctx
        self.M = M
        h = ctypes.c_void_p()
        _check(lib().vnm_graph_create(ctx.h, ctypes.byref(coo.c()), M, ctypes.byref(h)))
        self.h = h
        ctx._adopt(self)
        self.nrows0 = coo.nrows  # ids at creation: the length of perm()

    def _shape(self):
        r, c, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        _check(lib().vnm_graph_shape(self.h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(n)))
        return r.value, c.value, n.value

    @property
    def nrows(self):
        """Current rows: a VENOM reload resets the shape to (n, n), n the
        padded column count (reordering.py:180)."""
        return self._shape()[0]

    @property
    def ncols(self):
        return self._shape()[1]

    @property
    def nnz(self):
        return self._shape()[2]

    def close(self):
        if self.h:
            lib().vnm_graph_destroy(self.h)
            self.h = None
            self.ctx._release(self)

    def clone(self):
        """A device copy (current COO and composed relabel)."""
        h = ctypes.c_void_p()
        _check(lib().vnm_graph_clone(self.h, ctypes.byref(h)))
        g = Graph.__new__(Graph)
        g.ctx, g.M, g.h, g.nrows0 = self.ctx, self.M, h, self.nrows0
        self.ctx._adopt(g)
        return g

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nseg(self):
        return (self.nrows + self.M - 1) // self.M

    def download(self):
        c = Coo()
        _check(lib().vnm_graph_download(self.h, ctypes.byref(c)))
        return _take(c)

    def perm(self):
        out = np.zeros(self.nrows0, np.int32)
        _check(lib().vnm_graph_perm(self.h, _p(out, ctypes.c_int32)))
        return out

    def locate(self, N):
        cnt = np.zeros(self.nseg, np.int32)
        zero = np.zeros(self.nseg, np.int32)
        _check(lib().vnm_locate(self.h, N, _p(cnt, ctypes.c_int32), _p(zero, ctypes.c_int32)))
        return cnt, zero

    def swap_score(self, N, seg1, seg2):
        seg2 = np.ascontiguousarray(np.atleast_1d(seg2), np.int32)
        out = np.zeros((seg2.shape[0], 2 * self.M * self.M), np.int32)
        _check(lib().vnm_swap_score(self.h, N, seg1, _p(seg2, ctypes.c_int32), seg2.shape[0],
                                    _p(out, ctypes.c_int32)))
        return out

    def nm_reorder(self, N, sched, maxiter=10):
        st = NMStats()
        _check(lib().vnm_nm_reorder(self.h, N, sched, maxiter, ctypes.byref(st)))
        return st.as_list()

    @staticmethod
    def nm_reorder_batch(graphs, N, sched, maxiter=10):
        """nm_reorder of several graphs of one context in one launch chain;
        sched is one schedule for all or one per graph.  Returns the list of
        per-graph stats."""
        n = len(graphs)
        if isinstance(sched, int):
            sched = [sched] * n
        hs = (ctypes.c_void_p * n)(*[g.h for g in graphs])
        sc = np.ascontiguousarray(sched, dtype=np.int32)
        st = (NMStats * n)()
        _check(lib().vnm_nm_reorder_batch(hs, n, N, _p(sc, ctypes.c_int32), maxiter, st))
        return [s.as_list() for s in st]

    def nm_reorder_allpairs(self, N, sched, maxiter=10, clean=False):
        st = NMStats()
        _check(lib().vnm_nm_reorder_allpairs(self.h, N, sched, maxiter, int(clean), ctypes.byref(st)))
        return st.as_list()

    def nm_trace(self, N, sched, maxiter=10, round_cap=0, cap=1 << 20):
        """nm_reorder with a swap trace (9 ints per swap: round, primary,
        target, v1, v2, lgain, rgain, primary_cnt, target_cnt), optionally
        stopped after round_cap rounds.  Returns (stats, trace)."""
        st = NMStats()
        rec = np.zeros((max(cap, 1), 9), np.int32)
        n = ctypes.c_int64()
        _check(lib().vnm_nm_trace(self.h, N, sched, maxiter, round_cap, _p(rec, ctypes.c_int32), cap,
                                  ctypes.byref(n), ctypes.byref(st)))
        return st.as_list(), rec[:min(n.value, cap)].copy()

    def pair_score(self, N, segp, segq):
        """Exact gain matrices of segment pairs from their entries (same
        layout as swap_score: one row of 2*M*M ints per pair)."""
        segp = np.ascontiguousarray(np.atleast_1d(segp), np.int32)
        segq = np.ascontiguousarray(np.atleast_1d(segq), np.int32)
        out = np.zeros((segp.shape[0], 2 * self.M * self.M), np.int32)
        _check(lib().vnm_pair_score(self.h, N, _p(segp, ctypes.c_int32), _p(segq, ctypes.c_int32),
                                    segp.shape[0], _p(out, ctypes.c_int32)))
        return out

    def lp_features(self, N):
        X = np.zeros((self.nrows, 16), np.float32)
        _check(lib().vnm_lp_features(self.h, N, _p(X, ctypes.c_float)))
        return X

    def lp_reorder(self, model, N, mode=LP_LEARNED, topk=4096, window=16, maxiter=20, seed=1):
        st = LPStats()
        _check(lib().vnm_lp_reorder(self.h, model.h if model is not None else None, N, mode, topk,
                                    window, maxiter, seed, ctypes.byref(st)))
        return st

    @staticmethod
    def lp_reorder_batch(graphs, model, N, mode=LP_LEARNED, topk=4096, window=16, maxiter=20, seed=1):
        """lp_reorder of several graphs of one context as one launch chain (a
        captured iteration of all graphs, replayed); each result is
        lp_reorder's on the graph alone.  Returns the stats."""
        n = len(graphs)
        hs = (ctypes.c_void_p * n)(*[g.h for g in graphs])
        st = (LPStats * n)()
        _check(lib().vnm_lp_reorder_batch(hs, n, model.h if model is not None else None, N, mode, topk,
                                          window, maxiter, seed, st))
        return list(st)

    def venom_eval(self, V):
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().vnm_venom_eval(self.h, V, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def venom_reorder(self, V, maxiter=10):
        st = VenomStats()
        _check(lib().vnm_venom_reorder(self.h, V, maxiter, ctypes.byref(st)))
        return st.as_list()

    def hybrid(self, N, V, rounds, sched):
        st = HybridStats()
        _check(lib().vnm_hybrid(self.h, N, V, rounds, sched, ctypes.byref(st)))
        return st

    def hybrid_learned(self, model, N, V, rounds, mode=LP_LEARNED, topk=4096, window=16, maxiter=20):
        st = HybridStats()
        _check(lib().vnm_hybrid_learned(self.h, model.h if model is not None else None, N, V, rounds,
                                        mode, topk, window, maxiter, ctypes.byref(st)))
        return st

    def find_best_sched(self, N, nsched=4):
        b = ctypes.c_int32()
        _check(lib().vnm_find_best_sched(self.h, N, nsched, ctypes.byref(b)))
        return b.value

    def retained_nnz(self, V, N=2):
        out = ctypes.c_int64()
        _check(lib().vnm_retained_nnz(self.h, V, N, ctypes.byref(out)))
        return out.value

    def pack(self, V, dtype, dense_min=0):
        """V:N:M packing; dense_min > 0 puts blocks keeping at least that many
        values on the 2:4 sparse matrix cores (BF16 only)."""
        return Packed(self, V, dtype, dense_min)


class LPModel:
    """The learned pair proposer (parameters on the device, Adam state)."""

    def __init__(self, ctx, params=None, seed=1):
        self.ctx = ctx
        h = ctypes.c_void_p()
        pp = None
        if params is not None:
            self._init = np.ascontiguousarray(params, np.float32)
            assert self._init.shape == (LP_NPARAMS,)
            pp = _p(self._init, ctypes.c_float)
        _check(lib().vnm_lp_model_create(ctx.h, pp, seed, ctypes.byref(h)))
        self.h = h
        ctx._adopt(self)

    @classmethod
    def load(cls, ctx, path):
        return cls(ctx, np.fromfile(path, dtype="<f4"))

    def save(self, path):
        self.params().astype("<f4").tofile(path)

    def params(self):
        out = np.zeros(LP_NPARAMS, np.float32)
        _check(lib().vnm_lp_model_get(self.h, _p(out, ctypes.c_float)))
        return out

    def forward(self, graph, N, bf16=True):
        Y = np.zeros((graph.nrows, 64), np.float32)
        _check(lib().vnm_lp_forward(self.h, graph.h, N, int(bf16), _p(Y, ctypes.c_float)))
        return Y

    def loss_grad(self, graph, N, pairs, target, bf16=True, want_grad=True):
        pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        target = np.ascontiguousarray(target, np.float32)
        loss = ctypes.c_float()
        grad = np.zeros(LP_NPARAMS, np.float32) if want_grad else None
        _check(lib().vnm_lp_loss_grad(self.h, graph.h, N, int(bf16), _p(pairs, ctypes.c_int32),
                                      _p(target, ctypes.c_float), target.shape[0], ctypes.byref(loss),
                                      _p(grad, ctypes.c_float) if want_grad else None))
        return loss.value, grad

    def train_step(self, graph, N, pairs, target, lr):
        pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        target = np.ascontiguousarray(target, np.float32)
        loss = ctypes.c_float()
        _check(lib().vnm_lp_train_step(self.h, graph.h, N, _p(pairs, ctypes.c_int32),
                                       _p(target, ctypes.c_float), target.shape[0], lr,
                                       ctypes.byref(loss)))
        return loss.value

    def close(self):
        if self.h:
            lib().vnm_lp_model_destroy(self.h)
            self.h = None
            self.ctx._release(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Packed:
    def __init__(self, graph, V, dtype, dense_min=0):
        h = ctypes.c_void_p()
        if dense_min:
            _check(lib().vnm_pack_ex(graph.h, V, dtype, dense_min, ctypes.byref(h)))
        else:
            _check(lib().vnm_pack(graph.h, V, dtype, ctypes.byref(h)))
        self.h = h
        self.graph = graph
        graph.ctx._adopt(self)

    def info(self):
        i = PackedInfo()
        _check(lib().vnm_packed_info_get(self.h, ctypes.byref(i)))
        return i

    def spmm(self, B_ptr, ldb, n, b_dtype, C_ptr, ldc, b_scales_ptr=None):
        _check(lib().vnm_spmm(self.h, B_ptr, ldb, n, b_dtype, b_scales_ptr, C_ptr, ldc))

    def spmm_noalloc(self, B_ptr, ldb, n, b_dtype, C_ptr, ldc, b_scales_ptr=None):
        """spmm without any device allocation (static slot array of the library)."""
        _check(lib().vnm_spmm_noalloc(self.h, B_ptr, ldb, n, b_dtype, b_scales_ptr, C_ptr, ldc))

    def set_shape(self, b_dtype, n, cols_per_lane):
        """Fix the SpMM launch shape for (b_dtype, n); 0 re-enables tuning."""
        _check(lib().vnm_spmm_set_shape(self.h, b_dtype, n, cols_per_lane))

    def get_shape(self, b_dtype, n):
        g = ctypes.c_int32()
        _check(lib().vnm_spmm_get_shape(self.h, b_dtype, n, ctypes.byref(g)))
        return g.value

    def gcn_layer(self, X_ptr, ldx, K, W_ptr, N, Y_ptr, ldy, bias_ptr=None):
        """Y = A_vnm (X W) + bias (X, W BF16 device arrays; Y FP32)."""
        _check(lib().vnm_gcn_layer(self.h, X_ptr, ldx, K, W_ptr, N, bias_ptr, Y_ptr, ldy))

    def close(self):
        if self.h:
            lib().vnm_packed_destroy(self.h)
            self.h = None
            self.graph.ctx._release(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched SpMM over packed matrices (one persistent launch, HIP graph).
    B, C, scales and bias are lists of device pointers (ints) per matrix."""

    def __init__(self, ctx, packed, B, ldb, n, b_dtype, C, ldc, b_scales=None, bias=None, cols_per_lane=0):
        k = len(packed)
        VP = ctypes.c_void_p * k
        self._keep = packed
        ps = VP(*[p.h.value if isinstance(p.h, ctypes.c_void_p) else p.h for p in packed])
        bb, cc = VP(*B), VP(*C)
        lb = np.ascontiguousarray(ldb, np.int32)
        lc = np.ascontiguousarray(ldc, np.int32)
        sc = VP(*b_scales) if b_scales is not None else None
        bi = VP(*[x or 0 for x in bias]) if bias is not None else None
        h = ctypes.c_void_p()
        _check(lib().vnm_batch_create(ctx.h, ps, k, bb, _p(lb, ctypes.c_int32), n, b_dtype, sc, cc,
                                      _p(lc, ctypes.c_int32), bi, cols_per_lane, ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        ctx._adopt(self)

    def run(self):
        _check(lib().vnm_batch_run(self.h))

    def info(self):
        v, u, z = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().vnm_batch_info(self.h, ctypes.byref(v), ctypes.byref(u), ctypes.byref(z)))
        return {"cols_per_lane": v.value, "units": u.value, "zero_rows": z.value}

    def close(self):
        if self.h:
            lib().vnm_batch_destroy(self.h)
            self.h = None
            self.ctx._release(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ reference-shaped API

class Mtx:
    """Mtx<float>(mtxfile, aligned) (mtx/Mtx.h:21): load, pad, upload."""

    def __init__(self, ctx, mtxfile=None, aligned=8, coo=None):
        self.host = coo if coo is not None else mtx_read(mtxfile, aligned)
        self.graph = Graph(ctx, self.host, aligned)

    @property
    def nrows(self):
        return self.host.nrows

    @property
    def nnz(self):
        return self.host.nnz

    def write_to_file(self, path):
        mtx_write(path, self.graph.download())


class ReorderUtil:
    """reorderUtil<float>(Mtx*, maxiter) (reorder/reorder.h:319-390)."""

    def __init__(self, mtx, maxiter=10, N=2):
        self.mtx, self.maxiter, self.N = mtx, maxiter, N

    def _run(self, sched):
        return self.mtx.graph.nm_reorder(self.N, sched, self.maxiter)

    def run_onepair(self, verbose=False, stepmsg=False):
        return self._run(0)

    def run_onepair_double(self, verbose=False, stepmsg=False):
        return self._run(1)

    def run_onepair_reverse(self, verbose=False, stepmsg=False):
        return self._run(2)

    def run_onepair_reverse_double(self, verbose=False, stepmsg=False):
        return self._run(3)

    def run(self, verbose=False, stepmsg=False):
        return self._run(8)
