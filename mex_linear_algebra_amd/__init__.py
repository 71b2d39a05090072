"""MI355X-native GF(2) linear algebra: drop-in for HCPhy/mex_linear_algebra.

The six functions below mirror the reference's MATLAB entry points (README.md:
38-71): same names, same argument meaning (logical or double matrices, double
values taken as ((int)x) & 1), same outputs (logical matrix, double rank,
logical n x (n - rank) null basis) and the same error messages.  They call the
HIP library libmela.so through its C ABI (include/mela.h); there is no CPU
fallback -- a missing library or GPU raises MelaError.

    C = mela_matmul_gf2(A, B)      C = mela_matmul_m4ri(A, B)
    r = mela_rank_gf2(A)           r = mela_rank_m4ri(A)
    Z = mela_null_gf2(H)           Z = mela_null_m4ri(H)
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

__all__ = [
    "MelaError", "Context", "BitMatrix", "lib_path", "load", "default_context",
    "mela_matmul_gf2", "mela_matmul_m4ri", "mela_rank_gf2", "mela_rank_m4ri",
    "mela_null_gf2", "mela_null_m4ri", "mela_rref_gf2",
    "ALGO_AUTO", "ALGO_ANDPOP", "ALGO_M4RM", "ALGO_MFMA_I8", "ALGO_MFMA_FP4", "ALGO_ROWXOR",
    "ELIM_GE", "ELIM_M4RI",
    "ELIM_AUTO", "ELIM_BLOCKED", "ELIM_PANEL",
    "Batch", "BATCH_NULL", "BATCH_GE", "BATCH_GRAPH", "MatmulBatch",
    "Server", "Request", "Result", "OP_RANK", "OP_NULL", "OP_MATMUL",
]

_HERE = os.path.dirname(os.path.abspath(__file__))

MELA_LOGICAL, MELA_DOUBLE, MELA_OTHER, MELA_SPARSE = 0, 1, 2, 0x100
ALGO_AUTO, ALGO_ANDPOP, ALGO_M4RM, ALGO_MFMA_I8, ALGO_MFMA_FP4, ALGO_ROWXOR = 0, 1, 2, 3, 4, 5
ELIM_GE, ELIM_M4RI, ELIM_AUTO, ELIM_BLOCKED, ELIM_PANEL = 0, 1, 2, 3, 4

STATUS = {0: "OK", 1: "SPARSE", 2: "DIM", 3: "TYPE", 4: "NOMEM", 5: "INVALID",
          6: "DEVICE", 7: "ADMISSION", 8: "CANCELLED", 9: "TIMEOUT", 10: "ARGS"}


class MelaError(RuntimeError):
    """Raised with the reference's message text (and MATLAB id where it has one)."""

    def __init__(self, message, identifier="", status=None):
        super().__init__(message)
        self.identifier = identifier
        self.status = status


class BitMatrix(C.Structure):
    """mela_bitmat: device pointer + rows, cols, ld (64-bit words per row)."""
    _fields_ = [("data", C.c_void_p), ("rows", C.c_size_t), ("cols", C.c_size_t),
                ("ld", C.c_size_t)]


ALLOC_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t)

OP_RANK, OP_NULL, OP_MATMUL = 0, 1, 2


class Request(C.Structure):
    """mela_request: op, variant, device inputs A (and B for matmul)."""
    _fields_ = [("op", C.c_int), ("variant", C.c_int), ("A", BitMatrix), ("B", BitMatrix)]


class Result(C.Structure):
    """mela_result: status, rank, device output, latency."""
    _fields_ = [("status", C.c_int), ("rank", C.c_size_t), ("out", BitMatrix),
                ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cancelled", C.c_uint64), ("failed", C.c_uint64), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("hbm_in_use", C.c_size_t),
                ("hbm_peak", C.c_size_t), ("jobs", C.c_uint64), ("batched_requests", C.c_uint64),
                ("loop_launches", C.c_uint64)]

_lib = None
_lib_lock = threading.Lock()


def lib_path() -> str:
    return os.environ.get("MELA_LIB", os.path.join(_HERE, "libmela.so"))


def load():
    """Loads libmela.so (fails loudly if it is not built)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not os.path.exists(path):
            raise MelaError(f"libmela.so not built at {path} (run __graft_entry__.build())")
        L = C.CDLL(path)
        st, vp, sz = C.c_int, C.c_void_p, C.c_size_t
        ctxp = C.c_void_p
        L.mela_ctx_create.argtypes = [C.c_int, C.POINTER(ctxp)]
        L.mela_ctx_create.restype = st
        L.mela_ctx_destroy.argtypes = [ctxp]
        L.mela_ctx_destroy.restype = None
        L.mela_ctx_stream.argtypes = [ctxp]
        L.mela_ctx_stream.restype = vp
        L.mela_ctx_sync.argtypes = [ctxp]
        L.mela_ctx_sync.restype = st
        L.mela_ctx_elim_counts.argtypes = [ctxp] + [C.POINTER(C.c_ulonglong)] * 3
        L.mela_ctx_elim_counts.restype = st
        L.mela_ctx_profile_begin.argtypes = [ctxp]
        L.mela_ctx_profile_begin.restype = st
        L.mela_ctx_profile_end.argtypes = [ctxp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.mela_ctx_profile_end.restype = st
        L.mela_last_error.restype = C.c_char_p
        L.mela_last_error_id.restype = C.c_char_p
        L.mela_version.restype = C.c_char_p
        for f in (L.mela_matmul_gf2, L.mela_matmul_m4ri):
            f.argtypes = [ctxp, vp, C.c_int, sz, sz, vp, C.c_int, sz, sz, vp]
            f.restype = st
        for f in (L.mela_rank_gf2, L.mela_rank_m4ri):
            f.argtypes = [ctxp, vp, C.c_int, sz, sz, C.POINTER(C.c_double)]
            f.restype = st
        for f in (L.mela_null_gf2, L.mela_null_m4ri):
            f.argtypes = [ctxp, vp, C.c_int, sz, sz, ALLOC_FN, vp, C.POINTER(sz)]
            f.restype = st
        L.mela_rref_gf2.argtypes = [ctxp, vp, C.c_int, sz, sz, vp, C.POINTER(C.c_double),
                                    C.POINTER(sz)]
        L.mela_rref_gf2.restype = st
        bm = C.POINTER(BitMatrix)
        L.mela_pack_csc.argtypes = [ctxp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), vp,
                                    C.c_int, sz, sz, bm]
        L.mela_pack_csc.restype = st
        L.mela_dev_alloc.argtypes = [ctxp, sz, C.POINTER(vp)]
        L.mela_dev_alloc.restype = st
        L.mela_dev_free.argtypes = [ctxp, vp]
        L.mela_dev_free.restype = st
        L.mela_copy_h2d.argtypes = [ctxp, vp, vp, sz]
        L.mela_copy_h2d.restype = st
        L.mela_copy_d2h.argtypes = [ctxp, vp, vp, sz]
        L.mela_copy_d2h.restype = st
        L.mela_copy_d2d.argtypes = [ctxp, vp, vp, sz]
        L.mela_copy_d2d.restype = st
        L.mela_bitmat_alloc.argtypes = [ctxp, sz, sz, bm]
        L.mela_bitmat_alloc.restype = st
        L.mela_bitmat_free.argtypes = [ctxp, bm]
        L.mela_bitmat_free.restype = None
        L.mela_gen_bernoulli.argtypes = [ctxp, C.c_uint64, bm]
        L.mela_gen_bernoulli.restype = st
        L.mela_pack_dev.argtypes = [ctxp, vp, C.c_int, bm]
        L.mela_pack_dev.restype = st
        L.mela_unpack_dev.argtypes = [ctxp, bm, vp]
        L.mela_unpack_dev.restype = st
        L.mela_matmul_packed.argtypes = [ctxp, bm, bm, bm, C.c_int]
        L.mela_matmul_packed.restype = st
        L.mela_rank_packed.argtypes = [ctxp, bm, C.c_int, C.POINTER(sz)]
        L.mela_rank_packed.restype = st
        L.mela_rank_packed_noalloc.argtypes = [ctxp, bm, C.c_int, C.POINTER(sz)]
        L.mela_rank_packed_noalloc.restype = st
        L.mela_rref_packed.argtypes = [ctxp, bm, C.c_int, bm, C.POINTER(C.c_int), C.POINTER(sz)]
        L.mela_rref_packed.restype = st
        L.mela_null_packed.argtypes = [ctxp, bm, C.c_int, bm, C.POINTER(sz)]
        L.mela_null_packed.restype = st
        bp = C.c_void_p
        L.mela_batch_create.argtypes = [ctxp, bm, sz, C.c_int, C.POINTER(bp)]
        L.mela_batch_create.restype = st
        L.mela_batch_set_modes.argtypes = [bp, C.POINTER(C.c_int)]
        L.mela_batch_set_modes.restype = st
        L.mela_batch_set_tile.argtypes = [bp, C.c_int]
        L.mela_batch_set_tile.restype = st
        L.mela_batch_get_plan.argtypes = [bp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.mela_batch_get_plan.restype = st
        L.mela_batch_set_groups.argtypes = [bp, C.POINTER(C.c_int)]
        L.mela_batch_set_groups.restype = st
        L.mela_batch_autotune.argtypes = [bp, C.POINTER(C.c_float)]
        L.mela_batch_autotune.restype = st
        L.mela_batch_run.argtypes = [bp]
        L.mela_batch_run.restype = st
        L.mela_batch_results.argtypes = [bp, C.POINTER(C.c_uint32), bm]
        L.mela_batch_results.restype = st
        L.mela_batch_destroy.argtypes = [bp]
        L.mela_batch_destroy.restype = None
        L.mela_matmul_batch_create.argtypes = [ctxp, bm, bm, bm, sz, C.c_int, C.POINTER(bp)]
        L.mela_matmul_batch_create.restype = st
        L.mela_matmul_batch_set_algos.argtypes = [bp, C.POINTER(C.c_int)]
        L.mela_matmul_batch_set_algos.restype = st
        L.mela_matmul_batch_get_algos.argtypes = [bp, C.POINTER(C.c_int)]
        L.mela_matmul_batch_get_algos.restype = st
        L.mela_matmul_batch_run.argtypes = [bp]
        L.mela_matmul_batch_run.restype = st
        L.mela_matmul_batch_destroy.argtypes = [bp]
        L.mela_matmul_batch_destroy.restype = None
        srv = C.c_void_p
        L.mela_server_create.argtypes = [C.c_int, sz, C.c_int, C.c_int, C.POINTER(srv)]
        L.mela_server_create.restype = st
        L.mela_server_submit.argtypes = [srv, C.POINTER(Request), C.POINTER(C.c_uint64)]
        L.mela_server_submit.restype = st
        L.mela_server_wait.argtypes = [srv, C.c_uint64, C.c_double, C.POINTER(Result)]
        L.mela_server_wait.restype = st
        L.mela_server_cancel.argtypes = [srv, C.c_uint64]
        L.mela_server_cancel.restype = st
        L.mela_server_release.argtypes = [srv, C.c_uint64]
        L.mela_server_release.restype = st
        L.mela_server_stats_get.argtypes = [srv, C.POINTER(ServerStats)]
        L.mela_server_stats_get.restype = st
        L.mela_server_reset_stats.argtypes = [srv]
        L.mela_server_reset_stats.restype = st
        L.mela_server_destroy.argtypes = [srv]
        L.mela_server_destroy.restype = None
        _lib = L
        return L


def _raise(status):
    L = load()
    raise MelaError(L.mela_last_error().decode(), L.mela_last_error_id().decode(),
                    STATUS.get(status, status))


def _check(status):
    if status != 0:
        _raise(status)


class Context:
    """A device context (HIP stream + cached device scratch)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.mela_ctx_create(device, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.mela_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return self.lib.mela_ctx_stream(self.handle)

    def sync(self):
        _check(self.lib.mela_ctx_sync(self.handle))

    def elim_counts(self):
        """Eliminations run by this context: (blocked, fallback, panel)."""
        v = [C.c_ulonglong() for _ in range(3)]
        _check(self.lib.mela_ctx_elim_counts(self.handle, *[C.byref(x) for x in v]))
        return tuple(int(x.value) for x in v)

    def profile_begin(self):
        """Start timing the dominant kernel of every matmul_packed call."""
        _check(self.lib.mela_ctx_profile_begin(self.handle))

    def profile_end(self):
        """Stop timing; returns (summed kernel ms, number of timed launches)."""
        ms, cnt = C.c_double(), C.c_int()
        _check(self.lib.mela_ctx_profile_end(self.handle, C.byref(ms), C.byref(cnt)))
        return float(ms.value), int(cnt.value)

    # ---- device-resident packed helpers ----
    def bitmat(self, rows: int, cols: int) -> BitMatrix:
        M = BitMatrix()
        _check(self.lib.mela_bitmat_alloc(self.handle, rows, cols, C.byref(M)))
        return M

    def free(self, M: BitMatrix):
        self.lib.mela_bitmat_free(self.handle, C.byref(M))

    def upload(self, P: np.ndarray, cols: int) -> BitMatrix:
        P = np.ascontiguousarray(P, dtype=np.uint64)
        M = self.bitmat(P.shape[0], cols)
        if P.shape[1] != M.ld:
            Q = np.zeros((P.shape[0], M.ld), dtype=np.uint64)
            Q[:, : min(P.shape[1], M.ld)] = P[:, : M.ld]
            P = Q
        if P.size:
            _check(self.lib.mela_copy_h2d(self.handle, M.data, P.ctypes.data, P.nbytes))
        return M

    def download(self, M: BitMatrix) -> np.ndarray:
        out = np.zeros((M.rows, M.ld), dtype=np.uint64)
        if out.size:
            _check(self.lib.mela_copy_d2h(self.handle, out.ctypes.data, M.data, out.nbytes))
        return out

    def pack_dev(self, A) -> BitMatrix:
        """Packed device matrix from a logical or double array that is first
        copied to device memory as MATLAB's column-major bytes
        (mela_pack_dev: the device-side parity packer)."""
        buf, cls, m, n = _classify(A)
        if buf is None:
            raise MelaError("Inputs must be logical or double matrices.")
        M = self.bitmat(m, n)
        src = C.c_void_p()
        _check(self.lib.mela_dev_alloc(self.handle, max(buf.nbytes, 1), C.byref(src)))
        try:
            if buf.nbytes:
                _check(self.lib.mela_copy_h2d(self.handle, src, buf.ctypes.data, buf.nbytes))
            _check(self.lib.mela_pack_dev(self.handle, src, cls, C.byref(M)))
            self.sync()
        finally:
            self.lib.mela_dev_free(self.handle, src)
        return M

    def unpack_dev(self, M: BitMatrix) -> np.ndarray:
        """M as a column-major logical array unpacked on the device
        (mela_unpack_dev), copied back to the host."""
        out = np.zeros((M.rows, M.cols), dtype=np.uint8, order="F")
        dst = C.c_void_p()
        _check(self.lib.mela_dev_alloc(self.handle, max(out.nbytes, 1), C.byref(dst)))
        try:
            _check(self.lib.mela_unpack_dev(self.handle, C.byref(M), dst))
            if out.nbytes:
                _check(self.lib.mela_copy_d2h(self.handle, out.ctypes.data, dst, out.nbytes))
        finally:
            self.lib.mela_dev_free(self.handle, dst)
        return out.view(np.bool_)

    def gen_bernoulli(self, seed: int, M: BitMatrix):
        _check(self.lib.mela_gen_bernoulli(self.handle, seed, C.byref(M)))

    def matmul_packed(self, A: BitMatrix, B: BitMatrix, Cm: BitMatrix, algo: int = ALGO_AUTO):
        _check(self.lib.mela_matmul_packed(self.handle, C.byref(A), C.byref(B), C.byref(Cm), algo))

    def rank_packed(self, A: BitMatrix, variant: int = ELIM_AUTO) -> int:
        r = C.c_size_t()
        _check(self.lib.mela_rank_packed(self.handle, C.byref(A), variant, C.byref(r)))
        return r.value

    def rank_packed_noalloc(self, A: BitMatrix, variant: int = ELIM_AUTO) -> int:
        """rank through the allocation-free entry point (fixed 32 MiB device
        work array, no device allocation in the call)."""
        r = C.c_size_t()
        _check(self.lib.mela_rank_packed_noalloc(self.handle, C.byref(A), variant, C.byref(r)))
        return r.value

    def rref_packed(self, A: BitMatrix, variant: int = ELIM_AUTO):
        rmax = min(A.rows, A.cols)
        R = self.bitmat(rmax, A.cols)
        piv = (C.c_int * max(rmax, 1))()
        r = C.c_size_t()
        _check(self.lib.mela_rref_packed(self.handle, C.byref(A), variant, C.byref(R), piv,
                                         C.byref(r)))
        return R, [piv[i] for i in range(r.value)]

    def pack_csc(self, indptr, indices, data, m: int, n: int) -> BitMatrix:
        """Packed m x n device matrix from compressed sparse column arrays
        (MATLAB's jc / ir / pr, or scipy's indptr / indices / data).  data:
        None (every stored entry is 1), bool (nonzero) or float64 (the dense
        parity rule)."""
        jc = np.ascontiguousarray(indptr, dtype=np.uint64)
        ir = np.ascontiguousarray(indices, dtype=np.uint64)
        if jc.size != n + 1:
            raise MelaError("indptr must have n + 1 entries")
        vals, vcls = None, MELA_LOGICAL
        if data is not None:
            d = np.asarray(data)
            if d.dtype == np.bool_:
                vals, vcls = np.ascontiguousarray(d).view(np.uint8), MELA_LOGICAL
            elif d.dtype == np.float64:
                vals, vcls = np.ascontiguousarray(d), MELA_DOUBLE
            else:
                raise MelaError("Inputs must be logical or double matrices.")
        A = self.bitmat(m, n)
        try:
            _check(self.lib.mela_pack_csc(
                self.handle, jc.ctypes.data_as(C.POINTER(C.c_uint64)),
                ir.ctypes.data_as(C.POINTER(C.c_uint64)) if ir.size else None,
                None if vals is None or vals.size == 0 else vals.ctypes.data, vcls, m, n,
                C.byref(A)))
        except MelaError:
            self.free(A)
            raise
        return A

    def upload_sparse(self, S) -> BitMatrix:
        """Packed device matrix from a scipy.sparse matrix (any format)."""
        S = S.tocsc()
        S.sort_indices()
        data = S.data if S.dtype in (np.float64, np.bool_) else S.data.astype(np.float64)
        return self.pack_csc(S.indptr, S.indices, data, S.shape[0], S.shape[1])

    def null_packed(self, A: BitMatrix, variant: int = ELIM_AUTO) -> BitMatrix:
        Z = BitMatrix()
        f = C.c_size_t()
        _check(self.lib.mela_null_packed(self.handle, C.byref(A), variant, C.byref(Z), C.byref(f)))
        return Z


BATCH_NULL, BATCH_GE, BATCH_GRAPH = 1, 2, 4


class Batch:
    """Batched rank + canonical null basis of many packed device matrices in
    one persistent launch (mela_batch_* in include/mela.h)."""

    def __init__(self, ctx: Context, mats, flags: int = BATCH_NULL | BATCH_GRAPH):
        self.ctx = ctx
        self.lib = ctx.lib
        self.count = len(mats)
        self._mats = (BitMatrix * max(self.count, 1))(*mats)
        h = C.c_void_p()
        _check(self.lib.mela_batch_create(ctx.handle, self._mats, self.count, flags, C.byref(h)))
        self.handle = h

    def set_modes(self, modes):
        """Per matrix: 0 = GE, 1 = M4RI with k = 4, 2 = M4RI with k = 8
        (8-word tiles; k = 4 with wider tiles)."""
        arr = (C.c_int * max(self.count, 1))(*[int(x) for x in modes])
        _check(self.lib.mela_batch_set_modes(self.handle, arr))

    def set_tile(self, words: int):
        """Update tile width: 8, 16 or 32 words."""
        _check(self.lib.mela_batch_set_tile(self.handle, int(words)))

    def plan(self):
        """(tile words, per-matrix modes, per-matrix workgroups) of the next run."""
        tw = C.c_int()
        modes = (C.c_int * max(self.count, 1))()
        groups = (C.c_int * max(self.count, 1))()
        _check(self.lib.mela_batch_get_plan(self.handle, C.byref(tw), modes, groups))
        return tw.value, [modes[i] for i in range(self.count)], [groups[i] for i in range(self.count)]

    def set_groups(self, groups):
        """Workgroups (1..8) that eliminate each matrix together."""
        arr = (C.c_int * max(self.count, 1))(*[int(x) for x in groups])
        _check(self.lib.mela_batch_set_groups(self.handle, arr))

    def autotune(self) -> float:
        ms = C.c_float()
        _check(self.lib.mela_batch_autotune(self.handle, C.byref(ms)))
        return ms.value

    def run(self):
        _check(self.lib.mela_batch_run(self.handle))

    def results(self):
        ranks = (C.c_uint32 * max(self.count, 1))()
        nulls = (BitMatrix * max(self.count, 1))()
        _check(self.lib.mela_batch_results(self.handle, ranks, nulls))
        return [ranks[i] for i in range(self.count)], [nulls[i] for i in range(self.count)]

    def close(self):
        # destroying a batch needs its context (stream and lock); after the
        # context is closed the batch's device memory stays with the process
        if self.handle and self.ctx.handle:
            self.lib.mela_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MatmulBatch:
    """C_i = A_i * B_i for many uneven packed device products in one
    persistent launch (mela_matmul_batch_* in include/mela.h).  The C_i are
    caller-allocated device matrices."""

    def __init__(self, ctx: Context, As, Bs, Cs, flags: int = BATCH_GRAPH):
        if not (len(As) == len(Bs) == len(Cs)):
            raise MelaError("A, B and C lists must have the same length")
        self.ctx = ctx
        self.lib = ctx.lib
        self.count = len(As)
        cnt = max(self.count, 1)
        self._A = (BitMatrix * cnt)(*As)
        self._B = (BitMatrix * cnt)(*Bs)
        self._C = (BitMatrix * cnt)(*Cs)
        h = C.c_void_p()
        _check(self.lib.mela_matmul_batch_create(ctx.handle, self._A, self._B, self._C, self.count,
                                                 flags, C.byref(h)))
        self.handle = h

    def algos(self):
        """Engine of each item: ALGO_MFMA_FP4, ALGO_ANDPOP or ALGO_ROWXOR."""
        arr = (C.c_int * max(self.count, 1))()
        _check(self.lib.mela_matmul_batch_get_algos(self.handle, arr))
        return [arr[i] for i in range(self.count)]

    def set_algos(self, algos):
        """Force each item's engine (ALGO_MFMA_FP4 needs k < 2^23)."""
        arr = (C.c_int * max(self.count, 1))(*[int(a) for a in algos])
        _check(self.lib.mela_matmul_batch_set_algos(self.handle, arr))

    def run(self):
        _check(self.lib.mela_matmul_batch_run(self.handle))

    def close(self):
        if self.handle and self.ctx.handle:
            self.lib.mela_matmul_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request server (mela_server_* in include/mela.h)."""

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, workers: int = 4,
                 max_batch: int = 64):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.mela_server_create(device, hbm_budget, workers, max_batch, C.byref(h)))
        self.handle = h

    def submit(self, op: int, A: BitMatrix, B: BitMatrix | None = None, variant: int | None = None) -> int:
        """variant: ELIM_* for rank / null, ALGO_* for matmul; AUTO of the
        request's kind when None."""
        if variant is None:
            variant = ALGO_AUTO if op == OP_MATMUL else ELIM_AUTO
        req = Request(op, variant, A, B if B is not None else BitMatrix())
        t = C.c_uint64()
        _check(self.lib.mela_server_submit(self.handle, C.byref(req), C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_s: float = -1.0) -> Result:
        res = Result()
        _check(self.lib.mela_server_wait(self.handle, ticket, timeout_s, C.byref(res)))
        return res

    def cancel(self, ticket: int):
        _check(self.lib.mela_server_cancel(self.handle, ticket))

    def release(self, ticket: int):
        _check(self.lib.mela_server_release(self.handle, ticket))

    def stats(self) -> dict:
        st = ServerStats()
        _check(self.lib.mela_server_stats_get(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(self.lib.mela_server_reset_stats(self.handle))

    def close(self):
        if self.handle:
            self.lib.mela_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None
_default_lock = threading.Lock()


def default_context() -> Context:
    global _default
    with _default_lock:
        if _default is None:
            _default = Context(0)
        return _default


def _classify(A):
    """(column-major buffer, class code, rows, cols) for a MATLAB-style input."""
    if hasattr(A, "tocsc") and hasattr(A, "nnz"):  # scipy.sparse
        shape = A.shape
        code = MELA_DOUBLE if A.dtype == np.float64 else (MELA_LOGICAL if A.dtype == np.bool_ else MELA_OTHER)
        return None, code | MELA_SPARSE, shape[0], shape[1]
    A = np.asarray(A)
    if A.ndim > 2:
        raise MelaError("matrices must be 2-D")
    if A.ndim < 2:
        A = A.reshape((1, -1)) if A.ndim == 1 else A.reshape((1, 1))
    if A.dtype == np.bool_:
        buf = np.asfortranarray(A).view(np.uint8)
        return buf, MELA_LOGICAL, A.shape[0], A.shape[1]
    if A.dtype == np.float64:
        return np.asfortranarray(A), MELA_DOUBLE, A.shape[0], A.shape[1]
    return None, MELA_OTHER, A.shape[0], A.shape[1]


def _ptr(buf):
    return None if buf is None or buf.size == 0 else buf.ctypes.data


def _matmul(fn_name, A, B, ctx):
    ctx = ctx or default_context()
    a, ac, m, k = _classify(A)
    b, bc, k2, n = _classify(B)
    out = np.zeros((m, n), dtype=np.uint8, order="F")
    fn = getattr(ctx.lib, fn_name)
    _check(fn(ctx.handle, _ptr(a), ac, m, k, _ptr(b), bc, k2, n, _ptr(out)))
    return out.view(np.bool_)


def mela_matmul_gf2(A, B, *more, ctx: Context | None = None):
    """C = A*B over GF(2) (mela_matmul_gf2.c:72): logical m x n."""
    if more:
        raise MelaError("Need two input matrices A,B.")
    return _matmul("mela_matmul_gf2", A, B, ctx)


def mela_matmul_m4ri(A, B, *more, ctx: Context | None = None):
    """C = A*B over GF(2) by the Method of Four Russians (mela_matmul_m4ri.c:140)."""
    if more:
        raise MelaError("Need two input matrices A,B.")
    return _matmul("mela_matmul_m4ri", A, B, ctx)


def _rank(fn_name, A, ctx):
    ctx = ctx or default_context()
    a, ac, m, n = _classify(A)
    r = C.c_double()
    _check(getattr(ctx.lib, fn_name)(ctx.handle, _ptr(a), ac, m, n, C.byref(r)))
    return float(r.value)


def mela_rank_gf2(A, *more, ctx: Context | None = None):
    """rank over GF(2) (mela_rank_gf2.c:126), returned as a double."""
    if more:
        raise MelaError("One input required.", "bitrank_mex:nargin")
    return _rank("mela_rank_gf2", A, ctx)


def mela_rank_m4ri(A, *more, ctx: Context | None = None):
    """rank over GF(2) by block M4RI (mela_rank_m4ri.c:286), returned as a double."""
    if more:
        raise MelaError("One input required.")
    return _rank("mela_rank_m4ri", A, ctx)


def _null(fn_name, H, ctx):
    ctx = ctx or default_context()
    h, hc, m, n = _classify(H)
    holder = {}

    def _alloc(_user, rows, cols):
        arr = np.zeros((rows, cols), dtype=np.uint8, order="F")
        holder["Z"] = arr
        return arr.ctypes.data if arr.size else None

    cb = ALLOC_FN(_alloc)
    f = C.c_size_t()
    _check(getattr(ctx.lib, fn_name)(ctx.handle, _ptr(h), hc, m, n, cb, None, C.byref(f)))
    Z = holder.get("Z", np.zeros((n, f.value), dtype=np.uint8, order="F"))
    return Z.view(np.bool_)


def mela_rref_gf2(A, ctx: Context | None = None):
    """[R, jb] = gf2rref(A) (gf2null/gf2rref.m:49-73): R is the m x n logical
    reduced row echelon form, jb the 1-based pivot columns (a double row
    vector, empty when the rank is 0)."""
    ctx = ctx or default_context()
    a, ac, m, n = _classify(A)
    R = np.zeros((m, n), dtype=np.uint8, order="F")
    jb = np.zeros(max(min(m, n), 1), dtype=np.float64)
    r = C.c_size_t()
    _check(ctx.lib.mela_rref_gf2(ctx.handle, _ptr(a), ac, m, n, _ptr(R),
                                 jb.ctypes.data_as(C.POINTER(C.c_double)), C.byref(r)))
    return R.view(np.bool_), jb[: r.value].reshape((1, -1))


def mela_null_gf2(H, *more, ctx: Context | None = None):
    """canonical null-space basis over GF(2) (mela_null_gf2.c:142): logical n x (n-rank)."""
    if more:
        raise MelaError("One input required: H.")
    return _null("mela_null_gf2", H, ctx)


def mela_null_m4ri(H, *more, ctx: Context | None = None):
    """canonical null-space basis over GF(2) via M4RI (mela_null_m4ri.c:251)."""
    if more:
        raise MelaError("One input required: H.")
    return _null("mela_null_m4ri", H, ctx)
