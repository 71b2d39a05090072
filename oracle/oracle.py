"""ctypes front end of the CPU restatement (libmela_oracle_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product library never
calls it.

Matrices follow MATLAB semantics: an m x n numpy array (bool, uint8 or
float64) is handed over in column-major order, exactly as a mxArray's data.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

ORC_LOGICAL = 0
ORC_DOUBLE = 1


def _cpu_has_avx512f() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f" in line
    except OSError:
        pass
    return False


def lib_path() -> str:
    name = "libmela_oracle_avx512.so" if _cpu_has_avx512f() else "libmela_oracle_avx2.so"
    return os.path.join(_HERE, name)


def load():
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built (run make -C oracle)")
    L = C.CDLL(path)
    vp, sz, u64p, i32p = C.c_void_p, C.c_size_t, C.POINTER(C.c_uint64), C.POINTER(C.c_int)
    L.orc_parity_double.argtypes = [C.c_double]
    L.orc_parity_double.restype = C.c_int
    L.orc_pack.argtypes = [vp, C.c_int, sz, sz, vp, sz]
    L.orc_unpack.argtypes = [vp, sz, sz, sz, vp]
    for f in (L.orc_matmul_gf2, L.orc_matmul_m4ri):
        f.argtypes = [vp, C.c_int, sz, sz, vp, C.c_int, sz, vp]
        f.restype = None
    for f in (L.orc_rank_gf2, L.orc_rank_m4ri):
        f.argtypes = [vp, C.c_int, sz, sz]
        f.restype = C.c_double
    for f in (L.orc_null_gf2, L.orc_null_m4ri):
        f.argtypes = [vp, C.c_int, sz, sz, vp]
        f.restype = C.c_long
    for f in (L.orc_matmul_gf2_packed, L.orc_matmul_m4ri_packed):
        f.argtypes = [vp, sz, sz, sz, vp, sz, sz, vp, sz]
        f.restype = None
    for f in (L.orc_rank_gf2_packed, L.orc_rank_m4ri_packed):
        f.argtypes = [vp, sz, sz, sz]
        f.restype = C.c_double
    for f in (L.orc_rref_gf2_packed, L.orc_rref_m4ri_packed):
        f.argtypes = [vp, sz, sz, sz, vp]
        f.restype = C.c_long
    L.orc_null_backsub.argtypes = [vp, sz, sz, vp, C.c_long, vp]
    L.orc_null_backsub.restype = None
    L.orc_splitmix64.argtypes = [C.c_uint64]
    L.orc_splitmix64.restype = C.c_uint64
    L.orc_gen_bernoulli_packed.argtypes = [C.c_uint64, sz, sz, vp, sz]
    L.orc_gen_bernoulli_packed.restype = None
    L.orc_num_threads.restype = C.c_int
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _as_input(A):
    """(column-major buffer, kind) for a logical or double matrix."""
    A = np.asarray(A)
    if A.ndim != 2:
        raise ValueError("matrix must be 2-D")
    if A.dtype == np.float64:
        return np.asfortranarray(A), ORC_DOUBLE
    if A.dtype == np.bool_ or A.dtype == np.uint8:
        return np.asfortranarray(A.astype(np.uint8, copy=False)), ORC_LOGICAL
    raise TypeError("oracle accepts logical (bool/uint8) or float64 matrices")


def parity_double(x: float) -> int:
    return int(load().orc_parity_double(float(x)))


def words(n: int) -> int:
    return (n + 63) // 64


def pack(A, ld: int | None = None) -> np.ndarray:
    buf, kind = _as_input(A)
    m, n = buf.shape
    ld = words(n) if ld is None else ld
    out = np.zeros((max(m, 1), max(ld, 1)), dtype=np.uint64)
    load().orc_pack(_ptr(buf), kind, m, n, _ptr(out), ld)
    return out[:m, :ld] if ld else np.zeros((m, 0), np.uint64)


def unpack(P: np.ndarray, rows: int, cols: int) -> np.ndarray:
    P = np.ascontiguousarray(P, dtype=np.uint64)
    out = np.zeros((max(rows * cols, 1),), dtype=np.uint8)
    ld = P.shape[1] if P.ndim == 2 else 0
    load().orc_unpack(_ptr(P), ld, rows, cols, _ptr(out))
    return out[: rows * cols].reshape((rows, cols), order="F").astype(bool)


def matmul(A, B, method: str = "gf2") -> np.ndarray:
    a, ka = _as_input(A)
    b, kb = _as_input(B)
    m, k = a.shape
    k2, n = b.shape
    assert k == k2
    out = np.zeros((max(m * n, 1),), dtype=np.uint8)
    fn = load().orc_matmul_gf2 if method == "gf2" else load().orc_matmul_m4ri
    fn(_ptr(a), ka, m, k, _ptr(b), kb, n, _ptr(out))
    return out[: m * n].reshape((m, n), order="F").astype(bool)


def rank(A, method: str = "gf2") -> float:
    a, ka = _as_input(A)
    m, n = a.shape
    fn = load().orc_rank_gf2 if method == "gf2" else load().orc_rank_m4ri
    return float(fn(_ptr(a), ka, m, n))


def null(A, method: str = "gf2") -> np.ndarray:
    a, ka = _as_input(A)
    m, n = a.shape
    Z = np.zeros((max(n * n, 1),), dtype=np.uint8)
    fn = load().orc_null_gf2 if method == "gf2" else load().orc_null_m4ri
    f = int(fn(_ptr(a), ka, m, n, _ptr(Z)))
    return Z[: n * f].reshape((n, f), order="F").astype(bool)


# ---- packed (kernel-only) forms -------------------------------------------

def matmul_packed(Ap, m, k, Bp, n, method: str = "gf2") -> np.ndarray:
    Ap = np.ascontiguousarray(Ap, dtype=np.uint64)
    Bp = np.ascontiguousarray(Bp, dtype=np.uint64)
    ldc = words(n)
    Cp = np.zeros((max(m, 1), max(ldc, 1)), dtype=np.uint64)
    fn = load().orc_matmul_gf2_packed if method == "gf2" else load().orc_matmul_m4ri_packed
    fn(_ptr(Ap), Ap.shape[1], m, k, _ptr(Bp), Bp.shape[1], n, _ptr(Cp), Cp.shape[1])
    return Cp[:m, :ldc]


def rank_packed(P, m, n, method: str = "gf2") -> float:
    if method == "gf2":
        M = np.array(P, dtype=np.uint64, order="C", copy=True)
        return float(load().orc_rank_gf2_packed(_ptr(M), m, n, M.shape[1] if M.ndim == 2 else 0))
    ld = (words(n) + 7) // 8 * 8
    M = np.zeros((max(m, 1), max(ld, 8)), dtype=np.uint64)
    M[:m, : P.shape[1]] = P
    return float(load().orc_rank_m4ri_packed(_ptr(M), m, n, M.shape[1]))


def rref_packed(P, m, n, method: str = "gf2"):
    """(echelon matrix, pivot list) as the reference's RREF routines leave them."""
    M = np.array(P, dtype=np.uint64, order="C", copy=True)
    if M.size == 0:
        M = np.zeros((max(m, 1), max(words(n), 1)), dtype=np.uint64)
    piv = np.zeros((max(min(m, n), 1),), dtype=np.int32)
    fn = load().orc_rref_gf2_packed if method == "gf2" else load().orc_rref_m4ri_packed
    r = int(fn(_ptr(M), m, n, M.shape[1], _ptr(piv)))
    return M, piv[:r].copy()


def null_from_rref(M, n, piv) -> np.ndarray:
    M = np.ascontiguousarray(M, dtype=np.uint64)
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    r = len(piv)
    Z = np.zeros((max(n * (n - r), 1),), dtype=np.uint8)
    load().orc_null_backsub(_ptr(M), n, M.shape[1], _ptr(piv), r, _ptr(Z))
    return Z[: n * (n - r)].reshape((n, n - r), order="F").astype(bool)


def gen_bernoulli_packed(seed: int, rows: int, cols: int, ld: int | None = None) -> np.ndarray:
    ld = words(cols) if ld is None else ld
    out = np.zeros((max(rows, 1), max(ld, 1)), dtype=np.uint64)
    load().orc_gen_bernoulli_packed(seed, rows, cols, _ptr(out), out.shape[1])
    return out[:rows, :ld]


def num_threads() -> int:
    return int(load().orc_num_threads())
