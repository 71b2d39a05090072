"""Host-side column packing of the entry points' MATLAB arrays (csrc/hostpack.cpp),
built with g++ and checked against the oracle's packer on the transposed
matrix: column j of A packed = row j of pack(A^T).  Ragged sizes, logical bytes
other than 0 / 1, the doubles whose parity the x86-64 cast fixes (NaN, +-Inf,
the int range edges, negative and fractional values), sizes large enough to
run on several pool threads, and concurrent callers."""
import ctypes as C
import os
import subprocess
import threading

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mex_linear_algebra_amd", "csrc")
MELA_LOGICAL, MELA_DOUBLE = 0, 1  # include/mela.h


@pytest.fixture(scope="module")
def hp(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hostpack") / "hp.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-pthread",
                    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                    os.path.join(ROOT, "tests", "hostpack_capi.cpp"),
                    os.path.join(CSRC, "hostpack.cpp"), "-o", out], check=True)
    lib = C.CDLL(out)
    lib.hp_pack.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_size_t, C.c_void_p]
    lib.hp_unpack.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p]
    lib.hp_threads.argtypes = [C.c_size_t]
    lib.hp_threads.restype = C.c_uint
    return lib


def host_pack(hp, A):
    A = np.asfortranarray(A)
    rows, cols = A.shape
    cls = MELA_DOUBLE if A.dtype == np.float64 else MELA_LOGICAL
    ldw = (rows + 63) // 64
    out = np.full((max(cols, 1), max(ldw, 1)), 0xDEADBEEF, dtype=np.uint64)
    hp.hp_pack(A.ctypes.data, cls, rows, cols, out.ctypes.data)
    return out[:cols, :ldw]


def host_unpack(hp, P, rows, cols):
    P = np.ascontiguousarray(P, dtype=np.uint64)
    out = np.full(max(rows * cols, 1), 7, dtype=np.uint8)
    hp.hp_unpack(P.ctypes.data, rows, cols, out.ctypes.data)
    return out[: rows * cols].reshape((rows, cols), order="F")


def check_pack(hp, orc, A):
    got = host_pack(hp, A)
    want = orc.pack(np.asfortranarray(A).T)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


SHAPES = [(1, 1), (63, 3), (64, 2), (65, 5), (127, 1), (200, 130), (1000, 7), (3, 1000)]


@pytest.mark.parametrize("shape", SHAPES)
def test_pack_logical(hp, orc, shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    A = rng.integers(0, 2, shape).astype(bool)
    check_pack(hp, orc, A)
    # logical bytes other than 0 / 1 count as set
    B = rng.integers(0, 4, shape).astype(np.uint8) * 85
    check_pack(hp, orc, B)


@pytest.mark.parametrize("shape", SHAPES)
def test_pack_double(hp, orc, shape):
    rng = np.random.default_rng(shape[0] * 11 + shape[1])
    special = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 1.5, -1.5, 2.0, 3.0, -3.0, 2.9, -2.9,
                        2147483647.0, 2147483647.9, 2147483648.0, -2147483648.0,
                        -2147483648.5, -2147483649.0, 3e9, -3e9, 1e300, -1e300,
                        np.inf, -np.inf, np.nan, 4503599627370497.0, 5e-324],
                       dtype=np.float64)
    A = rng.choice(np.concatenate([special, rng.normal(0, 1e3, 64)]), size=shape)
    check_pack(hp, orc, A)
    want_bits = np.array([orc.parity_double(x) for x in A.ravel(order="F")], dtype=np.uint8)
    got_bits = host_unpack(hp, host_pack(hp, A), shape[0], shape[1]).ravel(order="F")
    assert np.array_equal(got_bits, want_bits)


@pytest.mark.parametrize("shape", SHAPES + [(0, 5), (5, 0)])
def test_unpack_inverts_pack(hp, shape):
    rng = np.random.default_rng(shape[0] * 13 + shape[1])
    A = rng.integers(0, 2, shape).astype(np.uint8)
    P = host_pack(hp, A)
    assert np.array_equal(host_unpack(hp, P, *shape), A)


def test_pool_threads_large_and_concurrent(hp, orc):
    """Arrays of several MiB go through the worker pool; calls from several
    threads at once (ctypes releases the GIL) take turns and stay exact."""
    assert hp.hp_threads(100) == 1
    assert hp.hp_threads(64 << 20) >= min(2, os.cpu_count() or 1)
    rng = np.random.default_rng(5)
    mats = [np.asfortranarray(rng.integers(0, 2, (2048 + 3 * i, 1500 + i)).astype(bool))
            for i in range(4)]
    want = [orc.pack(M.T) for M in mats]
    errors = []

    def worker(i):
        for _ in range(3):
            P = host_pack(hp, mats[i])
            if not np.array_equal(P, want[i]):
                errors.append(("pack", i))
            if not np.array_equal(host_unpack(hp, P, *mats[i].shape), mats[i]):
                errors.append(("unpack", i))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    D = np.asfortranarray(rng.integers(-5, 6, (1024, 700)).astype(np.float64) + 0.25)
    check_pack(hp, orc, D)
