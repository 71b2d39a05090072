"""Pins the CPU oracle (oracle/mela_oracle.c) before it is trusted as the
checker: the reference's own known-answer cases (tests/golden/kats.json),
closed forms (numpy matmul mod 2), an independent row-reduction, invariants
(rank(A) = rank(A^T), GE = M4RI, nullity = n - rank, A*Z = 0, Z(free,:) = I)
and exhaustive enumeration of every 4x4 matrix."""
import json
import math
import os

import numpy as np
import pytest

from tests import _brute as bf

GOLD = os.path.join(os.path.dirname(__file__), "golden", "kats.json")


def _kat_matrix(case, key="A"):
    if "zeros" in case and key == "A":
        return np.zeros(case["zeros"], dtype=bool)
    if "eye" in case and key == "A":
        return np.eye(case["eye"], dtype=bool)
    if "zerosA" in case and key == "A":
        return np.zeros(case["zerosA"], dtype=bool)
    M = np.array(case[key])
    return M.astype(np.float64) if case["kind"] == "double" else M.astype(bool)


def _kats(op):
    with open(GOLD) as f:
        return [c for c in json.load(f) if c["op"] == op]


@pytest.mark.parametrize("method", ["gf2", "m4ri"])
def test_kat_rank(orc, method):
    for case in _kats("rank"):
        assert orc.rank(_kat_matrix(case), method) == case["rank"], case["src"]


@pytest.mark.parametrize("method", ["gf2", "m4ri"])
def test_kat_null(orc, method):
    for case in _kats("null"):
        Z = orc.null(_kat_matrix(case), method)
        assert Z.dtype == bool
        assert np.array_equal(Z, np.array(case["Z"], dtype=bool)), case["src"]


def test_kat_rref(orc):
    # the standard RREF (null_gf2.c:88-131) leaves gf2rref's R and jb
    for case in _kats("rref"):
        A = _kat_matrix(case)
        m, n = A.shape
        M, piv = orc.rref_packed(orc.pack(A), m, n, "gf2")
        assert np.array_equal(orc.unpack(M, m, n), np.array(case["R"], dtype=bool)), case["src"]
        assert [int(c) + 1 for c in piv] == case["jb"], case["src"]


@pytest.mark.parametrize("method", ["gf2", "m4ri"])
def test_kat_matmul(orc, method):
    rng = np.random.default_rng(42)
    for case in _kats("matmul"):
        A = _kat_matrix(case)
        if "B" in case:
            B = _kat_matrix(case, "B")
        else:
            B = rng.integers(0, 2, case["Bshape"]).astype(bool)
        C = orc.matmul(A, B, method)
        assert np.array_equal(C, np.array(case["C"], dtype=bool)), case["src"]


def test_parity_double_in_range(orc):
    # (int)x & 1 is defined C for |x| < 2^31: truncate toward zero, low bit
    table = {0.0: 0, -0.0: 0, 1.0: 1, 2.0: 0, 3.0: 1, -1.0: 1, -2.0: 0, -3.0: 1,
             2.7: 0, 3.5: 1, -1.5: 1, -0.5: 0, 0.999: 0, 1e9 + 1: 1,
             2147483647.0: 1, 2147483647.9: 1, -2147483647.0: 1,
             -2147483648.0: 0, -2147483648.5: 0, 5e-324: 0}
    for x, want in table.items():
        assert orc.parity_double(x) == want, x


def test_parity_double_out_of_range_x86_convention(orc):
    # undefined behaviour in C; x86 cvttsd2si yields 0x80000000 (parity 0).
    # Parity unpinned against the reference binary (it cannot be built here).
    for x in (2147483648.0, 4294967297.0, -4294967297.0, 1e300, float("inf"),
              float("-inf"), float("nan")):
        assert orc.parity_double(x) == 0, x


@pytest.mark.parametrize("shape", [(37, 23, 31), (1, 1, 1), (0, 5, 3), (5, 0, 3), (5, 7, 0),
                                   (100, 50, 30), (256, 128, 64), (300, 65, 129), (64, 64, 64),
                                   (257, 9, 200)])
def test_matmul_closed_form(orc, shape):
    m, k, n = shape
    rng = np.random.default_rng(m * 1000 + k * 10 + n)
    A = rng.integers(0, 2, (m, k)).astype(bool)
    B = rng.integers(0, 2, (k, n)).astype(bool)
    want = bf.matmul(A, B)
    assert np.array_equal(orc.matmul(A, B, "gf2"), want)
    assert np.array_equal(orc.matmul(A, B, "m4ri"), want)
    # double inputs with odd/even integers (parity semantics)
    Ad = A * rng.integers(1, 4, A.shape) * 2 - A  # odd where A, even elsewhere
    Bd = rng.integers(-7, 8, B.shape).astype(np.float64)
    assert np.array_equal(orc.matmul(Ad.astype(np.float64), Bd, "gf2"), bf.matmul(Ad.astype(np.float64), Bd))
    assert np.array_equal(orc.matmul(Ad.astype(np.float64), Bd, "m4ri"), bf.matmul(Ad.astype(np.float64), Bd))


def test_matmul_packed_forms(orc):
    rng = np.random.default_rng(7)
    for m, k, n in [(70, 130, 65), (256, 256, 256), (3, 200, 1), (513, 64, 127)]:
        A = rng.integers(0, 2, (m, k)).astype(bool)
        B = rng.integers(0, 2, (k, n)).astype(bool)
        want = orc.pack(bf.matmul(A, B))
        for meth in ("gf2", "m4ri"):
            got = orc.matmul_packed(orc.pack(A), m, k, orc.pack(B), n, meth)
            assert np.array_equal(got, want), (m, k, n, meth)


def test_rank_exhaustive_4x4(orc):
    idx, want = bf.exhaustive_rank_4x4()
    # spot the closed-form distribution: #invertible 4x4 over GF(2) = 20160
    assert int((want == 4).sum()) == 20160
    # a stride through the 2^16 cases via the MATLAB-layout entry points
    for i in range(0, 1 << 16, 97):
        A = bf.matrix_from_index_4x4(i)
        assert orc.rank(A, "gf2") == want[i]
        assert orc.rank(A, "m4ri") == want[i]


def test_rank_exhaustive_4x4_packed_all(orc):
    idx, want = bf.exhaustive_rank_4x4()
    for i in range(1 << 16):
        P = np.array([[(i >> (4 * r)) & 0xF] for r in range(4)], dtype=np.uint64)
        if orc.rank_packed(P, 4, 4, "gf2") != want[i] or orc.rank_packed(P, 4, 4, "m4ri") != want[i]:
            raise AssertionError(f"rank mismatch on 4x4 matrix {i}")


@pytest.mark.parametrize("m,n", [(17, 29), (64, 64), (50, 50), (31, 47), (16, 32), (1, 70),
                                 (70, 1), (130, 200), (200, 130), (0, 5), (5, 0), (8, 8)])
def test_rank_vs_independent_reduction(orc, m, n):
    rng = np.random.default_rng(m * 31 + n)
    A = rng.integers(0, 2, (m, n)).astype(bool)
    want = bf.rank(A)
    assert orc.rank(A, "gf2") == want
    assert orc.rank(A, "m4ri") == want
    assert orc.rank(A.T.copy(), "gf2") == want


@pytest.mark.parametrize("m,n,r", [(120, 150, 40), (200, 200, 199), (64, 300, 10), (300, 64, 63),
                                   (257, 513, 100)])
def test_rank_low_rank_products(orc, m, n, r):
    rng = np.random.default_rng(m + n + r)
    L = rng.integers(0, 2, (m, r)).astype(bool)
    R = rng.integers(0, 2, (r, n)).astype(bool)
    H = bf.matmul(L, R)
    want = bf.rank(H)
    assert want <= r
    assert orc.rank(H, "gf2") == want
    assert orc.rank(H, "m4ri") == want


def _check_canonical_null(A, Z):
    M = bf.bits(A)
    n = M.shape[1]
    r = bf.rank(M)
    assert Z.shape == (n, n - r)
    assert not bf.matmul(M, Z).any()
    piv = bf.pivot_columns_by_prefix_rank(M)
    free = [c for c in range(n) if c not in set(piv)]
    assert np.array_equal(Z[free, :], np.eye(n - r, dtype=bool))


@pytest.mark.parametrize("m,n", [(16, 32), (31, 47), (50, 100), (2, 4), (33, 33), (100, 50),
                                 (0, 6), (6, 0), (1, 1), (64, 130)])
def test_null_canonical(orc, m, n):
    rng = np.random.default_rng(m * 7 + n)
    A = rng.integers(0, 2, (m, n)).astype(bool)
    want = bf.null_canonical(A)
    for meth in ("gf2", "m4ri"):
        Z = orc.null(A, meth)
        assert np.array_equal(Z, want), meth
    if m and n:
        _check_canonical_null(A, want)


def test_null_low_rank_and_doubles(orc):
    rng = np.random.default_rng(3)
    L = rng.integers(0, 2, (90, 20)).astype(bool)
    R = rng.integers(0, 2, (20, 140)).astype(bool)
    H = bf.matmul(L, R)
    _check_canonical_null(H, orc.null(H, "gf2"))
    assert np.array_equal(orc.null(H, "gf2"), orc.null(H, "m4ri"))
    Hd = H * 2.0 + 1.0 - 1.0 * H  # odd where H, even elsewhere (values 1 or 2)
    Hd = np.where(H, 3.0, -4.0)
    assert np.array_equal(orc.null(Hd, "m4ri"), orc.null(H, "gf2"))


def test_null_exhaustive_3x3(orc):
    for i in range(1 << 9):
        A = np.array([[(i >> (3 * r + c)) & 1 for c in range(3)] for r in range(3)], dtype=bool)
        want = bf.null_canonical(A)
        assert np.array_equal(orc.null(A, "gf2"), want), i
        assert np.array_equal(orc.null(A, "m4ri"), want), i


def test_rref_forms_agree(orc):
    # GE RREF is fully reduced; the M4RI echelon is not reduced inside a block,
    # but both give the same pivots and the same canonical null basis.
    rng = np.random.default_rng(11)
    A = rng.integers(0, 2, (80, 150)).astype(bool)
    P = orc.pack(A)
    M1, p1 = orc.rref_packed(P, 80, 150, "gf2")
    M2, p2 = orc.rref_packed(P, 80, 150, "m4ri")
    assert list(p1) == list(p2) == bf.rref(A)[1]
    R, _ = bf.rref(A)
    assert np.array_equal(orc.unpack(M1, 80, 150), R)
    assert np.array_equal(orc.null_from_rref(M1, 150, p1), orc.null_from_rref(M2, 150, p2))


def test_generator_words(orc):
    P = orc.gen_bernoulli_packed(5, 3, 130)
    assert P.shape == (3, 3)
    assert int(P[0, 2]) >> 2 == 0  # bits past column 130 are clear
    sm = orc.load().orc_splitmix64
    k1, k2 = int(sm(5)), int(sm(5 ^ 0xD1B54A32D192ED03))
    M64 = 2**64 - 1
    assert int(P[1, 1]) == int(sm(int(sm((k1 + 1 * 3 + 1) & M64)) ^ k2))
    # splitmix64 known answer for seed 0 (first output of the published generator)
    assert int(sm(0)) == 0xE220A8397B1DCDAF
    ones = np.unpackbits(orc.gen_bernoulli_packed(9, 256, 256).view(np.uint8)).mean()
    assert abs(ones - 0.5) < 0.01


def test_generator_has_no_gf2_column_dependences(orc):
    # a random 4000 x 3900 bit matrix has full column rank except with
    # probability ~2^-100; a single hash round on a counter failed this
    P = orc.gen_bernoulli_packed(12, 4000, 3900)
    assert orc.rank_packed(P, 4000, 3900, "m4ri") == 3900
