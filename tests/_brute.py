"""Independent closed-form and brute-force answers used to pin the oracle.

Nothing here shares code with oracle/: matmul is numpy integer matmul mod 2,
rank is a plain row-reduction over numpy bool arrays (the algorithm of the
reference's pure-MATLAB gf2_rank_reference, run_ci_tests.m:117-148), and the
canonical null basis is read off a numpy RREF exactly as gf2null.m:47-57
builds it (Z(nopiv,:) = I, Z(pivcol,:) = R(1:r,nopiv)).
"""
import numpy as np


def bits(A):
    A = np.asarray(A)
    if A.dtype == np.float64:
        # (int)x & 1 for the in-range values the tests use
        return (np.trunc(A).astype(np.int64) & 1).astype(bool)
    return A.astype(bool)


def matmul(A, B):
    return (bits(A).astype(np.int64) @ bits(B).astype(np.int64)) % 2 == 1


def rref(A):
    """(R, pivot columns) over GF(2) by Gauss-Jordan on a bool array."""
    M = bits(A).copy()
    m, n = M.shape
    piv = []
    row = 0
    for col in range(n):
        if row >= m:
            break
        nz = np.nonzero(M[row:, col])[0]
        if nz.size == 0:
            continue
        p = row + nz[0]
        if p != row:
            M[[row, p]] = M[[p, row]]
        hit = np.nonzero(M[:, col])[0]
        hit = hit[hit != row]
        M[hit] ^= M[row]
        piv.append(col)
        row += 1
    return M, piv


def rank(A):
    return len(rref(A)[1])


def null_canonical(A):
    M = bits(A)
    n = M.shape[1]
    R, piv = rref(M)
    r = len(piv)
    free = [c for c in range(n) if c not in set(piv)]
    Z = np.zeros((n, n - r), dtype=bool)
    for j, f in enumerate(free):
        Z[f, j] = True
        for i, c in enumerate(piv):
            Z[c, j] = R[i, f]
    return Z


def pivot_columns_by_prefix_rank(A):
    """c is a pivot column iff rank(A[:, :c+1]) > rank(A[:, :c])."""
    M = bits(A)
    out, prev = [], 0
    for c in range(M.shape[1]):
        rc = rank(M[:, : c + 1])
        if rc > prev:
            out.append(c)
        prev = rc
    return out


def exhaustive_rank_4x4():
    """rank of all 2^16 4x4 matrices from the size of their row space."""
    idx = np.arange(1 << 16, dtype=np.int64)
    rows = [(idx >> (4 * r)) & 0xF for r in range(4)]
    span = np.zeros_like(idx)
    for combo in range(16):
        v = np.zeros_like(idx)
        for r in range(4):
            if (combo >> r) & 1:
                v ^= rows[r]
        span |= np.int64(1) << v
    size = np.array([bin(int(s)).count("1") for s in span])
    return idx, np.log2(size).astype(int)


def matrix_from_index_4x4(i):
    return np.array([[(i >> (4 * r + c)) & 1 for c in range(4)] for r in range(4)], dtype=bool)


def parity_rows(P, x, chunk=4096):
    """y_i = parity(popcount(P_i AND x)) for packed rows P (uint64) and packed
    x, in row chunks (bounded temporaries for 128 MiB operands)."""
    P = np.asarray(P)
    xw = np.asarray(x, dtype=np.uint64)[None, : P.shape[1]]
    out = np.zeros(P.shape[0], dtype=np.uint8)
    for r0 in range(0, P.shape[0], chunk):
        v = np.bitwise_count(np.bitwise_and(P[r0:r0 + chunk], xw))
        out[r0:r0 + chunk] = (v.sum(axis=1, dtype=np.int64) & 1).astype(np.uint8)
    return out


def pack_bits(bits):
    """bool vector -> packed uint64 words (bit j at word j>>6, bit j&63)."""
    bits = np.asarray(bits, dtype=np.uint8)
    w = (len(bits) + 63) // 64
    pad = np.zeros(w * 64, dtype=np.uint8)
    pad[: len(bits)] = bits
    return np.packbits(pad, bitorder="little").view("<u8").astype(np.uint64)


def freivalds_vectors(Ap, Bp, n, trials, seed):
    """[(packed x, A (B x))] for `trials` random x: shared by the checks of
    several products of the same A and B."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(trials):
        xp = pack_bits(rng.integers(0, 2, n).astype(np.uint8))
        out.append((xp, parity_rows(Ap, pack_bits(parity_rows(Bp, xp)))))
    return out


def freivalds_check(Cp, vectors):
    """C x == A (B x) for every precomputed (x, A (B x)) pair."""
    return all(np.array_equal(parity_rows(Cp, xp), abx) for xp, abx in vectors)


def freivalds(Ap, Bp, Cp, k, n, trials=4, seed=0):
    """Checks C = A*B over GF(2) on packed operands by random projections:
    C x == A (B x) for random x (a wrong C survives one trial w.p. <= 1/2)."""
    rng = np.random.default_rng(seed)
    for _ in range(trials):
        x = rng.integers(0, 2, n).astype(np.uint8)
        xp = pack_bits(x)
        bx = parity_rows(Bp, xp)          # k bits
        abx = parity_rows(Ap, pack_bits(bx))
        cx = parity_rows(Cp, xp)
        if not np.array_equal(abx, cx):
            return False
    return True
