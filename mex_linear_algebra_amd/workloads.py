"""Seeded synthetic workloads of the benchmark configurations (host side).

Packed layout everywhere: row-major uint64 words, bit j of a row at word j>>6,
bit j&63.  The generators are deterministic in their seed and portable (numpy
PCG64), so the CPU baseline and the device path see identical inputs.
"""
from __future__ import annotations

import numpy as np


def words(n: int) -> int:
    return (n + 63) // 64


def pack_rows(bits: np.ndarray) -> np.ndarray:
    """bool (m, n) -> packed uint64 (m, ceil(n/64))."""
    m, n = bits.shape
    w = words(n)
    pad = np.zeros((m, w * 64), dtype=np.uint8)
    pad[:, :n] = bits
    return np.packbits(pad, axis=1, bitorder="little").view("<u8").astype(np.uint64).reshape(m, w)


def parity_check_matrix(rng, m: int, n: int, w: int) -> np.ndarray:
    """m x n, every column with exactly w ones at distinct uniform rows, packed."""
    w = min(w, m)
    rows = rng.integers(0, m, size=(n, w))
    while True:  # redraw the columns whose w rows collide
        srt = np.sort(rows, axis=1)
        bad = np.nonzero((srt[:, 1:] == srt[:, :-1]).any(axis=1))[0] if w > 1 else []
        if len(bad) == 0:
            break
        rows[bad] = rng.integers(0, m, size=(len(bad), w))
    P = np.zeros((m, words(n)), dtype=np.uint64)
    cols = np.repeat(np.arange(n), w)
    np.bitwise_or.at(P, (rows.reshape(-1), cols >> 6),
                     np.left_shift(np.uint64(1), (cols & 63).astype(np.uint64)))
    return P


def dense_matrix(rng, m: int, n: int) -> np.ndarray:
    """m x n Bernoulli(1/2), packed (bits past n cleared)."""
    P = rng.integers(0, 2**63, size=(m, words(n)), dtype=np.uint64) ^ \
        (rng.integers(0, 2, size=(m, words(n)), dtype=np.uint64) << np.uint64(63))
    if n % 64:
        P[:, -1] &= np.uint64((1 << (n % 64)) - 1)
    return P


def batch_c4(count: int = 4096, seed: int = 4, mmin: int = 64, mmax: int = 2048):
    """Uneven rank/null batch: m = floor(exp(U[ln mmin, ln mmax])), n = m + U{0..m};
    90% parity-check-like (column weight w in {3,4,6}), 10% dense Bernoulli(1/2).
    Returns a list of (packed matrix, m, n, kind)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        m = int(np.floor(np.exp(rng.uniform(np.log(mmin), np.log(mmax)))))
        n = m + int(rng.integers(0, m + 1))
        if rng.random() < 0.9:
            w = int(rng.choice([3, 4, 6]))
            out.append((parity_check_matrix(rng, m, n, min(w, m)), m, n, f"sparse{w}"))
        else:
            out.append((dense_matrix(rng, m, n), m, n, "dense"))
    return out


def batch_matmul_items(count: int = 2048, seed: int = 5, mmin: int = 64, mmax: int = 2048,
                       bmin: int = 32, bmax: int = 1024):
    """Uneven product batch C_i = A_i * B_i: A_i as in batch_c4 (m log-uniform,
    k = m + U{0..m}, 90% column weight 3/4/6, 10% dense), B_i a dense k x b
    block with b log-uniform in [bmin, bmax] (e.g. syndromes of b error
    patterns).  Returns a list of (A packed, m, k, B packed, b, kind)."""
    rng = np.random.default_rng(seed)
    out = []
    for P, m, k, kind in batch_c4(count, seed=seed, mmin=mmin, mmax=mmax):
        b = int(np.floor(np.exp(rng.uniform(np.log(bmin), np.log(bmax)))))
        out.append((P, m, k, dense_matrix(rng, k, b), b, kind))
    return out
