"""Parity at the workloads' own sizes, through the C ABI.

* 32768 x 32768 x 32768 products (the benchmarked shape) and the rectangular
  65536 x 32768 x 4096 product, every algorithm: 64 rows of C compared bit
  for bit with the oracle's product of those rows of A with the whole B
  (mela_matmul_gf2.c:110-208 restated), plus Freivalds projections of the
  whole C (a wrong C survives each trial with probability <= 1/2).
* 16384 x 16384 = L * R (rank 12000) and 16384 x 20000 dense: rank, the
  reduced row echelon rows and the canonical null basis of every
  elimination variant against the oracle's RREF (mela_null_gf2.c:88-131):
  R equal row for row, Z(free,:) = I and Z(pivot i,:) = R(i, free), which
  is the unique basis gf2null.m:53-55 describes.
* the device pack / unpack kernels against the oracle's packers.
"""
import os
import subprocess
import threading

import numpy as np
import pytest

from tests import _brute as bf

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mela():
    import mex_linear_algebra_amd as m
    m.default_context()
    return m


@pytest.fixture(scope="module")
def ctx(mela):
    return mela.default_context()


def _unpack_rows(P, cols):
    """packed rows -> bool matrix (rows x cols)."""
    P = np.ascontiguousarray(P, dtype=np.uint64)
    return np.unpackbits(P.view(np.uint8), axis=1, bitorder="little")[:, :cols].astype(bool)


def _row_sample_exact(orc, Ap, Bp, Cp, m, k, n, rows):
    want = orc.matmul_packed(np.ascontiguousarray(Ap[rows]), len(rows), k, Bp, n, "gf2")
    return np.array_equal(Cp[rows][:, : want.shape[1]], want)


ALGOS = [0, 1, 2, 3, 4]  # auto, AND+parity, M4RM, int8 MFMA, fp4 MFMA


@pytest.mark.parametrize("shape", [(32768, 32768, 32768), (65536, 32768, 4096)])
def test_matmul_full_size_every_algo(mela, ctx, orc, shape):
    m, k, n = shape
    print(f"[matmul full size] {m}x{k}x{n}", flush=True)
    dA, dB, dC = ctx.bitmat(m, k), ctx.bitmat(k, n), ctx.bitmat(m, n)
    ctx.gen_bernoulli(1, dA)
    ctx.gen_bernoulli(2, dB)
    Ap, Bp = ctx.download(dA), ctx.download(dB)
    rng = np.random.default_rng(m + n)
    rows = np.sort(rng.choice(m, 64, replace=False))
    rows[0], rows[-1] = 0, m - 1  # first and last row blocks too
    vecs = bf.freivalds_vectors(Ap, Bp, n, trials=8, seed=k)
    try:
        for algo in ALGOS:
            print(f"[matmul full size] {m}x{k}x{n} algo={algo}", flush=True)
            ctx.matmul_packed(dA, dB, dC, algo)
            Cp = ctx.download(dC)
            assert _row_sample_exact(orc, Ap, Bp, Cp, m, k, n, rows), algo
            assert bf.freivalds_check(Cp, vecs), algo
    finally:
        for M in (dA, dB, dC):
            ctx.free(M)


def _c2a(ctx):
    L, R = ctx.bitmat(16384, 12000), ctx.bitmat(12000, 16384)
    ctx.gen_bernoulli(11, L)
    ctx.gen_bernoulli(12, R)
    H = ctx.bitmat(16384, 16384)
    ctx.matmul_packed(L, R, H, 1)
    ctx.free(L)
    ctx.free(R)
    return H


def _c2b(ctx):
    D = ctx.bitmat(16384, 20000)
    ctx.gen_bernoulli(13, D)
    return D


def _check_elimination_full(ctx, mela, orc, H, m, n, want_rank):
    Hp = ctx.download(H)
    print(f"[C2 {m}x{n}] oracle rref", flush=True)
    M, opiv = orc.rref_packed(Hp, m, n, "gf2")
    r = len(opiv)
    assert r == want_rank
    opiv = np.asarray(opiv)
    free = np.setdiff1d(np.arange(n), opiv)
    f = n - r
    Rfree = _unpack_rows(M[:r], n)[:, free]  # R(1:r, free)
    vecs = None
    for variant in (mela.ELIM_GE, mela.ELIM_M4RI, mela.ELIM_AUTO, mela.ELIM_PANEL):
        print(f"[C2 {m}x{n}] variant={variant}", flush=True)
        assert ctx.rank_packed(H, variant) == r
        Rm, piv = ctx.rref_packed(H, variant)
        assert list(piv) == list(opiv)
        Rd = ctx.download(Rm)
        ctx.free(Rm)
        assert np.array_equal(Rd[:r, : M.shape[1]], M[:r]), "RREF rows"
        assert not Rd[r:].any()
        Z = ctx.null_packed(H, variant)
        try:
            assert Z.rows == n and Z.cols == f
            Zp = ctx.download(Z)
            Zb = _unpack_rows(Zp, f)
            assert np.array_equal(Zb[free], np.eye(f, dtype=bool)), "Z(free,:) = I"
            assert np.array_equal(Zb[opiv], Rfree), "Z(pivot,:) = R(:, free)"
            # and A * Z = 0 on the device (the product path on a tall Z)
            Cm = ctx.bitmat(m, f)
            ctx.matmul_packed(H, Z, Cm)
            assert not ctx.download(Cm).any()
            ctx.free(Cm)
            if vecs is None:  # A * (Z x) = 0 on the host
                vecs = bf.freivalds_vectors(Hp, Zp, f, trials=4, seed=variant)
            assert all(not abx.any() for _, abx in vecs)
        finally:
            ctx.free(Z)


def test_c2a_full_size_rank_rref_null_every_variant(ctx, mela, orc):
    H = _c2a(ctx)
    try:
        _check_elimination_full(ctx, mela, orc, H, 16384, 16384, 12000)
    finally:
        ctx.free(H)


def test_c2b_full_size_rank_rref_null_every_variant(ctx, mela, orc):
    D = _c2b(ctx)
    try:
        _check_elimination_full(ctx, mela, orc, D, 16384, 20000, 16384)
    finally:
        ctx.free(D)


def test_ge_rref_rows_strided_sweep(ctx, mela, orc):
    """3001 x 65536: each panel's sweep strides its rows over the grid (more
    than one pass per wave, an odd row count in early panels); the RREF rows
    of every variant match the oracle's row for row."""
    m, n = 3001, 65536
    P = orc.gen_bernoulli_packed(41, m, n)
    dA = ctx.upload(P, n)
    M, opiv = orc.rref_packed(P, m, n, "gf2")
    try:
        for variant in (mela.ELIM_GE, mela.ELIM_M4RI, mela.ELIM_AUTO):
            Rm, piv = ctx.rref_packed(dA, variant)
            Rd = ctx.download(Rm)
            ctx.free(Rm)
            assert list(piv) == list(opiv)
            assert np.array_equal(Rd[: len(piv), : M.shape[1]], M[: len(piv)]), variant
    finally:
        ctx.free(dA)


@pytest.mark.parametrize("shape", [(1, 1), (7, 3), (64, 64), (65, 129), (300, 1000), (1000, 70),
                                   (2049, 513)])
def test_pack_unpack_dev_match_oracle(ctx, orc, shape):
    m, n = shape
    rng = np.random.default_rng(m * 3 + n)
    L = rng.integers(0, 2, (m, n)).astype(bool)
    D = rng.integers(-9, 10, (m, n)).astype(np.float64) + rng.choice([0.0, 0.5, -0.25], (m, n))
    flat = D.reshape(-1)
    flat[:: 7] = np.nan
    flat[3:: 11] = np.inf
    flat[5:: 13] = -np.inf
    flat[2:: 17] = 2147483648.0
    flat[4:: 19] = -2147483649.0
    flat[6:: 23] = 2147483647.0
    flat[1:: 29] = -2147483648.0
    for X in (L, D):
        M = ctx.pack_dev(X)
        got = ctx.download(M)
        want = orc.pack(X, M.ld)
        assert np.array_equal(got, want), X.dtype
        back = ctx.unpack_dev(M)
        assert np.array_equal(back, orc.unpack(want, m, n)), X.dtype
        ctx.free(M)


def test_one_context_shared_by_two_threads(mela, orc):
    """Two host threads interleave null_packed and rank_packed of different
    shapes on one context (one context serialises its own calls, scratch
    regrowth included); every result matches the oracle."""
    c = mela.Context(0)
    A1 = orc.gen_bernoulli_packed(51, 700, 1500)
    A2 = orc.gen_bernoulli_packed(52, 1300, 2600)
    M1, p1 = orc.rref_packed(A1, 700, 1500, "gf2")
    Z1 = orc.null_from_rref(M1, 1500, p1)
    r2 = int(orc.rank_packed(A2, 1300, 2600, "gf2"))
    d1, d2 = c.upload(A1, 1500), c.upload(A2, 2600)
    errors = []

    def nulls():
        try:
            for _ in range(12):
                Z = c.null_packed(d1)
                got = _unpack_rows(c.download(Z), Z.cols)
                c.free(Z)
                if not np.array_equal(got, Z1):
                    errors.append("null basis differs")
        except Exception as e:  # reported below
            errors.append(repr(e))

    def ranks():
        try:
            for i in range(12):
                v = (mela.ELIM_GE, mela.ELIM_M4RI, mela.ELIM_AUTO)[i % 3]
                if c.rank_packed(d2, v) != r2:
                    errors.append("rank differs")
        except Exception as e:
            errors.append(repr(e))

    ts = [threading.Thread(target=nulls), threading.Thread(target=ranks)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    c.free(d1)
    c.free(d2)
    c.close()
    assert not errors, errors


def test_host_arrays_through_pageable_staging(mela, orc, monkeypatch):
    """With no pinned staging (MELA_FORCE_PAGEABLE=1 when the context is
    created) the host-array entry points copy through ordinary memory and
    still match the oracle."""
    monkeypatch.setenv("MELA_FORCE_PAGEABLE", "1")
    c = mela.Context(0)
    try:
        rng = np.random.default_rng(61)
        A = rng.integers(0, 2, (900, 1300)).astype(bool)
        B = rng.integers(-3, 4, (1300, 500)).astype(np.float64)
        assert np.array_equal(mela.mela_matmul_gf2(A, B, ctx=c), orc.matmul(A, B, "gf2"))
        assert np.array_equal(mela.mela_matmul_m4ri(A, B, ctx=c), orc.matmul(A, B, "gf2"))
        assert mela.mela_rank_gf2(A, ctx=c) == orc.rank(A, "gf2")
        assert np.array_equal(mela.mela_null_m4ri(A, ctx=c), orc.null(A, "gf2"))
    finally:
        c.close()


def test_c_harness_selftest():
    """The C harness (no Python, no MATLAB) runs its self-test of the six
    entry points against its own naive C answers."""
    hdir = os.path.join(ROOT, "mex_linear_algebra_amd", "harness")
    exe = os.path.join(hdir, "mela_run")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", hdir], check=True)
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.join(ROOT, "mex_linear_algebra_amd") + ":" + env.get(
        "LD_LIBRARY_PATH", "")
    out = subprocess.run([exe, "selftest"], capture_output=True, text=True, timeout=120, env=env)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
