"""The allocation-free rank entry point (mela_rank_packed_noalloc, no
reference counterpart) against the existing entry point mela_rank_packed and
the oracle, through the C ABI: two calls in a row with the second input
smaller (the shared device array keeps the first call's contents), every
elimination variant, and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mela():
    import mex_linear_algebra_amd as m
    return m


def _noalloc(ctx, A, variant):
    r = C.c_size_t()
    st = ctx.lib.mela_rank_packed_noalloc(ctx.handle, C.byref(A), variant, C.byref(r))
    return st, r.value


def test_noalloc_rank_matches_rank_packed(mela, orc):
    ctx = mela.Context(0)
    rng = np.random.default_rng(71)
    # a rank-1200 3000 x 5000 product, then a smaller full-rank 700 x 900
    # matrix, then the first one again
    L = rng.integers(0, 2, (3000, 1200)).astype(bool)
    R = rng.integers(0, 2, (1200, 5000)).astype(bool)
    H1 = orc.matmul(L, R, "gf2")
    P1 = orc.pack(H1)
    P2 = orc.gen_bernoulli_packed(72, 700, 900)
    want1 = int(orc.rank_packed(P1, 3000, 5000, "gf2"))
    want2 = int(orc.rank_packed(P2, 700, 900, "gf2"))
    assert want1 == 1200 and want2 == 700
    d1, d2 = ctx.upload(P1, 5000), ctx.upload(P2, 900)
    try:
        for variant in (mela.ELIM_GE, mela.ELIM_M4RI, mela.ELIM_AUTO):
            print(f"[noalloc] variant={variant}", flush=True)
            for d, want in ((d1, want1), (d2, want2), (d1, want1)):
                st, r = _noalloc(ctx, d, variant)
                assert st == 0, ctx.lib.mela_last_error()
                assert r == want
                assert r == ctx.rank_packed(d, variant)
    finally:
        ctx.free(d1)
        ctx.free(d2)
        ctx.close()


def test_noalloc_rank_size_limit_and_edges(mela):
    ctx = mela.Context(0)
    try:
        # validated before any device work (the data pointer is never read)
        big = mela.BitMatrix(1, 16384, 16384, 256)
        st, _ = _noalloc(ctx, big, mela.ELIM_AUTO)
        assert st == 5
        assert b"allocation-free rank holds at most" in ctx.lib.mela_last_error()
        st, _ = _noalloc(ctx, big, 7)
        assert st == 5 and b"unknown elimination variant" in ctx.lib.mela_last_error()
        st, r = _noalloc(ctx, mela.BitMatrix(None, 0, 5, 1), mela.ELIM_AUTO)
        assert st == 0 and r == 0
        # the largest square shape that fits: 8192 x 8192 (16 MiB of rows)
        M = ctx.bitmat(8192, 8192)
        ctx.gen_bernoulli(73, M)
        st, r = _noalloc(ctx, M, mela.ELIM_AUTO)
        assert st == 0 and r == ctx.rank_packed(M) and r >= 8190
        ctx.free(M)
    finally:
        ctx.close()
