"""The blocked elimination path (512-column blocks factored from the head of
the active rows, matrix-core trailing update; csrc/elim_block.hip) against
the oracle: rank, the RREF rows with their pivot columns and the canonical
null basis (mela_rank_m4ri.c:94-279, mela_null_gf2.c:88-274 restated in
oracle/mela_oracle.c).  Shapes cover dense full-rank and low-rank matrices,
ragged widths, heads that hold every row, and matrices whose pivots the head
misses (sparse, zero head, rows that repeat), where the call falls back to
the panel path; every result must be the oracle's whichever path ran."""
import numpy as np
import pytest

from mex_linear_algebra_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mela():
    import mex_linear_algebra_amd as m
    return m


@pytest.fixture(scope="module")
def ctx(mela):
    return mela.Context(0)


def _bits(P, cols):
    P = np.ascontiguousarray(P, dtype=np.uint64)
    return np.unpackbits(P.view(np.uint8), axis=1, bitorder="little")[:, :cols].astype(bool)


def _packed(A):
    m, n = A.shape
    wc = (n + 63) // 64
    B = np.zeros((m, wc * 64), dtype=bool)
    B[:, :n] = A
    return np.packbits(B, axis=1, bitorder="little").view(np.uint64).reshape(m, wc)


def _dense(rng, m, n):
    return rng.integers(0, 2, (m, n)).astype(bool)


def _low_rank(rng, m, n, r):
    L = rng.integers(0, 2, (m, r)).astype(np.int64)
    R = rng.integers(0, 2, (r, n)).astype(np.int64)
    return ((L @ R) % 2).astype(bool)


def _check(ctx, mela, orc, A, variant, tag):
    """Rank, RREF and null basis of A by `variant` against the oracle;
    returns how many of the three calls finished on the blocked path and how
    many fell back to the panel path."""
    c0 = ctx.elim_counts()
    m, n = A.shape
    P = _packed(A)
    print(f"[blocked] {tag} {m}x{n} variant={variant}", flush=True)
    M, opiv = orc.rref_packed(P, m, n, "gf2")
    r = len(opiv)
    dA = ctx.upload(P, n)
    try:
        assert ctx.rank_packed(dA, variant) == r, "rank"
        Rm, piv = ctx.rref_packed(dA, variant)
        Rd = ctx.download(Rm)
        ctx.free(Rm)
        assert list(piv) == list(opiv), "pivot columns"
        assert np.array_equal(Rd[:r, : M.shape[1]], M[:r]), "RREF rows"
        assert not Rd[r:].any()
        Z = ctx.null_packed(dA, variant)
        try:
            f = n - r
            assert Z.rows == n and Z.cols == f
            Zb = _bits(ctx.download(Z), f)
            free = np.setdiff1d(np.arange(n), np.asarray(opiv, dtype=np.int64))
            assert np.array_equal(Zb[free], np.eye(f, dtype=bool)), "Z(free,:) = I"
            assert np.array_equal(Zb[np.asarray(opiv, dtype=np.int64)], _bits(M[:r], n)[:, free])
        finally:
            ctx.free(Z)
    finally:
        ctx.free(dA)
    c1 = ctx.elim_counts()
    return c1[0] - c0[0], c1[1] - c0[1]


DENSE = [(600, 600), (1024, 1024), (1500, 2600), (2600, 1500), (2048, 2048), (700, 4100),
         (3000, 3001), (4096, 520), (513, 5000), (100, 3000)]


@pytest.mark.parametrize("shape", DENSE)
def test_blocked_dense_matches_oracle(ctx, mela, orc, shape):
    m, n = shape
    A = _dense(np.random.default_rng(m * 7 + n), m, n)
    assert _check(ctx, mela, orc, A, mela.ELIM_BLOCKED, "dense") == (3, 0)


@pytest.mark.parametrize("m,n,r", [(2048, 2048, 1500), (3000, 2500, 700), (1200, 4000, 1100),
                                   (2500, 2500, 512), (2000, 3000, 1)])
def test_blocked_low_rank_matches_oracle(ctx, mela, orc, m, n, r):
    A = _low_rank(np.random.default_rng(r + m), m, n, r)
    assert _check(ctx, mela, orc, A, mela.ELIM_BLOCKED, f"rank<={r}") == (3, 0)


def test_blocked_head_misses_pivots(ctx, mela, orc):
    """Pivots the head cannot see: a zero head, column-weight-3 rows, and a
    block whose pivot columns only the last rows have."""
    rng = np.random.default_rng(5)
    A = _dense(rng, 2000, 1800)
    A[:700] = False  # the first 576 active rows are zero
    assert _check(ctx, mela, orc, A, mela.ELIM_BLOCKED, "zero head")[1] >= 1
    H = W.parity_check_matrix(rng, 1800, 2400, 3)
    assert _check(ctx, mela, orc, _bits(H, 2400), mela.ELIM_BLOCKED, "column weight 3")[1] >= 1
    B = _dense(rng, 1500, 1600)
    B[:1400, 100:200] = False  # block 0's columns 100..199 live in the last 100 rows only
    assert _check(ctx, mela, orc, B, mela.ELIM_BLOCKED, "late pivots")[1] >= 1
    C = _dense(rng, 300, 1300)
    C = np.concatenate([C, C, C, C, C, C, C, C], axis=0)  # rows repeat: rank 300
    # the head (the first 576 rows) holds all 300 distinct rows
    assert _check(ctx, mela, orc, C, mela.ELIM_BLOCKED, "repeated rows") == (3, 0)


def test_blocked_auto_and_panel_agree(ctx, mela, orc):
    """AUTO takes the blocked path from 2048 x 2048 on; PANEL forces the panel
    path; both give the oracle's results, and a shape whose blocked attempt
    fell back keeps working on later calls."""
    rng = np.random.default_rng(17)
    A = _low_rank(rng, 2304, 2600, 2000)
    assert _check(ctx, mela, orc, A, mela.ELIM_AUTO, "auto") == (3, 0)
    assert _check(ctx, mela, orc, A, mela.ELIM_PANEL, "panel") == (0, 0)
    assert _check(ctx, mela, orc, A, mela.ELIM_AUTO, "auto") == (3, 0)
    S = _bits(W.parity_check_matrix(rng, 2100, 3000, 3), 3000)
    # the first round falls back (rank, and RREF whose shape entry the null
    # call shares); later rounds take the panel path at once
    nb, nf = _check(ctx, mela, orc, S, mela.ELIM_AUTO, "sparse auto")
    assert nb == 0 and nf >= 1
    for _ in range(2):
        assert _check(ctx, mela, orc, S, mela.ELIM_AUTO, "sparse auto") == (0, 0)


def test_blocked_repeated_calls_replay(ctx, mela, orc):
    """Calls of one shape are captured and replayed; different matrices of
    the shape (one of them of lower rank, one needing the fallback) all give
    the oracle's results."""
    rng = np.random.default_rng(23)
    mats = [_dense(rng, 2048, 2100), _low_rank(rng, 2048, 2100, 900), _dense(rng, 2048, 2100)]
    z = _dense(rng, 2048, 2100)
    z[:800] = False
    mats.append(z)
    mats.append(_dense(rng, 2048, 2100))
    for A in mats:
        P = _packed(A)
        M, opiv = orc.rref_packed(P, 2048, 2100, "gf2")
        dA = ctx.upload(P, 2100)
        try:
            for _ in range(3):
                assert ctx.rank_packed(dA, mela.ELIM_BLOCKED) == len(opiv)
            Rm, piv = ctx.rref_packed(dA, mela.ELIM_BLOCKED)
            Rd = ctx.download(Rm)
            ctx.free(Rm)
            assert list(piv) == list(opiv)
            assert np.array_equal(Rd[: len(opiv), : M.shape[1]], M[: len(opiv)])
        finally:
            ctx.free(dA)


def test_blocked_replay_after_the_block_state_moved(mela, orc):
    """A shape's rank sequence is captured, then an RREF of the same shape
    grows the block state (it holds the earlier pivot rows' operands too), so
    that buffer moves: later rank calls of the shape take a sequence of their
    own instead of replaying launches on the old block state.  A fresh
    context, so the RREF is what grows the buffer."""
    rng = np.random.default_rng(29)
    A = _dense(rng, 2048, 2048)
    L = _low_rank(rng, 2048, 2048, 1000)
    c = mela.Context(0)
    try:
        for X in (A, L, A):
            P = _packed(X)
            M, opiv = orc.rref_packed(P, 2048, 2048, "gf2")
            r = len(opiv)
            print(f"[blocked] moved block state, rank {r}", flush=True)
            dA = c.upload(P, 2048)
            try:
                for _ in range(3):
                    assert c.rank_packed(dA, mela.ELIM_BLOCKED) == r
                Rm, piv = c.rref_packed(dA, mela.ELIM_BLOCKED)
                Rd = c.download(Rm)
                c.free(Rm)
                assert list(piv) == list(opiv)
                assert np.array_equal(Rd[:r, : M.shape[1]], M[:r])
                for _ in range(3):
                    assert c.rank_packed(dA, mela.ELIM_BLOCKED) == r
            finally:
                c.free(dA)
        blocked, fallback, _ = c.elim_counts()
        assert fallback == 0 and blocked == 21
    finally:
        c.close()
