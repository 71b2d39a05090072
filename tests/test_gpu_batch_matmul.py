"""Batched uneven products in one persistent launch (mela_matmul_batch_*),
item by item against the oracle's product (mela_matmul_gf2.c:110-208
restated): the planned per-item engines (AND+parity, row XOR, fp4 matrix-core
tiles), each engine forced on every item, and replays of the captured graph
with other work in between."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mela():
    import mex_linear_algebra_amd as m
    return m


def _items(W, rng):
    items = W.batch_matmul_items(300, seed=7, mmin=16, mmax=1100, bmin=1, bmax=700)
    # edges: empty dimensions, k not a multiple of 64, single rows/columns,
    # a wide dense product
    extra = [(0, 5, 7), (5, 0, 7), (5, 7, 0), (1, 1, 1), (70, 129, 1), (3, 4097, 65), (257, 65, 1500)]
    for m, k, n in extra:
        A = W.dense_matrix(rng, m, k) if m and k else np.zeros((m, W.words(k)), np.uint64)
        B = W.dense_matrix(rng, k, n) if k and n else np.zeros((k, W.words(n)), np.uint64)
        items.append((A, m, k, B, n, "edge"))
    return items


def _upload(ctx, items):
    dA, dB, dC = [], [], []
    for A, m, k, B, n, _ in items:
        dA.append(ctx.upload(A, k) if A.size else ctx.bitmat(m, k))
        dB.append(ctx.upload(B, n) if B.size else ctx.bitmat(k, n))
        dC.append(ctx.bitmat(m, n))
    return dA, dB, dC


def _check(ctx, orc, items, dC, tag):
    bad = []
    for i, ((A, m, k, B, n, kind), C) in enumerate(zip(items, dC)):
        got = ctx.download(C)
        want = orc.matmul_packed(A, m, k, B, n, "gf2") if m and n else np.zeros((m, max(1, (n + 63) // 64)), np.uint64)
        if m and n and not np.array_equal(got[:, : want.shape[1]], want):
            bad.append((i, m, k, n, kind))
    assert not bad, (tag, bad[:5])


def test_batch_matmul_matches_oracle(mela, orc):
    from mex_linear_algebra_amd import workloads as W
    ctx = mela.Context(0)
    rng = np.random.default_rng(3)
    items = _items(W, rng)
    dA, dB, dC = _upload(ctx, items)
    b = None
    try:
        b = mela.MatmulBatch(ctx, dA, dB, dC)
        algos = b.algos()
        assert set(algos) <= {mela.ALGO_ANDPOP, mela.ALGO_ROWXOR, mela.ALGO_MFMA_FP4}
        assert mela.ALGO_ROWXOR in algos and mela.ALGO_MFMA_FP4 in algos
        for rep in range(3):
            print(f"[batch matmul] planned, replay {rep}", flush=True)
            for C in dC:  # clear the outputs between replays
                ctx.gen_bernoulli(rep + 100, C)
            b.run()
            # other work on the same stream between replays
            x = ctx.bitmat(300, 200)
            ctx.gen_bernoulli(9, x)
            ctx.free(x)
            ctx.sync()
            _check(ctx, orc, items, dC, f"planned replay {rep}")
        for forced in (mela.ALGO_ANDPOP, mela.ALGO_ROWXOR, mela.ALGO_MFMA_FP4):
            print(f"[batch matmul] forced {forced}", flush=True)
            b.set_algos([forced] * len(items))
            for C in dC:
                ctx.gen_bernoulli(7, C)
            b.run()
            ctx.sync()
            _check(ctx, orc, items, dC, f"forced {forced}")
    finally:
        if b is not None:
            b.close()
        for M in dA + dB + dC:
            ctx.free(M)
        ctx.close()


def test_batch_matmul_plans_by_density(mela):
    """A clearly sparse A (column weight 3) gets the row-XOR product, a dense
    one the matrix-core tiles; engines outside the three are refused."""
    from mex_linear_algebra_amd import workloads as W
    ctx = mela.Context(0)
    rng = np.random.default_rng(11)
    mats = []
    try:
        for m, k, n, sparse in ((2000, 2500, 600, True), (1500, 1500, 1500, False)):
            A = W.parity_check_matrix(rng, m, k, 3) if sparse else W.dense_matrix(rng, m, k)
            mats.append((ctx.upload(A, k), ctx.bitmat(k, n), ctx.bitmat(m, n)))
            ctx.gen_bernoulli(5, mats[-1][1])
        b = mela.MatmulBatch(ctx, [x[0] for x in mats], [x[1] for x in mats], [x[2] for x in mats])
        assert b.algos() == [mela.ALGO_ROWXOR, mela.ALGO_MFMA_FP4]
        with pytest.raises(mela.MelaError, match="batched products take"):
            b.set_algos([mela.ALGO_M4RM, mela.ALGO_ANDPOP])
        b.close()
    finally:
        for x in mats:
            for M in x:
                ctx.free(M)
        ctx.close()


def test_batch_matmul_validation(mela):
    ctx = mela.Context(0)
    try:
        A, B, Cm = ctx.bitmat(10, 20), ctx.bitmat(21, 5), ctx.bitmat(10, 5)
        with pytest.raises(mela.MelaError, match="Inner dimensions must agree."):
            mela.MatmulBatch(ctx, [A], [B], [Cm])
        B2 = ctx.bitmat(20, 5)
        C2 = ctx.bitmat(10, 6)
        with pytest.raises(mela.MelaError, match="output shape"):
            mela.MatmulBatch(ctx, [A], [B2], [C2])
        b = mela.MatmulBatch(ctx, [], [], [])
        b.run()
        b.close()
        for M in (A, B, Cm, B2, C2):
            ctx.free(M)
    finally:
        ctx.close()
