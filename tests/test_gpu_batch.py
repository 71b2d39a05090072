"""Batched persistent rank + null kernel: every matrix's rank and canonical
null basis equal the oracle's, in every elimination mode (GE, M4RI with
k = 4 and k = 8 tables) and tile width (8, 16, 32 words), with and without
hipGraph replay, with the plan the autotuner picks, and equal the
single-matrix path's result for the same matrix (same alone as in the
batch)."""
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
    return mela.default_context()


def _unpack_cols(P, rows, cols):
    bits = np.unpackbits(P.view(np.uint8), axis=1, bitorder="little")
    return bits[:rows, :cols].astype(bool)


def _oracle_null(orc, P, m, n):
    A = _unpack_cols(P, m, n)
    return orc.rank(A, "gf2"), orc.null(A, "gf2")


@pytest.mark.parametrize("flags_extra,modes,groups", [(0, None, None), (4, None, None), (4, "ge", None),
                                                      (0, "mix", None), (4, None, 4), (0, "mix", "mix")])
def test_batch_matches_oracle(mela, ctx, orc, flags_extra, modes, groups):
    items = W.batch_c4(120, seed=7, mmin=8, mmax=384)
    items += [(W.dense_matrix(np.random.default_rng(3), 70, 64), 70, 64, "dense"),
              (np.zeros((5, 2), np.uint64), 5, 100, "zero"),
              (W.dense_matrix(np.random.default_rng(4), 1, 65), 1, 65, "dense")]
    print(f"[batch] {len(items)} matrices flags+{flags_extra} modes={modes} groups={groups}", flush=True)
    dev = [ctx.upload(P, n) for P, m, n, _ in items]
    b = mela.Batch(ctx, dev, mela.BATCH_NULL | flags_extra)
    if modes == "ge":
        b.set_modes([0] * len(items))
    elif modes == "mix":
        b.set_modes([i % 2 for i in range(len(items))])
    if groups == "mix":  # 1..8 workgroups per matrix
        b.set_groups([1 + i % 8 for i in range(len(items))])
    elif groups:
        b.set_groups([groups] * len(items))
    for rep in range(2):  # graph replays must repeat the result
        b.run()
        ranks, nulls = b.results()
        for i, (P, m, n, kind) in enumerate(items):
            r, Z = _oracle_null(orc, P, m, n)
            assert ranks[i] == r, (i, kind, m, n)
            got = _unpack_cols(ctx.download(nulls[i]), n, n - ranks[i])
            assert np.array_equal(got, Z), (i, kind, m, n)
    b.close()
    for M in dev:
        ctx.free(M)


def test_batch_same_alone_as_in_batch(mela, ctx):
    items = W.batch_c4(60, seed=11, mmin=32, mmax=1024)
    dev = [ctx.upload(P, n) for P, m, n, _ in items]
    b = mela.Batch(ctx, dev, mela.BATCH_NULL | mela.BATCH_GRAPH)
    b.autotune()
    b.run()
    ranks, nulls = b.results()
    for i, M in enumerate(dev):
        r = ctx.rank_packed(M)
        Z = ctx.null_packed(M)
        assert r == ranks[i]
        fw = (Z.cols + 63) // 64  # the batch view keeps the input's row pitch
        assert np.array_equal(ctx.download(Z)[:, :fw], ctx.download(nulls[i])[:, :fw]), i
        ctx.free(Z)
    b.close()
    for M in dev:
        ctx.free(M)


@pytest.mark.parametrize("groups", [1, 3, 8])
def test_batch_full_c4_shapes_property(mela, ctx, groups):
    # largest shapes of the uneven batch: 2048-row parity-check matrices;
    # checked by A*Z = 0 on the device and nullity = n - rank, with 1, 3 and
    # 8 workgroups per matrix
    print(f"[batch full shapes] groups={groups}", flush=True)
    rng = np.random.default_rng(5)
    items = [(W.parity_check_matrix(rng, 2048, 4096, 3), 2048, 4096),
             (W.dense_matrix(rng, 2048, 3000), 2048, 3000),
             (W.parity_check_matrix(rng, 1500, 2000, 6), 1500, 2000)]
    dev = [ctx.upload(P, n) for P, m, n in items]
    b = mela.Batch(ctx, dev, mela.BATCH_NULL)
    b.set_groups([groups] * len(items))
    b.run()
    ranks, nulls = b.results()
    for i, (P, m, n) in enumerate(items):
        assert nulls[i].rows == n and nulls[i].cols == n - ranks[i]
        assert ranks[i] == ctx.rank_packed(dev[i])
        Cm = ctx.bitmat(m, nulls[i].cols)
        ctx.matmul_packed(dev[i], nulls[i], Cm)
        assert not ctx.download(Cm).any()
        ctx.free(Cm)
    b.close()
    for M in dev:
        ctx.free(M)


def test_batch_groups_rejects_bad_counts(mela, ctx):
    items = W.batch_c4(4, seed=2, mmin=16, mmax=64)
    dev = [ctx.upload(P, n) for P, m, n, _ in items]
    b = mela.Batch(ctx, dev, mela.BATCH_NULL)
    for bad in (0, 9, -1):
        with pytest.raises(mela.MelaError):
            b.set_groups([bad] * len(items))
    b.close()
    for M in dev:
        ctx.free(M)


def _check_batch(orc, ctx, b, items):
    ranks, nulls = b.results()
    for i, (P, m, n, kind) in enumerate(items):
        r, Z = _oracle_null(orc, P, m, n)
        assert ranks[i] == r, (i, kind, m, n)
        got = _unpack_cols(ctx.download(nulls[i]), n, n - ranks[i])
        assert np.array_equal(got, Z), (i, kind, m, n)


def _tile_items():
    # uneven sparse matrices plus dense ones wide enough for several tiles of
    # every width and tall enough for panels of hundreds of updated rows
    rng = np.random.default_rng(29)
    items = W.batch_c4(60, seed=13, mmin=8, mmax=512)
    for m, n in [(1500, 3000), (700, 1100), (2048, 520), (64, 8192), (300, 65), (1024, 2047)]:
        items.append((W.dense_matrix(rng, m, n), m, n, "dense"))
    items.append((W.parity_check_matrix(rng, 1800, 3000, 3), 1800, 3000, "cw3"))
    return items


@pytest.mark.parametrize("tw,modes,groups", [(8, "k8", None), (8, "mix", "mix"), (16, "k8", None),
                                             (32, "mix", None), (32, "k8", "mix"), (8, "k4", 3)])
def test_batch_tiles_and_table_bits_match_oracle(mela, ctx, orc, tw, modes, groups):
    """Every tile width with k = 8 requested (8-bit tables with 8-word
    tiles, 4-bit tables otherwise), GE / k = 4 / k = 8 mixed per matrix, and
    several workgroups per matrix."""
    items = _tile_items()
    print(f"[batch tiles] tw={tw} modes={modes} groups={groups}", flush=True)
    dev = [ctx.upload(P, n) for P, m, n, _ in items]
    b = mela.Batch(ctx, dev, mela.BATCH_NULL | mela.BATCH_GRAPH)
    b.set_tile(tw)
    md = {"k8": [2] * len(items), "k4": [1] * len(items), "mix": [i % 3 for i in range(len(items))]}[modes]
    b.set_modes(md)
    if groups == "mix":
        b.set_groups([1 + i % 8 for i in range(len(items))])
    elif groups:
        b.set_groups([groups] * len(items))
    assert b.plan()[0] == tw and b.plan()[1] == md
    for rep in range(2):
        b.run()
        _check_batch(orc, ctx, b, items)
    b.close()
    for M in dev:
        ctx.free(M)


def test_batch_autotuned_plan_matches_oracle(mela, ctx, orc):
    """The autotuner's choice of tile width, per-matrix mode (GE, k = 4,
    k = 8) and workgroups per matrix is reported by plan() and gives the
    oracle's results."""
    items = _tile_items()
    dev = [ctx.upload(P, n) for P, m, n, _ in items]
    b = mela.Batch(ctx, dev, mela.BATCH_NULL | mela.BATCH_GRAPH)
    ms = b.autotune()
    tw, md, gr = b.plan()
    print(f"[batch autotune] {ms:.3f} ms tile={tw} modes={np.bincount(md, minlength=3).tolist()} "
          f"groups>1: {sum(g > 1 for g in gr)}", flush=True)
    assert tw in (8, 16, 32) and set(md) <= {0, 1, 2} and all(1 <= g <= 8 for g in gr)
    assert ms > 0
    for rep in range(2):
        b.run()
        _check_batch(orc, ctx, b, items)
    b.close()
    for M in dev:
        ctx.free(M)


def test_batch_plan_setters_reject_bad_values(mela, ctx):
    items = W.batch_c4(4, seed=2, mmin=16, mmax=64)
    dev = [ctx.upload(P, n) for P, m, n, _ in items]
    b = mela.Batch(ctx, dev, mela.BATCH_NULL)
    for bad in (0, 4, 12, 64):
        with pytest.raises(mela.MelaError):
            b.set_tile(bad)
    for bad in (-1, 3):
        with pytest.raises(mela.MelaError):
            b.set_modes([bad] * len(items))
    assert b.plan()[0] == 16 and b.plan()[1] == [1] * len(items)
    b.close()
    for M in dev:
        ctx.free(M)
