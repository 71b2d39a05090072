"""GPU tests of the RREF entry point with pivot columns ([R, jb] as in
gf2null/gf2rref.m:49-73) and of sparse (compressed sparse column) ingestion,
against the oracle and the independent row reduction in tests/_brute.py."""
import numpy as np
import pytest

from tests import _brute as bf

pytestmark = pytest.mark.gpu

sp = pytest.importorskip("scipy.sparse")


@pytest.fixture(scope="module")
def mela():
    import mex_linear_algebra_amd as m
    m.default_context()
    return m


@pytest.fixture(scope="module")
def ctx(mela):
    return mela.default_context()


def _low_rank(rng, m, n, r):
    L = rng.integers(0, 2, (m, r)).astype(np.int64)
    R = rng.integers(0, 2, (r, n)).astype(np.int64)
    return ((L @ R) % 2).astype(bool)


RREF_SHAPES = [(1, 1), (3, 5), (17, 29), (64, 64), (65, 130), (300, 700), (700, 300),
               (1000, 1000), (2048, 3000)]


@pytest.mark.parametrize("shape", RREF_SHAPES)
def test_rref_gf2_matches_oracle(mela, orc, shape):
    m, n = shape
    print(f"[rref] shape={shape}", flush=True)
    rng = np.random.default_rng(m * 31 + n)
    A = rng.integers(0, 2, (m, n)).astype(bool)
    R, jb = mela.mela_rref_gf2(A)
    M, piv = orc.rref_packed(orc.pack(A), m, n, "gf2")
    assert R.shape == (m, n) and R.dtype == bool
    assert np.array_equal(R, orc.unpack(M, m, n))
    assert jb.shape == (1, len(piv))
    assert jb.ravel().tolist() == [float(c + 1) for c in piv]


@pytest.mark.parametrize("m,n,r", [(40, 70, 9), (500, 300, 120), (1500, 2500, 700)])
def test_rref_gf2_low_rank_vs_brute(mela, m, n, r):
    print(f"[rref low rank] m={m} n={n} r={r}", flush=True)
    A = _low_rank(np.random.default_rng(r), m, n, r)
    R, jb = mela.mela_rref_gf2(A)
    want_R, want_piv = bf.rref(A)
    assert np.array_equal(R, want_R.astype(bool))
    assert jb.ravel().astype(int).tolist() == [c + 1 for c in want_piv]
    # R(1:r, jb) is the identity, the rows below are zero
    k = len(want_piv)
    assert np.array_equal(R[:k][:, [c for c in want_piv]], np.eye(k, dtype=bool))
    assert not R[k:].any()


def test_rref_gf2_double_parity_and_empty(mela, orc):
    A = np.array([[2.0, 3.0, -1.0], [4.5, 7.0, 2.0], [np.nan, 1.0, 5.0]])
    R, jb = mela.mela_rref_gf2(A)
    Ab = np.array([[orc.parity_double(x) for x in row] for row in A], dtype=bool)
    want_R, want_piv = bf.rref(Ab)
    assert np.array_equal(R, want_R.astype(bool))
    assert jb.ravel().astype(int).tolist() == [c + 1 for c in want_piv]
    for shape in ((0, 4), (4, 0), (3, 3)):
        R, jb = mela.mela_rref_gf2(np.zeros(shape, dtype=bool))
        assert R.shape == shape and not R.any() and jb.size == 0


def _dense_bits(S):
    """Bits of a scipy sparse matrix by the dense rule: stored doubles by
    ((int)x) & 1 (computed with numpy truncation for the finite in-range
    values used here), bools as is."""
    D = S.toarray()
    if D.dtype == np.bool_:
        return D
    return (np.trunc(D).astype(np.int64) & 1).astype(bool)


@pytest.mark.parametrize("m,n,density", [(1, 1, 1.0), (100, 77, 0.05), (3000, 5000, 0.002),
                                         (640, 2100, 0.3), (4096, 300, 0.0)])
def test_pack_csc_matches_dense(ctx, orc, m, n, density):
    print(f"[pack csc] m={m} n={n} density={density}", flush=True)
    rng = np.random.default_rng(m + n)
    S = sp.random(m, n, density=density, format="csc", random_state=rng,
                  data_rvs=lambda k: rng.integers(-5, 6, k).astype(np.float64))
    M = ctx.upload_sparse(S)
    got = ctx.download(M)
    ctx.free(M)
    want = orc.pack(_dense_bits(S))
    assert np.array_equal(got[:, : want.shape[1]], want)


def test_pack_csc_value_classes(ctx, orc):
    # pattern (every stored entry is 1), logical and double values in one shape
    rng = np.random.default_rng(5)
    S = sp.random(300, 200, density=0.1, format="csc", random_state=rng)
    pattern = (S != 0).astype(bool)
    M = ctx.pack_csc(S.indptr, S.indices, None, 300, 200)
    assert np.array_equal(ctx.download(M)[:, :4], orc.pack(pattern.toarray())[:, :4])
    ctx.free(M)
    B = S.copy()
    B.data = (rng.integers(0, 2, B.nnz)).astype(bool)
    M = ctx.pack_csc(B.indptr, B.indices, B.data, 300, 200)
    assert np.array_equal(ctx.download(M)[:, :4], orc.pack(B.toarray().astype(bool))[:, :4])
    ctx.free(M)
    D = S.copy()
    D.data = np.array([2.0, 3.0, -3.0, 2.7, 3.5, -1.5, 1e9 + 1, 0.5] * (D.nnz // 8 + 1))[: D.nnz]
    M = ctx.pack_csc(D.indptr, D.indices, D.data, 300, 200)
    want = np.zeros((300, 200), dtype=bool)
    Dc = D.tocoo()
    for i, j, v in zip(Dc.row, Dc.col, Dc.data):
        want[i, j] = orc.parity_double(float(v)) == 1
    assert np.array_equal(ctx.download(M)[:, :4], orc.pack(want)[:, :4])
    ctx.free(M)


def test_pack_csc_rejects_bad_input(mela, ctx):
    jc = np.array([0, 2, 3], dtype=np.uint64)
    ir = np.array([0, 5, 1], dtype=np.uint64)  # row 5 >= m = 4: flagged by the kernel, not written
    with pytest.raises(mela.MelaError):
        ctx.pack_csc(jc, ir, None, 4, 2)
    with pytest.raises(mela.MelaError):
        ctx.pack_csc(np.array([0, 2, 1], dtype=np.uint64), np.array([0, 1], dtype=np.uint64), None, 4, 2)
    with pytest.raises(mela.MelaError):
        ctx.pack_csc(np.array([1, 2, 3], dtype=np.uint64), ir, None, 8, 2)


def test_sparse_parity_check_rank_and_null(ctx, orc):
    # a parity-check-like matrix (column weight 3) straight from CSC to rank
    # and the canonical null basis on the device
    rng = np.random.default_rng(11)
    m, n, w = 1500, 2400, 3
    rows = np.concatenate([rng.choice(m, w, replace=False) for _ in range(n)])
    cols = np.repeat(np.arange(n), w)
    S = sp.csc_matrix((np.ones(n * w, dtype=np.float64), (rows, cols)), shape=(m, n))
    A = ctx.upload_sparse(S)
    dense = S.toarray().astype(bool)
    r = ctx.rank_packed(A)
    assert r == orc.rank(dense, "gf2")
    Z = ctx.null_packed(A)
    Zb = orc.unpack(ctx.download(Z), n, n - r)
    assert Zb.shape == (n, n - r)
    prod = (dense.astype(np.int64) @ Zb.astype(np.int64)) % 2
    assert not prod.any()
    ctx.free(Z)
    ctx.free(A)
