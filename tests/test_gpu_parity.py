"""Parity of the HIP path (through the C ABI) with the oracle, bit-exact.

Small and medium shapes are compared element for element with the CPU
restatement of the reference; large shapes are checked through
size-independent properties (Freivalds projections for products, rank of
L*R constructions, A*Z = 0 and Z(free,:) = I for null bases)."""
import json
import os

import numpy as np
import pytest

from tests import _brute as bf

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "kats.json")


@pytest.fixture(scope="module")
def mela():
    import mex_linear_algebra_amd as m
    m.default_context()
    return m


@pytest.fixture(scope="module")
def ctx(mela):
    return mela.default_context()


def _kat_matrix(case, key="A"):
    if "zeros" in case and key == "A":
        return np.zeros(case["zeros"], dtype=bool)
    if "eye" in case and key == "A":
        return np.eye(case["eye"], dtype=bool)
    if "zerosA" in case and key == "A":
        return np.zeros(case["zerosA"], dtype=bool)
    M = np.array(case[key])
    return M.astype(np.float64) if case["kind"] == "double" else M.astype(bool)


def test_kats(mela):
    rng = np.random.default_rng(42)
    with open(GOLD) as f:
        kats = json.load(f)
    for case in kats:
        A = _kat_matrix(case)
        if case["op"] == "rank":
            assert mela.mela_rank_gf2(A) == case["rank"], case["src"]
            assert mela.mela_rank_m4ri(A) == case["rank"], case["src"]
        elif case["op"] == "null":
            for fn in (mela.mela_null_gf2, mela.mela_null_m4ri):
                Z = fn(A)
                assert Z.dtype == bool
                assert np.array_equal(Z, np.array(case["Z"], dtype=bool)), case["src"]
        elif case["op"] == "rref":
            R, jb = mela.mela_rref_gf2(A)
            assert np.array_equal(R, np.array(case["R"], dtype=bool)), case["src"]
            assert jb.ravel().tolist() == case["jb"], case["src"]
        else:
            B = _kat_matrix(case, "B") if "B" in case else rng.integers(0, 2, case["Bshape"]).astype(bool)
            for fn in (mela.mela_matmul_gf2, mela.mela_matmul_m4ri):
                assert np.array_equal(fn(A, B), np.array(case["C"], dtype=bool)), case["src"]


MM_SHAPES = [(37, 23, 31), (1, 1, 1), (5, 7, 3), (64, 64, 64), (100, 50, 30), (256, 128, 64),
             (300, 257, 190), (1025, 129, 2049), (70, 1000, 65), (2048, 2048, 2048),
             (3000, 70, 4100), (0, 5, 3), (5, 0, 3), (5, 7, 0)]


@pytest.mark.parametrize("shape", MM_SHAPES)
def test_matmul_parity(mela, orc, shape):
    m, k, n = shape
    print(f"[matmul parity] shape={shape}", flush=True)
    rng = np.random.default_rng(m * 7 + k * 3 + n)
    A = rng.integers(0, 2, (m, k)).astype(bool)
    B = rng.integers(0, 2, (k, n)).astype(bool)
    want = orc.matmul(A, B, "m4ri")
    assert np.array_equal(mela.mela_matmul_gf2(A, B), want)
    assert np.array_equal(mela.mela_matmul_m4ri(A, B), want)


def test_matmul_double_parity(mela, orc):
    rng = np.random.default_rng(5)
    A = rng.integers(-9, 10, (77, 91)).astype(np.float64) + rng.choice([0.0, 0.25, 0.75], (77, 91))
    B = rng.integers(-9, 10, (91, 66)).astype(np.float64)
    A[0, 0], A[1, 1], A[2, 2], A[3, 3] = np.nan, np.inf, -np.inf, 3e10
    B[0, 0], B[1, 0] = 2147483648.0, -2147483649.0
    want = orc.matmul(A, B, "gf2")
    assert np.array_equal(mela.mela_matmul_gf2(A, B), want)
    assert np.array_equal(mela.mela_matmul_m4ri(A, B.astype(bool)), orc.matmul(A, B.astype(bool), "gf2"))


@pytest.mark.parametrize("algo", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("shape", [(64, 64, 64), (129, 4097, 300), (2100, 520, 4160), (1000, 8, 9),
                                   (33, 31, 95), (257, 2000, 129)])
def test_matmul_packed_algos(mela, ctx, orc, algo, shape):
    m, k, n = shape
    print(f"[matmul packed] algo={algo} shape={shape}", flush=True)
    Ap = orc.gen_bernoulli_packed(11, m, k)
    Bp = orc.gen_bernoulli_packed(12, k, n)
    dA, dB = ctx.upload(Ap, k), ctx.upload(Bp, n)
    dC = ctx.bitmat(m, n)
    ctx.matmul_packed(dA, dB, dC, algo)
    got = ctx.download(dC)
    want = orc.matmul_packed(Ap, m, k, Bp, n, "m4ri")
    for M in (dA, dB, dC):
        ctx.free(M)
    assert np.array_equal(got[:, : want.shape[1]], want)


@pytest.mark.parametrize("algo", [0, 3, 4])
@pytest.mark.parametrize("shape", [(4100, 1000, 4000), (4096, 4160, 4293), (4352, 64, 4097),
                                   (4100, 1, 4100), (4096, 129, 4352), (4100, 2500, 4200)])
def test_matmul_mfma_large_tiles_exact(mela, ctx, orc, algo, shape):
    # grids of >= 256 tiles of 256 x 256 take the stage-tiled kernels; ragged
    # m, k and n exercise the zero padding and the last-word mask, odd row
    # lengths the 8-byte loads of the operand re-layouts, and 2500 x 4200 (40
    # and 66 words per row) their 16-byte loads with a partial last group
    m, k, n = shape
    print(f"[matmul mfma large tiles] algo={algo} shape={shape}", flush=True)
    Ap = orc.gen_bernoulli_packed(21, m, k)
    Bp = orc.gen_bernoulli_packed(22, k, n)
    dA, dB = ctx.upload(Ap, k), ctx.upload(Bp, n)
    dC = ctx.bitmat(m, n)
    ctx.matmul_packed(dA, dB, dC, algo)
    got = ctx.download(dC)
    want = orc.matmul_packed(Ap, m, k, Bp, n, "m4ri")
    for M in (dA, dB, dC):
        ctx.free(M)
    assert np.array_equal(got[:, : want.shape[1]], want)


@pytest.mark.parametrize("algo", [0, 5])
@pytest.mark.parametrize("m,k,n,w", [(6000, 8000, 5000, 3), (4096, 4096, 9000, 6), (300, 20000, 700, 4)])
def test_matmul_sparse_operand(mela, ctx, orc, algo, m, k, n, w):
    """A with exactly w ones per column (parity-check-like), B dense: AUTO
    (which counts the bits of A and takes the row-XOR product when A is
    sparse) and the row-XOR product itself equal the oracle."""
    from mex_linear_algebra_amd import workloads as W
    print(f"[matmul sparse] algo={algo} {m}x{k}x{n} w={w}", flush=True)
    rng = np.random.default_rng(m + k + n + w)
    Ap = W.parity_check_matrix(rng, m, k, w)
    Bp = orc.gen_bernoulli_packed(91, k, n)
    dA, dB, dC = ctx.upload(Ap, k), ctx.upload(Bp, n), ctx.bitmat(m, n)
    ctx.matmul_packed(dA, dB, dC, algo)
    got = ctx.download(dC)
    want = orc.matmul_packed(Ap, m, k, Bp, n, "gf2")
    for M in (dA, dB, dC):
        ctx.free(M)
    assert np.array_equal(got[:, : want.shape[1]], want)


def test_device_generator_matches_oracle(ctx, orc):
    M = ctx.bitmat(300, 1000)
    ctx.gen_bernoulli(77, M)
    got = ctx.download(M)
    ctx.free(M)
    assert np.array_equal(got, orc.gen_bernoulli_packed(77, 300, 1000))


@pytest.mark.parametrize("algo", [1, 2, 3, 4])
def test_matmul_large_freivalds(ctx, algo):
    # 8192 x 8192 x 8192 product checked by random projections
    m = k = n = 8192
    print(f"[matmul freivalds] algo={algo} n={n}", flush=True)
    dA, dB, dC = ctx.bitmat(m, k), ctx.bitmat(k, n), ctx.bitmat(m, n)
    ctx.gen_bernoulli(1, dA)
    ctx.gen_bernoulli(2, dB)
    ctx.matmul_packed(dA, dB, dC, algo)
    Ap, Bp, Cp = ctx.download(dA), ctx.download(dB), ctx.download(dC)
    for M in (dA, dB, dC):
        ctx.free(M)
    assert bf.freivalds(Ap, Bp, Cp, k, n, trials=8, seed=algo)


RANK_SHAPES = [(8, 8), (17, 29), (64, 64), (65, 63), (1, 200), (200, 1), (300, 700),
               (700, 300), (1000, 1000), (2048, 2048), (130, 4000)]


@pytest.mark.parametrize("shape", RANK_SHAPES)
def test_rank_parity(mela, orc, shape):
    m, n = shape
    print(f"[rank parity] shape={shape}", flush=True)
    rng = np.random.default_rng(m + 31 * n)
    A = rng.integers(0, 2, (m, n)).astype(bool)
    want = orc.rank(A, "gf2")
    assert mela.mela_rank_gf2(A) == want
    assert mela.mela_rank_m4ri(A) == want


@pytest.mark.parametrize("m,n,r", [(300, 400, 100), (1000, 1200, 999), (700, 2000, 64),
                                   (2000, 1500, 1000), (128, 128, 1), (500, 500, 0)])
def test_rank_low_rank(mela, orc, m, n, r):
    print(f"[rank low-rank] m={m} n={n} r={r}", flush=True)
    rng = np.random.default_rng(m * n + r)
    if r == 0:
        H = np.zeros((m, n), dtype=bool)
    else:
        H = orc.matmul(rng.integers(0, 2, (m, r)).astype(bool),
                       rng.integers(0, 2, (r, n)).astype(bool), "gf2")
    want = orc.rank(H, "m4ri")
    assert mela.mela_rank_gf2(H) == want
    assert mela.mela_rank_m4ri(H) == want


def test_rank_structured(mela):
    for A, want in [(np.eye(300, dtype=bool), 300), (np.zeros((0, 5), bool), 0),
                    (np.zeros((5, 0), bool), 0), (np.ones((70, 90), bool), 1),
                    (np.tril(np.ones((200, 200), bool)), 200)]:
        assert mela.mela_rank_gf2(A) == want
        assert mela.mela_rank_m4ri(A) == want


NULL_SHAPES = [(2, 4), (16, 32), (31, 47), (50, 100), (100, 50), (64, 130), (300, 450),
               (1000, 1500), (5, 1), (1, 5), (0, 7), (7, 0)]


@pytest.mark.parametrize("shape", NULL_SHAPES)
def test_null_parity(mela, orc, shape):
    m, n = shape
    print(f"[null parity] shape={shape}", flush=True)
    rng = np.random.default_rng(m * 13 + n)
    A = rng.integers(0, 2, (m, n)).astype(bool)
    want = orc.null(A, "gf2")
    for fn in (mela.mela_null_gf2, mela.mela_null_m4ri):
        Z = fn(A)
        assert Z.dtype == bool and Z.shape == want.shape
        assert np.array_equal(Z, want), fn.__name__


@pytest.mark.parametrize("m,n,r", [(200, 300, 50), (900, 1400, 700), (1100, 1100, 1050),
                                   (2048, 3072, 1500)])
def test_null_low_rank(mela, orc, m, n, r):
    print(f"[null low-rank] m={m} n={n} r={r}", flush=True)
    rng = np.random.default_rng(r)
    H = orc.matmul(rng.integers(0, 2, (m, r)).astype(bool),
                   rng.integers(0, 2, (r, n)).astype(bool), "gf2")
    want = orc.null(H, "m4ri")
    for fn in (mela.mela_null_gf2, mela.mela_null_m4ri):
        assert np.array_equal(fn(H), want), fn.__name__


def _blocks(rng, orc, m, parts):
    """Columns side by side: ("lr", n, r) a rank-r product, ("zero", n), ("rand", n)."""
    cols = []
    for p in parts:
        if p[0] == "lr":
            cols.append(orc.matmul(rng.integers(0, 2, (m, p[2])).astype(bool),
                                   rng.integers(0, 2, (p[2], p[1])).astype(bool), "gf2"))
        elif p[0] == "zero":
            cols.append(np.zeros((m, p[1]), dtype=bool))
        else:
            cols.append(rng.integers(0, 2, (m, p[1])).astype(bool))
    return np.concatenate(cols, axis=1)


# rank stalls (a low-rank block, then zero panels) and resumes, leading zero
# panels, and a rank exhausted exactly at a panel boundary followed by zeros:
# the elimination skips panels once every active row is zero, and must not
# do so while later columns still hold pivots
@pytest.mark.parametrize("parts", [
    [("lr", 256, 64), ("zero", 320), ("rand", 100)],
    [("zero", 200), ("rand", 150)],
    [("rand", 128), ("zero", 400)],
    [("lr", 300, 37), ("zero", 70), ("lr", 200, 20), ("zero", 130)],
])
def test_rank_null_block_structured(mela, orc, parts):
    print(f"[block structured] {parts}", flush=True)
    H = _blocks(np.random.default_rng(len(parts) * 7 + parts[0][1]), orc, 500, parts)
    want_r = orc.rank(H, "gf2")
    want_z = orc.null(H, "gf2")
    assert mela.mela_rank_gf2(H) == want_r
    assert mela.mela_rank_m4ri(H) == want_r
    for fn in (mela.mela_null_gf2, mela.mela_null_m4ri):
        assert np.array_equal(fn(H), want_z), fn.__name__


def test_null_double_input(mela, orc):
    rng = np.random.default_rng(9)
    H = rng.integers(-5, 6, (40, 70)).astype(np.float64)
    assert np.array_equal(mela.mela_null_gf2(H), orc.null(H, "gf2"))
    assert np.array_equal(mela.mela_null_m4ri(H), orc.null(H, "gf2"))


@pytest.mark.parametrize("variant", ["gf2", "m4ri"])
def test_captured_panels_continue_for_a_higher_rank(mela, orc, variant):
    """Repeated eliminations of one shape replay a captured panel sequence
    that stops where the earlier run found the elimination complete.  A later
    matrix of the same shape that needs more panels continues past it, and
    the sequence after that is captured whole; every result matches the
    oracle (rank, and the null basis through the RREF path)."""
    rng = np.random.default_rng(23)
    m, n = 600, 1300
    low = orc.matmul(rng.integers(0, 2, (m, 100)).astype(bool),
                     rng.integers(0, 2, (100, n)).astype(bool), "gf2")
    full = rng.integers(0, 2, (m, n)).astype(bool)
    rank_fn = getattr(mela, f"mela_rank_{variant}")
    null_fn = getattr(mela, f"mela_null_{variant}")
    want_rank = {id(low): orc.rank(low, "gf2"), id(full): orc.rank(full, "gf2")}
    want_null = {id(low): orc.null(low, "gf2"), id(full): orc.null(full, "gf2")}
    assert want_rank[id(low)] == 100 and want_rank[id(full)] == 600
    for A in (low, low, low, full, full, full, low, full):
        assert rank_fn(A) == want_rank[id(A)]
    for A in (low, low, low, full, full, full, low, full):
        assert np.array_equal(null_fn(A), want_null[id(A)])


def test_host_arrays_in_several_staging_chunks(mela, orc):
    """Host arrays above 8 MiB are packed and copied in column chunks (a
    1500-row double column is 12000 bytes, so these go in 3-4 chunks):
    every entry point still matches the oracle bit for bit, double specials
    included, and the results come back through the packed transpose."""
    rng = np.random.default_rng(17)
    H = rng.integers(-7, 8, (1500, 2100)).astype(np.float64)
    H[::97, ::89] = np.nan
    H[5::101, 3::83] = 2147483649.0
    H[7::103, 1::79] = -0.75
    B = rng.integers(0, 2, (2100, 700)).astype(bool)
    assert mela.mela_rank_m4ri(H) == orc.rank(H, "m4ri")
    assert mela.mela_rank_gf2(H) == orc.rank(H, "gf2")
    assert np.array_equal(mela.mela_null_gf2(H), orc.null(H, "gf2"))
    assert np.array_equal(mela.mela_matmul_gf2(H, B), orc.matmul(H, B, "gf2"))
    # logical: 4096-byte columns, 8 MiB per chunk -> two chunks of 2048 columns
    L = np.asfortranarray(rng.integers(0, 2, (4096, 2600)).astype(bool))
    B3 = rng.integers(0, 2, (2600, 3)).astype(bool)
    assert np.array_equal(mela.mela_matmul_m4ri(L, B3), orc.matmul(L, B3, "gf2"))
    assert mela.mela_rank_gf2(L) == orc.rank(L, "gf2")
    R, jb = mela.mela_rref_gf2(L[:, :2300])
    M, piv = orc.rref_packed(orc.pack(L[:, :2300]), 4096, 2300, "gf2")
    assert np.array_equal(R, orc.unpack(M, 4096, 2300))
    assert jb.ravel().tolist() == [float(c + 1) for c in piv]


def test_rref_packed_matches_oracle(ctx, orc):
    m, n = 500, 900
    P = orc.gen_bernoulli_packed(3, m, n)
    dA = ctx.upload(P, n)
    R, piv = ctx.rref_packed(dA)
    got = ctx.download(R)
    M, opiv = orc.rref_packed(P, m, n, "gf2")
    ctx.free(dA)
    ctx.free(R)
    assert list(piv) == list(opiv)
    assert np.array_equal(got[: len(piv), : M.shape[1]], M[: len(piv)])


def test_null_packed_large_property(ctx, orc):
    # 4096 x 6144 Bernoulli: A * Z = 0 by Freivalds, Z(free,:) = I, nullity
    m, n = 4096, 6144
    print(f"[null packed large] m={m} n={n}", flush=True)
    dA = ctx.bitmat(m, n)
    ctx.gen_bernoulli(21, dA)
    r = ctx.rank_packed(dA)
    Z = ctx.null_packed(dA)
    assert Z.rows == n and Z.cols == n - r
    dC = ctx.bitmat(m, Z.cols)
    ctx.matmul_packed(dA, Z, dC)
    Cp = ctx.download(dC)
    Zp = ctx.download(Z)
    for M in (dA, Z, dC):
        ctx.free(M)
    assert not Cp.any()
    assert r == m  # a random 4096 x 6144 matrix has full row rank w.h.p.
    assert Zp.shape[0] == n


def test_error_messages(mela):
    A = np.zeros((3, 4), dtype=bool)
    with pytest.raises(mela.MelaError, match="Inner dimensions must agree."):
        mela.mela_matmul_gf2(A, A)
    with pytest.raises(mela.MelaError, match="Inputs must be logical or double matrices."):
        mela.mela_matmul_m4ri(A.astype(np.int32), A.T.copy())
    with pytest.raises(mela.MelaError) as ei:
        mela.mela_rank_gf2(A.astype(np.float32))
    assert str(ei.value) == "Input must be a logical or double matrix."
    assert ei.value.identifier == "bitrank_mex:inputNotLogical"
    with pytest.raises(mela.MelaError, match="Input must be double or logical."):
        mela.mela_rank_m4ri(A.astype(np.int8))
    with pytest.raises(mela.MelaError, match="H must be double or logical."):
        mela.mela_null_gf2(A.astype(np.uint16))
    with pytest.raises(mela.MelaError, match="Need two input matrices A,B."):
        mela.mela_matmul_gf2(A, A, A)


_EXPANDED_CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import mex_linear_algebra_amd as mela
from oracle import oracle as orc
orc.load()
ctx = mela.Context(0)
for m, k, n in [(4100, 1000, 4000), (4096, 4160, 4293), (4352, 2500, 4200)]:
    print(f"[matmul fp4 expanded] {m}x{k}x{n}", flush=True)
    Ap = orc.gen_bernoulli_packed(31, m, k)
    Bp = orc.gen_bernoulli_packed(32, k, n)
    dA, dB = ctx.upload(Ap, k), ctx.upload(Bp, n)
    dC = ctx.bitmat(m, n)
    ctx.matmul_packed(dA, dB, dC, 4)
    got = ctx.download(dC)
    want = orc.matmul_packed(Ap, m, k, Bp, n, "m4ri")
    for M in (dA, dB, dC):
        ctx.free(M)
    assert np.array_equal(got[:, : want.shape[1]], want), (m, k, n)
print("expanded ok", flush=True)
"""


def test_matmul_fp4_expanded_kernels_exact():
    """The expanded-operand fp4 kernels (operands stored as e2m1 nibbles, LDS
    ring), which grids of >= 256 tiles take when k >= 2^23 (the direct
    kernel's accumulators start at 2^23), forced with MELA_MFMA_EXPANDED=1 in
    a child process at sizes the oracle checks in seconds."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, MELA_MFMA_EXPANDED="1")
    out = subprocess.run([sys.executable, "-c", _EXPANDED_CHILD, root], capture_output=True, text=True,
                         timeout=240, env=env)
    print(out.stdout[-2000:], out.stderr[-2000:], flush=True)
    assert out.returncode == 0 and "expanded ok" in out.stdout
