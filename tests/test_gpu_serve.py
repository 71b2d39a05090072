"""Serving loop on the device: a mix of rank, null-space and matmul requests
submitted together (so they are admitted, micro-batched and run concurrently
on several streams) gives, request by request, exactly the result of the same
request run alone through the single-matrix entry points (tolerance 0)."""
import numpy as np
import pytest

from mex_linear_algebra_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mela():
    import mex_linear_algebra_amd as m
    return m


def test_mix_equals_alone(mela):
    ctx = mela.default_context()
    rng = np.random.default_rng(123)
    pool = []
    for i in range(48):
        m = int(np.exp(rng.uniform(np.log(64), np.log(1500))))
        n = int(np.exp(rng.uniform(np.log(64), np.log(1500))))
        P = W.dense_matrix(rng, m, n) if i % 2 else W.parity_check_matrix(rng, m, n, 3 + i % 4)
        pool.append(ctx.upload(P, n))
    big = ctx.bitmat(4500, 5000)  # beyond the batch kernel: single-matrix path
    ctx.gen_bernoulli(9, big)
    pool.append(big)
    mm = []
    for i in range(6):
        m, k, n = (int(x) for x in rng.integers(64, 2000, 3))
        A, B = ctx.bitmat(m, k), ctx.bitmat(k, n)
        ctx.gen_bernoulli(100 + i, A)
        ctx.gen_bernoulli(200 + i, B)
        mm.append((A, B))
    ctx.sync()
    srv = mela.Server(hbm_budget=64 << 30, workers=4, max_batch=16)
    tickets = []
    for i, M in enumerate(pool):
        op = mela.OP_NULL if i % 3 == 0 else mela.OP_RANK
        tickets.append((op, M, None, srv.submit(op, M, variant=i % 2)))
        if i % 8 == 0 and mm:
            A, B = mm.pop()
            tickets.append((mela.OP_MATMUL, A, B, srv.submit(mela.OP_MATMUL, A, B)))
    for A, B in mm:
        tickets.append((mela.OP_MATMUL, A, B, srv.submit(mela.OP_MATMUL, A, B)))
    for op, A, B, t in tickets:
        res = srv.wait(t, 120.0)
        assert res.status == 0
        if op == mela.OP_RANK:
            assert res.rank == ctx.rank_packed(A)
        elif op == mela.OP_NULL:
            Z = ctx.null_packed(A)
            assert res.out.cols == Z.cols and res.rank == A.cols - Z.cols
            assert np.array_equal(ctx.download(res.out), ctx.download(Z))
            ctx.free(Z)
        else:
            Cm = ctx.bitmat(A.rows, B.cols)
            ctx.matmul_packed(A, B, Cm)
            assert np.array_equal(ctx.download(res.out), ctx.download(Cm))
            ctx.free(Cm)
        srv.release(t)
    st = srv.stats()
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert st["batched_requests"] > 0
    srv.close()
