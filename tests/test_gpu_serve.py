"""The serving loop on the device, every result checked against the oracle.

A mixed stream (70% rank, 20% null space, 10% products; sides log-uniform
128..8192 for rank, up to 4096 for null bases and products; dense or column
weight 3..6) is submitted while earlier requests are still on the GPU, so
requests join the device-resident loop and the worker streams as they
arrive.  Ranks equal the oracle's (mela_rank_m4ri.c:94-279 restated), null
bases the canonical basis built from the oracle's RREF (Z(free,:) = I,
Z(pivot i,:) = R(i, free), mela_null_gf2.c:88-274), products the oracle's
product (mela_matmul_gf2.c:110-208); a sample of requests run again alone
gives the same results; a tight HBM budget is never exceeded."""
import time

import numpy as np
import pytest

from mex_linear_algebra_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mela():
    import mex_linear_algebra_amd as m
    return m


def _side(rng, lo, hi):
    return int(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def _matrix(rng, m, n):
    if rng.random() < 0.5:
        return W.dense_matrix(rng, m, n)
    return W.parity_check_matrix(rng, m, n, int(rng.integers(3, 7)))


def _bits(P, cols):
    P = np.ascontiguousarray(P, dtype=np.uint64)
    return np.unpackbits(P.view(np.uint8), axis=1, bitorder="little")[:, :cols].astype(bool)


def _check_null(orc, P, m, n, rank, Zp):
    M, piv = orc.rref_packed(P, m, n, "gf2")
    r = len(piv)
    assert rank == r
    f = n - r
    Zb = _bits(Zp, f)
    piv = np.asarray(piv, dtype=np.int64)
    free = np.setdiff1d(np.arange(n), piv)
    assert np.array_equal(Zb[free], np.eye(f, dtype=bool))
    assert np.array_equal(Zb[piv], _bits(M[:r], n)[:, free])


def _requests(rng, count):
    out = []
    for _ in range(count):
        u = rng.random()
        if u < 0.7:
            m, n = _side(rng, 128, 8192), _side(rng, 128, 8192)
            out.append(("rank", _matrix(rng, m, n), m, n, None, None))
        elif u < 0.9:
            m, n = _side(rng, 128, 4096), _side(rng, 128, 4096)
            out.append(("null", _matrix(rng, m, n), m, n, None, None))
        else:
            m, k, n = _side(rng, 128, 3000), _side(rng, 128, 3000), _side(rng, 128, 3000)
            out.append(("matmul", _matrix(rng, m, k), m, k, W.dense_matrix(rng, k, n), n))
    return out


def _submit(mela, srv, ctx, req):
    kind, P, m, n, Bp, nb = req
    dA = ctx.upload(P, n)
    if kind == "matmul":
        dB = ctx.upload(Bp, nb)
        return srv.submit(mela.OP_MATMUL, dA, dB), (dA, dB)
    op = mela.OP_RANK if kind == "rank" else mela.OP_NULL
    # AUTO (blocked elimination for large dense matrices, its fallback for
    # sparse ones) and M4RI (the panel path) alternate
    variant = mela.ELIM_AUTO if (m + n) % 2 == 0 else mela.ELIM_M4RI
    return srv.submit(op, dA, variant=variant), (dA,)


def _result(ctx, req, res):
    kind = req[0]
    if kind == "rank":
        return res.rank
    return res.rank, ctx.download(res.out)


def _check(orc, req, got):
    kind, P, m, n, Bp, nb = req
    if kind == "rank":
        assert got == int(orc.rank_packed(P, m, n, "m4ri")), (m, n)
    elif kind == "null":
        _check_null(orc, P, m, n, got[0], got[1])
    else:
        want = orc.matmul_packed(P, m, n, Bp, nb, "gf2")
        assert np.array_equal(got[1][:, : want.shape[1]], want), (m, n, nb)


def test_mixed_stream_against_oracle(mela, orc):
    rng = np.random.default_rng(2026)
    reqs = _requests(rng, 90)
    ctx = mela.Context(0)
    srv = mela.Server(hbm_budget=64 << 30, workers=4)
    inputs, tickets = [], []
    try:
        # submitted in a stream: later requests arrive while earlier ones run
        for i, q in enumerate(reqs):
            t, dev = _submit(mela, srv, ctx, q)
            tickets.append(t)
            inputs.append(dev)
            if i % 10 == 9:
                time.sleep(0.002)
        results = []
        for q, t in zip(reqs, tickets):
            res = srv.wait(t, 120.0)
            assert res.status == 0
            results.append(_result(ctx, q, res))
            srv.release(t)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["batched_requests"] > 0 and st["jobs"] > 0  # both paths ran
        print(f"[serve] loop {st['batched_requests']} workers {st['jobs']} "
              f"launches {st['loop_launches']} p50 {st['p50_ms']:.3f} p99 {st['p99_ms']:.3f}", flush=True)
        for q, got in zip(reqs, results):
            _check(orc, q, got)
        # the same requests alone, one at a time, give the same results
        for i in range(0, len(reqs), 9):
            t, dev = _submit(mela, srv, ctx, reqs[i])
            res = srv.wait(t, 120.0)
            alone = _result(ctx, reqs[i], res)
            srv.release(t)
            for M in dev:
                ctx.free(M)
            if reqs[i][0] == "rank":
                assert alone == results[i]
            else:
                assert alone[0] == results[i][0] and np.array_equal(alone[1], results[i][1])
    finally:
        srv.close()
        for dev in inputs:
            for M in dev:
                ctx.free(M)
        ctx.close()


def test_tight_budget_is_never_exceeded(mela, orc):
    rng = np.random.default_rng(77)
    budget = 48 << 20
    ctx = mela.Context(0)
    srv = mela.Server(hbm_budget=budget, workers=2)
    inputs = []
    try:
        reqs = []
        for i in range(60):
            m, n = _side(rng, 128, 1500), _side(rng, 128, 2500)
            reqs.append(("null" if i % 4 == 0 else "rank", _matrix(rng, m, n), m, n, None, None))
        reqs.append(("rank", W.dense_matrix(rng, 3000, 6000), 3000, 6000, None, None))  # a worker job
        tickets = []
        for q in reqs:
            t, dev = _submit(mela, srv, ctx, q)
            tickets.append(t)
            inputs.append(dev)
        for q, t in zip(reqs, tickets):
            res = srv.wait(t, 120.0)
            got = _result(ctx, q, res)
            srv.release(t)
            _check(orc, q, got)
        st = srv.stats()
        assert 0 < st["hbm_peak"] <= budget, st
        # one request that could never fit (working and pivot rows of 64 MiB
        # plus a 32 MiB basis) is refused at submission
        big = ctx.bitmat(16384, 16384)
        with pytest.raises(mela.MelaError) as ei:
            srv.submit(mela.OP_NULL, big)
        assert ei.value.status == "ADMISSION"
        ctx.free(big)
    finally:
        srv.close()
        for dev in inputs:
            for M in dev:
                ctx.free(M)
        ctx.close()


def test_loop_restarts_after_idle(mela, orc):
    rng = np.random.default_rng(5)
    ctx = mela.Context(0)
    srv = mela.Server(hbm_budget=1 << 30, workers=1)
    inputs = []
    try:
        for round_ in range(3):
            P = W.dense_matrix(rng, 300, 500)
            d = ctx.upload(P, 500)
            inputs.append(d)
            t = srv.submit(mela.OP_RANK, d)
            assert srv.wait(t, 30.0).rank == int(orc.rank_packed(P, 300, 500, "m4ri"))
            srv.release(t)
            time.sleep(0.05)  # longer than the loop's idle limit and lifetime
        assert srv.stats()["loop_launches"] >= 2
    finally:
        srv.close()
        for d in inputs:
            ctx.free(d)
        ctx.close()


def test_server_on_missing_device_leaves_no_error(mela):
    """A server on a device index that does not exist fails its requests;
    closing it leaves no HIP error behind for the caller's next call."""
    s = mela.Server(device=99, hbm_budget=1 << 30, workers=1)
    s.close()
    ctx = mela.Context(0)
    M = ctx.bitmat(64, 64)  # would raise "invalid device ordinal" from a stale error
    ctx.free(M)
