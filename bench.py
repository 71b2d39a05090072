#!/usr/bin/env python3
"""Benchmarks of the GF(2) hot path on one MI355X.

Default (the headline line): GF(2) matrix product C = A*B at 32768^3.  One
step = one packed product of two 32768 x 32768 Bernoulli(1/2) bit matrices
(128 MiB each, generated on the device and resident in HBM before the timed
region) through the C ABI (mela_matmul_packed).  Work per step is W = 2*m*n*k
bit operations (one AND and one XOR per (i,j,k) triple), reported as
Tbit-op/s.  Prints one JSON line, which also carries (unless --headline-only)
the rest of the metric: `batched` (4096 uneven rank + null matrices in one
persistent launch against one null_packed call per matrix), `batched_matmul`
(2048 uneven products in one persistent launch against one matmul_packed call
per product), `serve` (p50/p99 of the request server at two offered loads)
and `c1` (the six entry points end to end on 2048 x 2048 logical arrays, GPU
and CPU oracle side by side).

Other workloads (--workload): elim (16384-scale rank + null space, GE and
M4RI), batch, batch_matmul, serve (a sweep of offered loads), c1, copy (a
plain copy kernel of known size, used to scale HBM byte counters).  Each
prints its own JSON line(s).

CPU figures run with OMP_DYNAMIC=FALSE and, unless the environment already
sets it, OMP_NUM_THREADS = the CPUs this process may run on (the reference
benchmark's setting, benchmark_mex.m:163-164); the CPU model, CPU counts and
threads used are recorded with them.

--gpus N>1 is refused: this build measures one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# before any OpenMP runtime is loaded (it reads these once, at load)
os.environ["OMP_DYNAMIC"] = "FALSE"
os.environ.setdefault("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0))))

# v_bitop3_b32 retires one AND and one XOR per bit lane: 256 CU x 64 lanes x
# 2.4 GHz x 32 bits x 2 = 2516 Tbit-op/s of W on the bitwise VALU path.
VALU_BITOP3_PEAK = 2516.0
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dominant product kernel, measured with two
# rocprofv3 --pmc passes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a
# 1 GiB copy kernel) on this workload; profiles/k_mfma_direct_fp4_32768.md.
MEASURED_TRAFFIC = {("mfma_fp4", 32768): 6.611e9}


class HipEvents:
    """hipEvent timing on a given stream (the library's own stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def reference_timing(fn):
    """The reference benchmark's timing policy (benchmark_mex.m:671-707): one
    warm-up call; double the iteration count until a round takes >= 10 ms;
    then rounds of ceil(0.10 s / per-call) iterations, 5 rounds (2 when a call
    takes >= 0.5 s); the median per-call time in seconds, and the rounds."""
    fn()
    iters = 1
    while True:
        t0 = time.perf_counter()
        for _ in range(iters):
            fn()
        dt = time.perf_counter() - t0
        if dt >= 0.010:
            break
        iters *= 2
    per = dt / iters
    iters = max(1, int(-(-0.10 // per)))
    rounds = 2 if per >= 0.5 else 5
    times = []
    for _ in range(rounds):
        t0 = time.perf_counter()
        for _ in range(iters):
            fn()
        times.append((time.perf_counter() - t0) / iters)
    times.sort()
    return times[len(times) // 2], rounds, iters


def host_metadata(orc):
    """CPU model, CPU counts and the OpenMP settings the CPU figures ran with."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "cpus_allowed": len(os.sched_getaffinity(0)),
            "threads": orc.num_threads(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_DYNAMIC": os.environ.get("OMP_DYNAMIC")}


def cpu_matmul_baseline(m, k, n, budget_s=25.0, ge_budget_s=8.0):
    """The oracle's restatement of the reference M4RI product (matmul_m4ri.c:
    178-248, kernel-only on packed operands, OpenMP over 256-row blocks) timed
    on the host cores with the reference's timing policy: the full n^3
    workload when it fits the budget, else an m x n x m sample (m halved until
    it does).  Beside it, the restated AND+popcount product
    (mela_matmul_gf2.c:110-208) on a row sample sized from one timed call."""
    from oracle import oracle as orc
    m_full = m
    while True:
        # warm-up + doubling + 2 rounds of one call each ~ 4 calls at this size
        est_s = 2.0 * m * n * k / 40e12
        if est_s * 4 <= budget_s or m <= 1024:
            break
        m //= 2
    Ap = orc.gen_bernoulli_packed(1, m, k)
    Bp = orc.gen_bernoulli_packed(2, k, n)
    dt, rounds, iters = reference_timing(lambda: orc.matmul_packed(Ap, m, k, Bp, n, "m4ri"))
    out = {"value": round(2.0 * m * n * k / dt / 1e12, 3), "unit": "Tbitop/s",
           "cores": orc.num_threads(), "kind": "port",
           "sample": f"oracle restatement of the reference M4RI product (k=8 tables, 256-row "
                     f"blocks, OpenMP) on packed {m}x{k} * {k}x{n} Bernoulli(1/2); timing policy "
                     f"of benchmark_mex.m:671-707 (warm-up, {rounds} rounds x {iters} call(s), "
                     f"median): {dt:.2f} s per call"}
    out.update(host_metadata(orc))
    # GE restatement: time 64 rows once, then the largest power-of-two row
    # count whose ~4 calls fit ge_budget_s
    g = 64
    Ag = orc.gen_bernoulli_packed(3, g, k)
    t0 = time.perf_counter()
    orc.matmul_packed(Ag, g, k, Bp, n, "gf2")
    per_row = (time.perf_counter() - t0) / g
    while g * 2 <= m_full and 4 * per_row * g * 2 <= ge_budget_s:
        g *= 2
    Ag = orc.gen_bernoulli_packed(3, g, k)
    dg, rounds, iters = reference_timing(lambda: orc.matmul_packed(Ag, g, k, Bp, n, "gf2"))
    out["ge"] = {"value": round(2.0 * g * n * k / dg / 1e12, 3), "unit": "Tbitop/s",
                 "sample": f"oracle restatement of the reference AND+popcount product on packed "
                           f"{g}x{k} * {k}x{n}; {rounds} rounds x {iters} call(s), median "
                           f"{dg:.3f} s per call"}
    return out


def resolve_auto(m, k, n):
    """The library's AUTO product choice (csrc/matmul.hip)."""
    return "mfma_fp4" if m * n * k >= 2 ** 33 and k < 2 ** 24 else "andpop"


def roofline_for(algo, achieved, traffic, big_tiles=True):
    """Roofline of the dominant kernel of each product algorithm (W = 2mnk).
    achieved = W / average duration of that kernel's launches, timed with HIP
    events around the kernel alone on the library stream."""
    if algo in ("mfma_fp4", "mfma_i8"):
        # dense MFMA: fp4 10.1 PFLOP/s, int8 5.0 POP/s; one AND+XOR pair of W
        # is one multiply-add (2 ops), so W/s compares directly with ops/s
        peak = 10100.0 if algo == "mfma_fp4" else 5000.0
        if algo == "mfma_fp4":
            kern = "k_mfma_direct (packed fp4 planes)" if big_tiles else "k_mfma_small<fp4>"
        else:
            kern = "k_mfma_ring<int8>" if big_tiles else "k_mfma_small<int8>"
        return {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak,
                "unit": "Tbitop/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                "kernel": kern,
                "note": "achieved = 2mnk / HIP-event duration of the MFMA kernel launch (the "
                        "operand re-layout kernels are excluded; they are in ms_per_step)"}
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": VALU_BITOP3_PEAK,
            "unit": "Tbitop/s", "frac": round(achieved / VALU_BITOP3_PEAK, 4), "traffic": traffic,
            "kernel": "k_andpop" if algo == "andpop" else "k_m4rm",
            "note": "achieved = 2mnk / HIP-event duration of the kernel launch on the library "
                    "stream; peak = bitwise VALU with AND+XOR fused in v_bitop3_b32 (2 x 1258)"}


def bench_matmul(args, mela):
    algo = {"auto": mela.ALGO_AUTO, "m4rm": mela.ALGO_M4RM, "andpop": mela.ALGO_ANDPOP,
            "mfma_i8": mela.ALGO_MFMA_I8, "mfma_fp4": mela.ALGO_MFMA_FP4}[args.algo]
    if args.shape:
        m, k, n = (int(v) for v in args.shape.lower().split("x"))
    else:
        m = k = n = args.n
    ctx = mela.Context(0)
    A, B, Cm = ctx.bitmat(m, k), ctx.bitmat(k, n), ctx.bitmat(m, n)
    ctx.gen_bernoulli(1, A)
    ctx.gen_bernoulli(2, B)
    ctx.sync()
    for _ in range(args.warmup):
        ctx.matmul_packed(A, B, Cm, algo)
    ctx.sync()
    ev = HipEvents()
    e0, e1 = ev.event(), ev.event()
    stream = ctx.stream
    t0 = time.perf_counter()
    ev.record(e0, stream)
    for _ in range(args.steps):
        ctx.matmul_packed(A, B, Cm, algo)
    ev.record(e1, stream)
    ctx.sync()
    t1 = time.perf_counter()
    dev_ms = ev.elapsed_ms(e0, e1) / args.steps
    wall_ms = (t1 - t0) * 1e3 / args.steps
    # the dominant kernel alone, in a second pass of the same calls
    ctx.profile_begin()
    for _ in range(args.steps):
        ctx.matmul_packed(A, B, Cm, algo)
    kern_total, launches = ctx.profile_end()
    kern_ms = kern_total / max(launches, 1)
    check = freivalds_check(ctx, mela, A, B, Cm, m, k, n)
    progress(f"product {m}x{k}x{n}: {check}")
    work = 2.0 * m * k * n
    tbit = work / (wall_ms * 1e-3) / 1e12
    achieved = work / (kern_ms * 1e-3) / 1e12
    resolved = resolve_auto(m, k, n) if args.algo == "auto" else args.algo
    square = m == k == n
    out = {
        "metric": ("GF(2) matmul Tbitop/s at 32768^3" if square and n == 32768
                   else f"GF(2) matmul Tbitop/s at {m}x{k}x{n}"),
        "value": round(tbit, 2),
        "unit": "Tbitop/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp4 e2m1 operands (GF(2) bits as 0/1), fp32 accumulation, result mod 2"
                  if resolved == "mfma_fp4" else
                  "int8 operands (GF(2) bits as 0/1), int32 accumulation, result mod 2"
                  if resolved == "mfma_i8" else
                  "u1 (GF(2) bits packed in u64 words; AND/XOR/popcount)"),
        "data": "synthetic Bernoulli(1/2) bits, splitmix64 seeds 1 (A) and 2 (B), generated on device",
        "config": {"workload": f"packed GF(2) matmul {m}x{k} * {k}x{n}, operands resident in HBM",
                   "algo": args.algo, "algo_resolved": resolved, "parallelism": "1 GPU"},
        "roofline": roofline_for(resolved, achieved,
                                 args.traffic if args.traffic is not None
                                 else MEASURED_TRAFFIC.get((resolved, n)) if square else None,
                                 big_tiles=((m + 255) // 256) * ((n + 255) // 256) >= 256),
        "call_ms": round(dev_ms, 3),
        "kernel_ms": round(kern_ms, 3),
        "check": check,
    }
    for M in (A, B, Cm):
        ctx.free(M)
    progress(f"product {m}x{k}x{n}: {wall_ms:.3f} ms per step")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_matmul_baseline(m, k, n)
        progress("cpu baseline done")
    if args.workload == "matmul" and not args.headline_only:
        out["batched"] = bench_batch(args, mela)
        progress("batched rank+null done")
        out["batched_matmul"] = bench_batch_matmul(args, mela)
        progress("batched matmul done")
        saved = (args.serve_rates, args.serve_seconds)
        args.serve_rates, args.serve_seconds = args.headline_serve_rates, 1.0
        out["serve"] = bench_serve(args, mela)
        args.serve_rates, args.serve_seconds = saved
        progress("serve done")
        out["c1"] = bench_c1(args, mela)
        progress("c1 done")
    print(json.dumps(out), flush=True)


def freivalds_check(ctx, mela, A, B, Cm, m, k, n, vectors=64):
    """C = A*B on the device, checked by 64 simultaneous Freivalds vectors:
    X (n x 64 random bits), C*X against A*(B*X), the narrow products by the
    AND+parity kernel (not the kernel timed).  A wrong C passes with
    probability 2^-64.  Raises on a mismatch."""
    X = ctx.bitmat(n, vectors)
    Y = ctx.bitmat(k, vectors)
    Z1 = ctx.bitmat(m, vectors)
    Z2 = ctx.bitmat(m, vectors)
    try:
        ctx.gen_bernoulli(99, X)
        ctx.matmul_packed(B, X, Y, mela.ALGO_ANDPOP)
        ctx.matmul_packed(A, Y, Z1, mela.ALGO_ANDPOP)
        ctx.matmul_packed(Cm, X, Z2, mela.ALGO_ANDPOP)
        ok = bool((ctx.download(Z1) == ctx.download(Z2)).all())
    finally:
        for M in (X, Y, Z1, Z2):
            ctx.free(M)
    if not ok:
        raise SystemExit(f"product {m}x{k}x{n}: C*X != A*(B*X), the timed product is wrong")
    return f"Freivalds, {vectors} vectors: C*X == A*(B*X)"


def elim_bytes_alg(m, n, pivots, rref):
    """HBM bytes of the panel sweeps: for each group of 64 consecutive pivots
    (one packed-word panel) 2*8*rows*(wc - first_pivot_col/64), rows = m - 64g
    (rank) or m (RREF)."""
    wc = (n + 63) // 64
    tot = 0
    for g in range(0, len(pivots), 64):
        rows = m if rref else m - g
        tot += 16 * rows * (wc - pivots[g] // 64)
    return tot


def bench_elim(args, mela):
    from mex_linear_algebra_amd import workloads as W  # noqa: F401
    ctx = mela.Context(0)
    ev = HipEvents()
    cases = []
    # C2a: 16384 x 16384 = L * R with L 16384 x 12000, R 12000 x 16384 (rank <= 12000)
    L = ctx.bitmat(16384, 12000)
    R = ctx.bitmat(12000, 16384)
    ctx.gen_bernoulli(11, L)
    ctx.gen_bernoulli(12, R)
    H = ctx.bitmat(16384, 16384)
    ctx.matmul_packed(L, R, H, mela.ALGO_ANDPOP)
    ctx.free(L)
    ctx.free(R)
    cases.append(("C2a L*R 16384x16384 (inner 12000)", H, 16384, 16384))
    D = ctx.bitmat(16384, 20000)
    ctx.gen_bernoulli(13, D)
    cases.append(("C2b dense 16384x20000", D, 16384, 20000))
    for name, M, m, n in cases:
        Rm, piv = ctx.rref_packed(M)
        ctx.free(Rm)
        names = {"ge": mela.ELIM_GE, "m4ri": mela.ELIM_M4RI, "panel": mela.ELIM_PANEL,
                 "auto": mela.ELIM_AUTO, "blocked": mela.ELIM_BLOCKED}
        for vname in args.elim_variants.split(","):
            variant = names[vname]
            for op in ("rank", "null"):
                # two untimed calls: the second one captures the panel sequence
                # as a hipGraph, which the timed calls replay
                for _ in range(2):
                    if op == "rank":
                        ctx.rank_packed(M, variant)
                    else:
                        ctx.free(ctx.null_packed(M, variant))
                ctx.sync()
                e0, e1 = ev.event(), ev.event()
                reps = max(1, args.elim_reps)
                c0 = ctx.elim_counts()
                t0 = time.perf_counter()
                ev.record(e0, ctx.stream)
                for _ in range(reps):
                    if op == "rank":
                        r = ctx.rank_packed(M, variant)
                    else:
                        Z = ctx.null_packed(M, variant)
                        r = n - Z.cols
                        ctx.free(Z)
                ev.record(e1, ctx.stream)
                ctx.sync()
                wall = (time.perf_counter() - t0) * 1e3 / reps
                dev = ev.elapsed_ms(e0, e1) / reps
                c1 = ctx.elim_counts()
                path = ("blocked" if c1[0] > c0[0] else "") + ("fallback" if c1[1] > c0[1] else "") + \
                       ("panel" if c1[2] > c0[2] else "")
                bytes_alg = elim_bytes_alg(m, n, piv, op == "null")
                if op == "null":
                    bytes_alg += n * ((n - len(piv) + 63) // 64) * 8
                print(json.dumps({
                    "workload": f"{op} {vname} {name}", "path": path, "rank": r,
                    "ms": round(wall, 3),
                    "device_ms": round(dev, 3), "bytes_alg_GB": round(bytes_alg / 1e9, 3),
                    "GBps_alg": round(bytes_alg / (dev * 1e-3) / 1e9, 1),
                    "hbm_frac": round(bytes_alg / (dev * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}),
                    flush=True)
        ctx.free(M)


def bench_batch(args, mela):
    from mex_linear_algebra_amd import workloads as W
    ctx = mela.Context(0)
    t0 = time.perf_counter()
    items = W.batch_c4(args.batch_count, seed=4)
    gen_s = time.perf_counter() - t0
    dev = [ctx.upload(P, n) for P, m, n, _ in items]
    b = mela.Batch(ctx, dev, mela.BATCH_NULL | mela.BATCH_GRAPH)
    tuned_ms = b.autotune()
    tw, md, gr = b.plan()
    # modes as they run: k = 8 tables only with 8-word tiles
    eff = [0 if x == 0 else (8 if x == 2 and tw == 8 else 4) for x in md]
    plan = {"tile_words": tw, "ge": eff.count(0), "m4ri_k4": eff.count(4), "m4ri_k8": eff.count(8),
            "multi_workgroup_matrices": sum(g > 1 for g in gr)}
    b.run()
    b.results()
    ev = HipEvents()
    e0, e1 = ev.event(), ev.event()
    ctx.sync()
    t0 = time.perf_counter()
    ev.record(e0, ctx.stream)
    for _ in range(args.steps):
        b.run()
    ev.record(e1, ctx.stream)
    ranks, nulls = b.results()
    wall = (time.perf_counter() - t0) / args.steps
    dev_ms = ev.elapsed_ms(e0, e1) / args.steps
    rate = len(items) / wall
    # plain form: one elimination per matrix through the single-matrix API
    # (null_packed gives the rank, n - Z.cols, and the basis)
    sidx = list(range(0, len(items), max(1, len(items) // args.plain_sample)))
    ctx.sync()
    t0 = time.perf_counter()
    for i in sidx:
        Z = ctx.null_packed(dev[i])
        ctx.free(Z)
    ctx.sync()
    plain_s = (time.perf_counter() - t0) / len(sidx)
    plain_rate = 1.0 / plain_s
    # sweep bytes of rank + null basis (RREF rows = m) with each matrix's
    # pivots taken in its leading columns (the most words a sweep can cover),
    # plus the packed null basis written
    bytes_alg = 0
    for (P, m, n, _), r in zip(items, ranks):
        wc = (n + 63) // 64
        bytes_alg += sum(16 * m * (wc - j) for j in range((r + 63) // 64))
        bytes_alg += n * ((n - r + 63) // 64) * 8
    row = {
        "workload": f"batched rank+null, {len(items)} uneven matrices (m log-uniform 64..2048, "
                    f"n=m+U[0,m], 90% column-weight 3/4/6, 10% dense)",
        "batched_ms": round(wall * 1e3, 3), "batched_device_ms": round(dev_ms, 3),
        "matrices_per_s": round(rate, 1), "autotune_best_ms": round(tuned_ms, 3), "plan": plan,
        "plain_one_per_launch_matrices_per_s": round(plain_rate, 1),
        "plain_sample": len(sidx), "speedup_vs_plain": round(rate / plain_rate, 2),
        "generate_s": round(gen_s, 2), "mean_rank": sum(ranks) / len(ranks),
        "bytes_alg_GB": round(bytes_alg / 1e9, 3),
        "GBps_alg": round(bytes_alg / (dev_ms * 1e-3) / 1e9, 1),
        "hbm_frac": round(bytes_alg / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    b.close()
    for M in dev:
        ctx.free(M)
    return row


def bench_batch_matmul(args, mela):
    """2048 uneven products C_i = A_i * B_i (workloads.batch_matmul_items: A_i
    as in the rank/null batch, B_i dense k x b, b log-uniform 32..1024) in one
    persistent, graph-captured launch (mela_matmul_batch_run) against one
    mela_matmul_packed call (AUTO) per product on a sample."""
    from mex_linear_algebra_amd import workloads as W
    ctx = mela.Context(0)
    items = W.batch_matmul_items(args.batch_mm_count, seed=5)
    dA = [ctx.upload(A, k) for A, m, k, B, n, _ in items]
    dB = [ctx.upload(B, n) for A, m, k, B, n, _ in items]
    dC = [ctx.bitmat(m, n) for A, m, k, B, n, _ in items]
    b = mela.MatmulBatch(ctx, dA, dB, dC)
    algos = b.algos()
    for _ in range(max(1, args.warmup)):
        b.run()
    ev = HipEvents()
    e0, e1 = ev.event(), ev.event()
    ctx.sync()
    t0 = time.perf_counter()
    ev.record(e0, ctx.stream)
    for _ in range(args.steps):
        b.run()
    ev.record(e1, ctx.stream)
    ctx.sync()
    wall = (time.perf_counter() - t0) / args.steps
    dev_ms = ev.elapsed_ms(e0, e1) / args.steps
    sidx = list(range(0, len(items), max(1, len(items) // args.plain_sample)))
    for i in sidx[:8]:
        ctx.matmul_packed(dA[i], dB[i], dC[i])
    ctx.sync()
    t0 = time.perf_counter()
    for i in sidx:
        ctx.matmul_packed(dA[i], dB[i], dC[i])
    ctx.sync()
    plain_rate = len(sidx) / (time.perf_counter() - t0)
    work = sum(2.0 * m * k * n for A, m, k, B, n, _ in items)
    rate = len(items) / wall
    row = {"workload": f"batched matmul, {len(items)} uneven products (A as the rank/null batch, "
                       f"B dense k x b, b log-uniform 32..1024)",
           "batched_ms": round(wall * 1e3, 3), "batched_device_ms": round(dev_ms, 3),
           "products_per_s": round(rate, 1), "Tbitops": round(work / wall / 1e12, 2),
           "plain_one_per_launch_products_per_s": round(plain_rate, 1),
           "plain_sample": len(sidx), "speedup_vs_plain": round(rate / plain_rate, 2),
           "items_andpop": algos.count(mela.ALGO_ANDPOP), "items_rowxor": algos.count(mela.ALGO_ROWXOR),
           "items_mfma": algos.count(mela.ALGO_MFMA_FP4)}
    b.close()
    for M in dA + dB + dC:
        ctx.free(M)
    return row


def bench_serve(args, mela):
    """Poisson arrivals of mixed requests (70% rank, 20% null, 10% matmul),
    sides log-uniform 128..8192, Bernoulli(1/2) or column weight 3..6, HBM
    admission cap 64 GB; offered load swept upward to saturation."""
    import threading
    import numpy as np
    from mex_linear_algebra_amd import workloads as W
    ctx = mela.Context(0)
    rng = np.random.default_rng(55)

    def side():
        return int(np.exp(rng.uniform(np.log(128), np.log(8192))))

    pool = []
    for i in range(args.serve_pool):
        m, n = side(), side()
        if i % 2:
            M = ctx.bitmat(m, n)
            ctx.gen_bernoulli(1000 + i, M)
        else:
            M = ctx.upload(W.parity_check_matrix(rng, m, n, int(rng.integers(3, 7))), n)
        pool.append(M)
    mm = []
    for i in range(max(4, args.serve_pool // 8)):
        m, k, n = side(), side(), side()
        A, B = ctx.bitmat(m, k), ctx.bitmat(k, n)
        ctx.gen_bernoulli(2000 + i, A)
        ctx.gen_bernoulli(3000 + i, B)
        mm.append((A, B))
    ctx.sync()
    srv = mela.Server(hbm_budget=64 * 10**9, workers=args.serve_workers, max_batch=64)
    # warm-up: every pool entry once
    for M in pool[:8]:
        t = srv.submit(mela.OP_RANK, M)
        srv.wait(t)
        srv.release(t)
    results = []
    for rate in args.serve_rates:
        srv.reset_stats()
        nreq = max(20, int(rate * args.serve_seconds))
        gaps = rng.exponential(1.0 / rate, nreq)
        ops = rng.choice([mela.OP_RANK, mela.OP_NULL, mela.OP_MATMUL], nreq, p=[0.7, 0.2, 0.1])
        picks = rng.integers(0, 1 << 30, nreq)
        tickets = []
        done_at = []
        errors = []
        lock = threading.Lock()

        def drain():
            i = 0
            while True:
                with lock:
                    if i >= nreq:
                        return
                    have = len(tickets)
                if i >= have:
                    time.sleep(0.0005)
                    continue
                t = tickets[i]
                try:
                    srv.wait(t)
                except mela.MelaError as e:  # counted, reported below
                    errors.append(str(e))
                done_at.append(time.perf_counter())
                srv.release(t)
                i += 1

        th = threading.Thread(target=drain)
        th.start()
        t0 = time.perf_counter()
        nxt = t0
        for j in range(nreq):
            nxt += gaps[j]
            dl = nxt - time.perf_counter()
            if dl > 0:
                time.sleep(dl)
            if ops[j] == mela.OP_MATMUL:
                A, B = mm[picks[j] % len(mm)]
                t = srv.submit(mela.OP_MATMUL, A, B)
            else:
                t = srv.submit(int(ops[j]), pool[picks[j] % len(pool)], variant=mela.ELIM_AUTO)
            with lock:
                tickets.append(t)
        th.join()
        span = max(done_at) - t0
        st = srv.stats()
        row = {"offered_req_per_s": rate, "requests": nreq, "achieved_req_per_s": round(nreq / span, 1),
               "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
               "mean_ms": round(st["mean_ms"], 3), "jobs": st["jobs"],
               "batched_requests": st["batched_requests"], "hbm_peak_GB": round(st["hbm_peak"] / 1e9, 3),
               "failed": len(errors)}
        results.append(row)
        if args.workload == "serve":
            print(json.dumps({"workload": "serve", **row}), flush=True)
    srv.close()
    for M in pool:
        ctx.free(M)
    for A, B in mm:
        ctx.free(A)
        ctx.free(B)
    return {"loads": results, "saturation_req_per_s": max(r["achieved_req_per_s"] for r in results),
            "workers": args.serve_workers, "hbm_cap_GB": 64,
            "mix": "Poisson arrivals; 70% rank / 20% null / 10% matmul, sides log-uniform "
                   "128..8192, half Bernoulli(1/2), half column weight 3..6"}


def bench_c1(args, mela):
    """n x n logical A, B (n = 2048 by default; column-major host arrays, as
    MATLAB passes them): matmul (both entry points), rank(A) and null(A) end to
    end through the C ABI -- host packing, copies, kernels and unpacking
    included -- next to the oracle's restatement of the reference on the same
    inputs (skipped with --no-cpu), both timed with the reference benchmark's
    policy."""
    import numpy as np
    rng = np.random.default_rng(1)
    n = args.c1_n
    cpu = not args.no_cpu
    if cpu:
        from oracle import oracle as orc
    # column-major like a MATLAB array, so no layout conversion is timed
    A = np.asfortranarray(rng.integers(0, 2, (n, n)).astype(bool))
    B = np.asfortranarray(rng.integers(0, 2, (n, n)).astype(bool))
    ctx = mela.Context(0)
    rows = []
    for name, gfn, cfn, work in (
            ("matmul_gf2", lambda: mela.mela_matmul_gf2(A, B, ctx=ctx),
             lambda: orc.matmul(A, B, "gf2"), 2.0 * n ** 3),
            ("matmul_m4ri", lambda: mela.mela_matmul_m4ri(A, B, ctx=ctx),
             lambda: orc.matmul(A, B, "m4ri"), 2.0 * n ** 3),
            ("rank_gf2", lambda: mela.mela_rank_gf2(A, ctx=ctx), lambda: orc.rank(A, "gf2"), None),
            ("rank_m4ri", lambda: mela.mela_rank_m4ri(A, ctx=ctx), lambda: orc.rank(A, "m4ri"), None),
            ("null_gf2", lambda: mela.mela_null_gf2(A, ctx=ctx), lambda: orc.null(A, "gf2"), None),
            ("null_m4ri", lambda: mela.mela_null_m4ri(A, ctx=ctx), lambda: orc.null(A, "m4ri"), None)):
        g = reference_timing(gfn)[0] * 1e3
        row = {"workload": f"C1 {name} {n}x{n} logical, end to end (host in/out)",
               "gpu_ms": round(g, 3)}
        if cpu:
            c = reference_timing(cfn)[0] * 1e3
            row.update({"cpu_oracle_ms": round(c, 3), "speedup": round(c / g, 1),
                        "cpu_threads": orc.num_threads()})
        if work:
            row["gpu_Tbitops_incl_pcie"] = round(work / (g * 1e-3) / 1e12, 2)
        rows.append(row)
        if args.workload == "c1":
            print(json.dumps(row), flush=True)
    return rows


def bench_copy(args, mela):
    ctx = mela.Context(0)
    nbytes = args.bytes
    src, dst = C.c_void_p(), C.c_void_p()
    L = ctx.lib
    assert L.mela_dev_alloc(ctx.handle, nbytes, C.byref(src)) == 0
    assert L.mela_dev_alloc(ctx.handle, nbytes, C.byref(dst)) == 0
    for _ in range(args.steps):
        assert L.mela_copy_d2d(ctx.handle, dst, src, nbytes) == 0
    ctx.sync()
    print(json.dumps({"workload": "copy", "bytes": nbytes, "launches": args.steps}), flush=True)
    L.mela_dev_free(ctx.handle, src)
    L.mela_dev_free(ctx.handle, dst)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="matmul",
                    choices=["matmul", "elim", "batch", "batch_matmul", "serve", "c1", "copy"])
    ap.add_argument("--headline-only", action="store_true",
                    help="the 32768^3 product line without the batched / serve / c1 keys")
    ap.add_argument("--algo", default="auto",
                    choices=["auto", "m4rm", "andpop", "mfma_i8", "mfma_fp4"])
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--shape", default=None,
                    help="MxKxN for a rectangular product (e.g. 65536x32768x4096)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch measured by rocprofv3 --pmc (scaled)")
    ap.add_argument("--elim-reps", type=int, default=2)
    ap.add_argument("--elim-variants", default="ge,m4ri,panel,auto",
                    help="elimination variants timed by --workload elim (ge, m4ri, panel, auto, blocked)")
    ap.add_argument("--c1-n", type=int, default=2048,
                    help="side of the square host arrays of the c1 workload")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-mm-count", type=int, default=2048)
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--serve-rates", type=float, nargs="+",
                    default=[50, 100, 200, 400, 800, 1600, 3200, 6400, 12800])
    ap.add_argument("--headline-serve-rates", type=float, nargs=2, default=[1600, 6400],
                    help="the two offered loads of the headline line's serve key")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-pool", type=int, default=48)
    ap.add_argument("--serve-workers", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)
    import mex_linear_algebra_amd as mela
    if args.workload in ("batch", "batch_matmul"):
        fn = bench_batch if args.workload == "batch" else bench_batch_matmul
        print(json.dumps(fn(args, mela)), flush=True)
    elif args.workload == "serve":
        res = bench_serve(args, mela)
        print(json.dumps({"workload": "serve summary",
                          **{k: v for k, v in res.items() if k != "loads"}}), flush=True)
    else:
        {"matmul": bench_matmul, "elim": bench_elim, "c1": bench_c1,
         "copy": bench_copy}[args.workload](args, mela)


if __name__ == "__main__":
    main()
