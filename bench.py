#!/usr/bin/env python3
"""Benchmark: GF(2) matrix product C = A*B at 32768 x 32768 x 32768.

One step = one packed GF(2) product of two 32768 x 32768 Bernoulli(1/2) bit
matrices (128 MiB each, generated on the device and resident in HBM before
the timed region), through the C ABI (mela_matmul_packed).  Work per step is
W = 2*m*n*k bit operations (one AND and one XOR per (i,j,k) triple), reported
as Tbit-op/s.  Prints one JSON line.

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

VALU_PEAK_TBITOPS = 1258.0  # 256 CU x 64 lanes x 2.4 GHz x 32 bits (bitwise VALU)


class HipEvents:
    """hipEvent timing on a given stream (the library's own stream)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

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


def cpu_baseline(seconds_target=12.0):
    """Times the oracle's restatement of the reference M4RI product (kernel-only,
    packed, OpenMP on all host threads) on a bounded sample of the workload:
    m x 32768 x n with m = n chosen so the run takes ~seconds_target."""
    from oracle import oracle as orc
    k = 32768
    m = 512
    best = None
    while True:
        Ap = orc.gen_bernoulli_packed(1, m, k)
        Bp = orc.gen_bernoulli_packed(2, k, m)
        t0 = time.perf_counter()
        orc.matmul_packed(Ap, m, k, Bp, m, "m4ri")
        dt = time.perf_counter() - t0
        best = (m, dt)
        if dt * 4 > seconds_target or m >= 8192:
            break
        m *= 2
    m, dt = best
    tb = 2.0 * m * m * k / dt / 1e12
    return {"value": round(tb, 4), "unit": "Tbitop/s", "cores": orc.num_threads(),
            "kind": "port",
            "sample": f"oracle M4RM (matmul_m4ri.c:178-248 restated, k=8 tables, 256-row blocks) "
                      f"packed {m}x{k} * {k}x{m}, one call, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--algo", default="m4rm", choices=["auto", "m4rm", "andpop"])
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1"}))
        sys.exit(2)

    import mex_linear_algebra_amd as mela
    algo = {"auto": mela.ALGO_AUTO, "m4rm": mela.ALGO_M4RM, "andpop": mela.ALGO_ANDPOP}[args.algo]
    n = args.n
    ctx = mela.Context(0)
    A, B, Cm = ctx.bitmat(n, n), ctx.bitmat(n, n), ctx.bitmat(n, n)
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

    work = 2.0 * n * n * n
    tbit = work / (wall_ms * 1e-3) / 1e12
    achieved = work / (dev_ms * 1e-3) / 1e12
    out = {
        "metric": "GF(2) matmul Tbitop/s at 32768^3",
        "value": round(tbit, 2),
        "unit": "Tbitop/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u1 (packed GF(2) bits in u64 words)",
        "data": "synthetic Bernoulli(1/2) bits, splitmix64 seeds 1 (A) and 2 (B), device-generated",
        "config": {"workload": f"packed GF(2) matmul {n}x{n} * {n}x{n}, operands resident in HBM",
                   "algo": args.algo, "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": VALU_PEAK_TBITOPS,
                     "unit": "Tbitop/s", "frac": round(achieved / VALU_PEAK_TBITOPS, 4),
                     "traffic": None,
                     "note": "W = 2mnk bit-ops per launch / HIP-event launch time; M4RM does ~8x "
                             "fewer VALU ops than W, so frac may exceed 1"},
        "kernel_ms": round(dev_ms, 3),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    for M in (A, B, Cm):
        ctx.free(M)


if __name__ == "__main__":
    main()
