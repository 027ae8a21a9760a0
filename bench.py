#!/usr/bin/env python3
"""Benchmark of MTJ's CSR hot path on one MI355X.

A step is one FP64 CompRowMatrix.mult (y = A x, S/CompRowMatrix.java:294-314)
over the 3-D 7-point Poisson matrix on a 256^3 grid (n = 16,777,216,
nnz = 117,047,296), with A, x and y resident in HBM.  `value` is the
algorithmic SpMV bandwidth 12*nnz + 20*n + 4 bytes per product divided by the
mean device time per product (HIP events on the library's stream).  The CPU
baseline is the oracle's restatement of MTJ's loop on one host core over a
bounded number of products of the same matrix.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "matrix-toolkits-java_amd"))

import numpy as np  # noqa: E402

METRIC = "SpMV GB/s (% HBM peak); solves/s + p99 latency, served vs plain, rtol 1e-8"
HBM_PEAK_GBPS = 8000.0


class HipTimer:
    """hipEvent pair on a given stream (the library's own stream)."""

    def __init__(self, stream):
        path = "/opt/rocm/lib/libamdhip64.so"
        self.hip = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        self.stream = C.c_void_p(stream)
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(self.e0)) == 0
        assert self.hip.hipEventCreate(C.byref(self.e1)) == 0

    def start(self):
        assert self.hip.hipEventRecord(self.e0, self.stream) == 0

    def stop_ms(self):
        assert self.hip.hipEventRecord(self.e1, self.stream) == 0
        assert self.hip.hipEventSynchronize(self.e1) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), self.e0, self.e1) == 0
        return ms.value


def workload(name):
    from mtjgpu import problems as P
    if name == "poisson3d_256":
        return "3D 7-point Poisson 256^3, FP64 CSR, y = A x", P.poisson3d(256)
    if name == "powerlaw_4m":
        return "random CSR n=4M, power-law rows (alpha 2.1, 1-2000), FP64, y = A x", P.powerlaw(4_000_000, seed=0)
    if name == "poisson2d_256":
        return "2D 5-point Poisson 256^2, FP64 CSR, y = A x", P.poisson2d(256)
    raise SystemExit("unknown workload " + name)


def cpu_baseline(sysm, seconds):
    """Oracle restatement of MTJ's loop (one core) on the same matrix."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from mtjgpu import problems as P
    n, m, rp, ci, va = sysm
    O = oracle.Csr(n, m, rp, ci, va)
    x = P.uniform(m, 1)
    O.mult(x)  # warm
    t0 = time.perf_counter()
    calls = 0
    while True:
        O.mult(x)
        calls += 1
        el = time.perf_counter() - t0
        if el >= seconds or calls >= 1000:
            break
    per = el / calls
    return {"value": round(P.spmv_bytes(n, len(va)) / per / 1e9, 3), "unit": "GB/s", "cores": 1,
            "kind": "port",
            "sample": "%d oracle CompRowMatrix.mult calls on the same matrix (%.1f s, %.3f s/call)"
                      % (calls, el, per)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="poisson3d_256")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import mtjgpu as M
    from mtjgpu import problems as P

    desc, sysm = workload(args.workload)
    n, m, rp, ci, va = sysm
    nnz = len(va)
    bytes_spmv = P.spmv_bytes(n, nnz)
    ctx = M.Context(0)
    A = M.CompRowMatrix(n, m, rp, ci, va, ctx=ctx)
    x = M.DenseVector(P.uniform(m, 1), ctx)
    y = M.DenseVector(n, ctx)
    for _ in range(args.warmup):
        A.mult(x, y)
    ctx.synchronize()
    timer = HipTimer(ctx.stream)
    t0 = time.perf_counter()
    timer.start()
    for _ in range(args.steps):
        A.mult(x, y)
    dev_ms = timer.stop_ms()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    per_ms = dev_ms / args.steps
    gbps = bytes_spmv / (per_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gbps, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "n": n, "nnz": nnz, "bytes_per_spmv": bytes_spmv,
                   "pct_hbm_peak": round(100 * gbps / HBM_PEAK_GBPS, 2)},
        "roofline": {"bound": "hbm", "achieved": round(gbps, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "k_spmv<double,double,SPMV_MULT>",
                     "kernel_ms": round(per_ms, 5)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sysm, args.cpu_seconds)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
