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


def measured_traffic(workload):
    """HBM bytes per SpMV launch from the committed rocprofv3 PMC summary of
    this workload (FETCH_SIZE and WRITE_SIZE, each scaled by the same
    counter's reading for a 1 GiB copy), or None if there is none."""
    import re
    path = os.path.join(ROOT, "profiles", "spmv_%s_hbm_traffic.md" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9]+)", f.read())
    return (int(m.group(1)), os.path.relpath(path, ROOT)) if m else None


def cpu_info():
    """The host the CPU baselines ran on: logical CPUs visible to this
    process, the machine's count, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count()
    return {"nproc": visible, "machine_cpus": os.cpu_count(), "cpu_model": model}


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
    out = {"value": round(P.spmv_bytes(n, len(va)) / per / 1e9, 3), "unit": "GB/s", "cores": 1,
           "kind": "port",
           "sample": "%d oracle CompRowMatrix.mult calls on the same matrix (%.1f s, %.3f s/call)"
                     % (calls, el, per)}
    out.update(cpu_info())
    return out


def copy_calibration(steps):
    """Device-to-device copies of 1 GiB: the known-size reference whose HBM
    counters (FETCH_SIZE, WRITE_SIZE) scale the SpMV's counter readings.
    Also times the library's dot product over two 1 GiB vectors, a
    read-only stream, as the achievable-bandwidth reference for the SpMV."""
    import mtjgpu as M
    ctx = M.Context(0)
    n = 1 << 27  # 2^27 doubles = 1 GiB
    a = M.DenseVector(n, ctx)
    b = M.DenseVector(n, ctx)
    timer = HipTimer(ctx.stream)
    hip = timer.hip
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    timer.start()
    for _ in range(steps):
        assert hip.hipMemcpyAsync(b.device_ptr(), a.device_ptr(), n * 8, 3, timer.stream) == 0
    ms = timer.stop_ms() / steps
    for _ in range(3):
        a.dot(b)
    timer.start()
    for _ in range(steps):
        a.dot(b)
    ms_dot = timer.stop_ms() / steps
    print(json.dumps({"mode": "copy", "bytes_read": n * 8, "bytes_written": n * 8,
                      "ms_per_copy": round(ms, 5), "GBps_rw": round(2 * n * 8 / (ms * 1e-3) / 1e9, 1),
                      "ms_per_dot": round(ms_dot, 5),
                      "dot_read_GBps": round(2 * n * 8 / (ms_dot * 1e-3) / 1e9, 1)}))
    return 0


SOLVE_WORKLOADS = {
    # name: (description, generator, solver, preconditioner, rtol, oracle-feasible)
    "cg_poisson2d_256": ("2D 5-point Poisson 256^2, Jacobi-CG to rtol 1e-8", "p2d", "cg", "diag", 1e-8, True),
    "cg_poisson3d_256": ("3D 7-point Poisson 256^3, Jacobi-CG to rtol 1e-8", "p3d", "cg", "diag", 1e-8, False),
    "bicgstab_ilu_convdiff_128": ("3D upwind convection-diffusion 128^3 (Pe 10), BiCGstab + ILU(0) to rtol 1e-8",
                                  "cd", "bicgstab", "ilu", 1e-8, False),
    "gmres_ilu_convdiff_128": ("3D upwind convection-diffusion 128^3 (Pe 10), GMRES(30) + ILU(0) to rtol 1e-8",
                               "cd", "gmres", "ilu", 1e-8, False),
    "mixed_varcoef_192": ("3D variable-kappa Poisson 192^3, FP32 inner Jacobi-CG (1e-4) under FP64 refinement to 1e-10",
                          "vk", "mixed", "diag", 1e-10, False),
    "ir_cg32_varcoef_192": ("3D variable-kappa Poisson 192^3, IR with the FP32 inner Jacobi-CG preconditioner (1e-4) "
                            "to 1e-10", "vk", "ir", "cg32", 1e-10, False),
    "cg_powerlaw_spd_4m": ("random SPD CSR n=4M, S = B + B^T + diagonal dominance over the power-law B "
                           "(alpha 2.1, 1-2000), Jacobi-CG to rtol 1e-8", "pls", "cg", "diag", 1e-8, False),
    "cg_varcoef_192": ("3D variable-kappa Poisson 192^3, FP64 Jacobi-CG to 1e-10 (oracle of the mixed solve)",
                       "vk", "cg", "diag", 1e-10, False),
    "bicg_ilu_convdiff_64": ("3D upwind convection-diffusion 64^3 (Pe 10), BiCG + ILU(0) to rtol 1e-8",
                             "cd64", "bicg", "ilu", 1e-8, True),
    "cgs_ilu_convdiff_64": ("3D upwind convection-diffusion 64^3 (Pe 10), CGS + ILU(0) to rtol 1e-8",
                            "cd64", "cgs", "ilu", 1e-8, True),
    "qmr_ilu_convdiff_64": ("3D upwind convection-diffusion 64^3 (Pe 10), QMR + ILU(0) to rtol 1e-8",
                            "cd64", "qmr", "ilu", 1e-8, True),
    "cg_icc_poisson2d_256": ("2D 5-point Poisson 256^2, CG + ICC to rtol 1e-8", "p2d", "cg", "icc", 1e-8, True),
    "mg_amg_poisson2d_256": ("2D 5-point Poisson 256^2, AMG() cycles x = M(b, x) as a multigrid iteration "
                             "to rtol 1e-8", "p2d", "mg", "amg", 1e-8, True),
}


def _solve_system(gen):
    from mtjgpu import problems as P
    return {"p2d": lambda: P.poisson2d(256), "p3d": lambda: P.poisson3d(256),
            "cd": lambda: P.convdiff3d(128, peclet=10.0, seed=1),
            "cd64": lambda: P.convdiff3d(64, peclet=10.0, seed=1),
            "vk": lambda: P.varcoef_poisson(3, 192, seed=1),
            "pls": lambda: P.powerlaw_spd(4_000_000, seed=0)}[gen]()


# Full-size oracle solves recorded in tests/golden by make_fullsize_fixtures.py
# (the reference solution of each workload; the mixed-precision ones compare
# with the FP64 Jacobi-CG solve of the same system).
FIXTURE_OF = {"cg_poisson3d_256": "cg_poisson3d_256", "cg_powerlaw_spd_4m": "cg_powerlaw_spd_4m",
              "bicgstab_ilu_convdiff_128": "bicgstab_ilu_convdiff_128",
              "gmres_ilu_convdiff_128": "gmres_ilu_convdiff_128", "cg_varcoef_192": "cg_varcoef_192",
              "mixed_varcoef_192": "cg_varcoef_192", "ir_cg32_varcoef_192": "cg_varcoef_192"}
FIXTURE_FIELDS = ("iterations", "code", "residual", "b_norm", "x_norm", "err_norm", "true_rel_residual", "n",
                  "nnz", "cpu_s", "rtol", "restart")


def load_fixture(name):
    path = os.path.join(ROOT, "tests", "golden", "fullsize_%s.npz" % name)
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as f:
        out = dict(zip(FIXTURE_FIELDS, f["meta"].tolist()))
        out["idx"], out["x_samples"] = f["idx"], f["x_samples"]
    return out


def oracle_solve_sample(name, sysm, bh, iterations_full, budget_s):
    """The oracle's one-core solve of the same system: in full when its
    recorded time fits the budget, else the first k iterations (k sized from
    the recorded time per iteration), extrapolated to the full count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    desc, gen, method, prec, rtol, _ = SOLVE_WORKLOADS[name]
    fx = load_fixture(FIXTURE_OF.get(name, name))
    O = oracle.Csr(*sysm)
    n = sysm[0]
    if method in ("mixed", "ir"):
        # the mixed-precision schemes are measured against the FP64 Jacobi-CG
        # solve of the same system
        method, prec = "cg", "diag"
    kind = {"diag": oracle.PREC_DIAGONAL, "ilu": oracle.PREC_ILU}[prec]
    t0 = time.perf_counter()
    op = oracle.Precond(kind, O)
    setup = time.perf_counter() - t0
    per_it = fx["cpu_s"] / max(fx["iterations"], 1) if fx else None
    k = iterations_full
    if per_it is not None and per_it * iterations_full > budget_s:
        k = max(1, int(budget_s / per_it))
        if method == "gmres":
            k = max(31, k // 31 * 31)  # whole restart cycles (30 Arnoldi steps + the restart)
    t0 = time.perf_counter()
    code, xo, it_o, _ = oracle.solve(method, O, bh, np.zeros(n), op, oracle.monitor(k, rtol))
    dt = time.perf_counter() - t0
    full = it_o < k or code == oracle.OK
    out = {"kind": "port", "cores": 1, "method": method + "+" + prec, "setup_s": round(setup, 3),
           "iterations_timed": int(it_o), "seconds": round(dt, 3)}
    if full:
        out.update({"solve_s": round(dt, 3), "iterations": int(it_o), "sample": "the whole solve"})
    else:
        out.update({"s_per_iteration": round(dt / max(it_o, 1), 5),
                    "solve_s_est": round(dt / max(it_o, 1) * iterations_full, 2),
                    "iterations": int(iterations_full),
                    "sample": "the first %d of %d iterations, extrapolated" % (it_o, iterations_full)})
    out.update(cpu_info())
    return out


def _mg_gpu(M, Mp, A, b, x, bh, rtol, max_cycles=200):
    """x = M(b, x) until ||b - A x|| < rtol ||b||; returns the cycle count."""
    r = M.DenseVector(len(bh), x.ctx)
    nb = float(np.linalg.norm(bh))
    for it in range(1, max_cycles + 1):
        Mp.apply(b, x)
        r.set(b)
        A.multAdd(-1.0, x, r)
        if r.norm(M.Norm.Two) < rtol * nb:
            return it
    return max_cycles


def _mg_oracle(oracle, O, g, bh, rtol, max_cycles=200):
    x = np.zeros(O.nr)
    nb = float(np.linalg.norm(bh))
    for it in range(1, max_cycles + 1):
        _, x = g.apply(bh, x)
        if np.linalg.norm(O.mult_add(-1.0, x, bh)) < rtol * nb:
            return it, x
    return max_cycles, x


def solve_mode(name, repeats, cpu_budget=0.0):
    """Times one solve of a named configuration: x_true ~ U[-1,1) (seed 1),
    b = A x_true, x0 = 0.  Reports iterations, device time, true relative
    residual and error, and (where it runs in seconds) the oracle's solve."""
    import mtjgpu as M
    from mtjgpu import problems as P
    desc, gen, method, prec, rtol, with_oracle = SOLVE_WORKLOADS[name]
    sysm = _solve_system(gen)
    n, m, rp, ci, va = sysm
    import scipy.sparse as sp
    Ah = sp.csr_matrix((va, ci, rp), shape=(n, n))
    x_true = P.uniform(n, 1)
    bh = Ah @ x_true
    ctx = M.Context(0)
    A = M.CompRowMatrix(n, m, rp, ci, va, ctx=ctx)
    b = M.DenseVector(bh, ctx)
    x = M.DenseVector(n, ctx)
    S = None
    if method != "mg":
        cls = {"cg": M.CG, "bicgstab": M.BiCGstab, "gmres": M.GMRES, "mixed": M.MixedCG, "ir": M.IR, "bicg": M.BiCG,
               "cgs": M.CGS, "qmr": M.QMR}[method]
        S = cls(x)
    Mp = {"diag": lambda: M.DiagonalPreconditioner(n, ctx), "ilu": lambda: M.ILU(A), "icc": lambda: M.ICC(A),
          "amg": lambda: M.AMG(ctx=ctx), "cg32": lambda: M.FP32CGPreconditioner(n, inner_rtol=1e-4, ctx=ctx)}[prec]()
    t0 = time.perf_counter()
    Mp.setMatrix(A)
    ctx.synchronize()
    setup = time.perf_counter() - t0
    if S is not None:
        if method != "mixed":
            S.setPreconditioner(Mp)
        S.setIterationMonitor(M.DefaultIterationMonitor(100 if method == "mixed" or prec == "cg32" else 100000, rtol))
    times = []
    for _ in range(repeats):
        x.zero()
        ctx.synchronize()
        t0 = time.perf_counter()
        if S is None:
            iters = _mg_gpu(M, Mp, A, b, x, bh, rtol)
        else:
            S.solve(A, b, x)
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    xs = x.getData()
    if S is not None:
        iters = S.getIterationMonitor().iterations()
    out = {"mode": "solve", "workload": name, "config": desc, "n": n, "nnz": len(va),
           "method": method, "preconditioner": prec, "rtol": rtol,
           "iterations": iters,
           "gpu_s": round(min(times), 5), "gpu_s_all": [round(t, 5) for t in times],
           "setup_s": round(setup, 4),
           "true_rel_residual": float(np.linalg.norm(bh - Ah @ xs) / np.linalg.norm(bh)),
           "rel_error_vs_x_true": float(np.linalg.norm(xs - x_true) / np.linalg.norm(x_true))}
    if prec == "cg32":
        out["inner_iterations"] = Mp.inner_iterations() // max(repeats, 1)
    fx = load_fixture(FIXTURE_OF.get(name, ""))
    if fx is not None:
        xt = float(np.linalg.norm(x_true))
        out["oracle_fixture"] = {
            "workload": FIXTURE_OF[name], "iterations": int(fx["iterations"]),
            "true_rel_residual": fx["true_rel_residual"], "rel_err_vs_x_true": fx["err_norm"] / xt,
            "rel_diff_x_sampled": float(np.linalg.norm(xs[fx["idx"]] - fx["x_samples"]) /
                                        np.linalg.norm(fx["x_samples"])),
            "cpu_s_build_container": round(fx["cpu_s"], 2)}
        if cpu_budget > 0:
            out["cpu_baseline"] = oracle_solve_sample(name, sysm, bh, int(fx["iterations"]), cpu_budget)
    if method == "mixed":
        out["inner_iterations"] = S.inner_iterations()
        # the FP64 reference: Jacobi-CG on the device, same system and tolerance
        x64 = M.DenseVector(n, ctx)
        C64 = M.CG(x64)
        C64.setPreconditioner(M.DiagonalPreconditioner(n, ctx))
        C64.getPreconditioner().setMatrix(A)
        C64.setIterationMonitor(M.DefaultIterationMonitor(100000, rtol))
        C64.solve(A, b, x64)
        xr = x64.getData()
        out["fp64_cg"] = {"iterations": C64.getIterationMonitor().iterations(),
                          "true_rel_residual": float(np.linalg.norm(bh - Ah @ xr) / np.linalg.norm(bh)),
                          "rel_error_vs_x_true": float(np.linalg.norm(xr - x_true) / np.linalg.norm(x_true)),
                          "rel_diff_mixed_vs_fp64_cg": float(np.linalg.norm(xs - xr) / np.linalg.norm(xr))}
    if with_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        O = oracle.Csr(n, m, rp, ci, va)
        t0 = time.perf_counter()
        if prec == "amg":
            g = oracle.Amg(O)
        else:
            op = oracle.Precond({"diag": oracle.PREC_DIAGONAL, "ilu": oracle.PREC_ILU,
                                 "icc": oracle.PREC_ICC}[prec], O)
        setup_o = time.perf_counter() - t0
        t0 = time.perf_counter()
        if method == "mg":
            it_o, xo = _mg_oracle(oracle, O, g, bh, rtol)
        else:
            code, xo, it_o, _ = oracle.solve(method, O, bh, np.zeros(n), op, oracle.monitor(100000, rtol))
        out["oracle"] = {"kind": "port", "cores": 1, "iterations": it_o, "cpu_s": round(time.perf_counter() - t0, 4),
                         "setup_s": round(setup_o, 4),
                         "rel_diff_x": float(np.linalg.norm(xs - xo) / np.linalg.norm(xo)),
                         "bitwise_x": bool(xs.tobytes() == xo.tobytes())}
    print(json.dumps(out))
    return 0


PRECOND_WORKLOADS = {
    "p2d_256": ("2D 5-point Poisson 256^2", lambda P: P.poisson2d(256)),
    "p3d_64": ("3D 7-point Poisson 64^3", lambda P: P.poisson3d(64)),
    "cd_64": ("3D upwind convection-diffusion 64^3 (Pe 10)", lambda P: P.convdiff3d(64, peclet=10.0, seed=1)),
    "cd_128": ("3D upwind convection-diffusion 128^3 (Pe 10)", lambda P: P.convdiff3d(128, peclet=10.0, seed=1)),
}


def precond_mode(prec, wl, steps, with_cpu=True):
    """One preconditioner apply on the device, timed with HIP events over
    `steps` applies, beside the oracle's apply on one core; the two results
    are compared bit for bit.  SSOR and AMG start from x's incoming values,
    so every apply starts from the same x (copied in, inside the timing)."""
    import mtjgpu as M
    from mtjgpu import problems as P
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    desc, gen = PRECOND_WORKLOADS[wl]
    n, m, rp, ci, va = gen(P)
    ctx = M.Context(0)
    A = M.CompRowMatrix(n, m, rp, ci, va, ctx=ctx)
    Mp = {"ilu": lambda: M.ILU(A), "ssor": lambda: M.SSOR(A), "icc": lambda: M.ICC(A),
          "amg": lambda: M.AMG(ctx=ctx)}[prec]()
    t0 = time.perf_counter()
    Mp.setMatrix(A)
    ctx.synchronize()
    setup = time.perf_counter() - t0
    bh = P.uniform(n, 4)
    x0 = P.uniform(n, 5)
    b = M.DenseVector(bh, ctx)
    xs = M.DenseVector(x0, ctx)
    x = M.DenseVector(n, ctx)
    for _ in range(2):
        x.set(xs)
        Mp.apply(b, x)
    ctx.synchronize()
    timer = HipTimer(ctx.stream)
    timer.start()
    for _ in range(steps):
        x.set(xs)
        Mp.apply(b, x)
    ms = timer.stop_ms() / steps
    xg = x.getData()
    if not with_cpu:
        out = {"mode": "precond", "preconditioner": prec, "workload": wl, "config": desc, "n": n, "nnz": len(va),
               "apply_ms": round(ms, 4), "setup_s": round(setup, 4)}
        if prec == "amg":
            out["levels"] = Mp.levels()
        print(json.dumps(out))
        return 0
    O = oracle.Csr(n, m, rp, ci, va)
    t0 = time.perf_counter()
    if prec == "amg":
        g = oracle.Amg(O)
    else:
        op = oracle.Precond({"ilu": oracle.PREC_ILU, "ssor": oracle.PREC_SSOR, "icc": oracle.PREC_ICC}[prec], O)
    setup_o = time.perf_counter() - t0
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        xo = g.apply(bh, x0)[1] if prec == "amg" else op.apply(bh, x0)
    cpu_ms = (time.perf_counter() - t0) / reps * 1e3
    out = {"mode": "precond", "preconditioner": prec, "workload": wl, "config": desc, "n": n, "nnz": len(va),
           "apply_ms": round(ms, 4), "setup_s": round(setup, 4),
           "cpu_baseline": {"apply_ms": round(cpu_ms, 3), "setup_s": round(setup_o, 4), "cores": 1,
                            "kind": "port"},
           "speedup_apply": round(cpu_ms / ms, 2), "bitwise_equal": bool(xg.tobytes() == xo.tobytes())}
    if prec == "amg":
        out["levels"] = Mp.levels()
    print(json.dumps(out))
    return 0


def serve_mode(count, lam, seed, tune=True, plain=True):
    """Served stream: `count` requests of the 2D/3D variable-kappa Poisson
    stream (Poisson arrivals at `lam`/s, n log-uniform 1k-262k), solved to
    relative residual 1e-8 (a) by the batched mixed-precision engine with
    continuous admission and (b) one at a time by FP64 Jacobi-CG.  Systems are
    generated before the clock starts; uploads are inside the timed work."""
    import mtjgpu as M
    from mtjgpu import serving as S
    reqs = S.request_stream(count, lam=lam, seed=seed)
    t0 = time.perf_counter()
    systems = [S.make_system(r) for r in reqs]
    gen_s = time.perf_counter() - t0
    print("serve: %d systems generated in %.1f s" % (count, gen_s), file=sys.stderr, flush=True)
    ctx = M.Context(0)
    best, tinfo = tune_batch(ctx, lam, seed, tune)
    print("serve: launch shape %s" % best, file=sys.stderr, flush=True)
    st_b, res_b = S.serve_batched(ctx, reqs, systems, **best)
    print("serve: batched %.1f solves/s" % st_b["solves_per_s"], file=sys.stderr, flush=True)
    if not plain:
        worst_b = max(S.true_residual(systems[k], res_b[k]["x"]) for k in range(count))
        print(json.dumps({"mode": "serve", "requests": count, "lambda_per_s": lam,
                          "sum_n": int(sum(s[0] for s in systems)),
                          "batched_mixed": {k: round(v, 4) for k, v in st_b.items()},
                          "worst_true_rel_residual": {"batched": worst_b},
                          "all_converged": all(res_b[k]["code"] == 0 for k in range(count)), "autotune": tinfo}))
        return 0
    st_p, res_p = S.serve_plain(ctx, reqs, systems)
    print("serve: plain %.1f solves/s" % st_p["solves_per_s"], file=sys.stderr, flush=True)
    worst_b = max(S.true_residual(systems[k], res_b[k]["x"]) for k in range(count))
    worst_p = max(S.true_residual(systems[k], res_p[k]["x"]) for k in range(count))
    out = {"mode": "serve", "requests": count, "lambda_per_s": lam,
           "sum_n": int(sum(s[0] for s in systems)), "generation_s": round(gen_s, 2),
           "batched_mixed": {k: round(v, 4) for k, v in st_b.items()},
           "plain_fp64_cg": {k: round(v, 4) for k, v in st_p.items()},
           "speedup_solves_per_s": round(st_b["solves_per_s"] / st_p["solves_per_s"], 3),
           "worst_true_rel_residual": {"batched": worst_b, "plain": worst_p},
           "all_converged": all(res_b[k]["code"] == 0 and res_p[k]["code"] == 0 for k in range(count)),
           "autotune": tinfo}
    print(json.dumps(out))
    return 0


def mmload_mode(grid, repeats, cpu_seconds):
    """Matrix Market input: the 3D 7-point Poisson matrix grid^3 written in
    MTJ's MatrixVectorWriter layout to a temporary file, then read into a
    device CompRowMatrix (parse + CSR assembly + upload + SpMV schedule).
    The CPU baseline is the oracle's single-threaded restatement of
    CompRowMatrix(MatrixVectorReader) on a leading slice of the same file
    (its size line rewritten to the slice's entry count)."""
    import tempfile
    import mtjgpu as M
    from mtjgpu import matrix_market as MM
    from mtjgpu import problems as P
    n, _, rp, ci, va = P.poisson3d(grid)
    ctx = M.Context(0)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "poisson3d_%d.mtx" % grid)
        P.write_matrix_market(path, n, n, rp, ci, va)
        size = os.path.getsize(path)
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            A = MM.load_matrix(path, ctx)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
            A.close()
        t0 = time.perf_counter()
        MM.read_matrix(path)
        host_dt = time.perf_counter() - t0
        out = {"mode": "mmload", "workload": "poisson3d_%d" % grid, "file_bytes": size, "nnz": int(ci.size),
               "best_s_to_device": round(best, 4), "MB_per_s_to_device": round(size / best / 1e6, 1),
               "host_parse_s": round(host_dt, 4), "threads": min(os.cpu_count() or 1, 16)}
        if cpu_seconds > 0:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            with open(path, "rb") as f:
                head = f.readline()
                f.readline()
                # about 19 MB/s for the restatement: size the slice to the budget
                want = int(cpu_seconds * 18e6)
                body = f.read(want)
            body = body[:body.rfind(b"\n") + 1]
            cnt = body.count(b"\n")
            text = head + b"%10d %10d %19d\n" % (n, n, cnt) + body
            t0 = time.perf_counter()
            code = oracle.mm_matrix(text)[0]
            dt = time.perf_counter() - t0
            assert code == oracle.OK
            out["cpu_baseline"] = {"value": round(len(text) / dt / 1e6, 2), "unit": "MB/s", "cores": 1,
                                   "kind": "port", "sample": "first %d entries of the same file" % cnt}
    print(json.dumps(out))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="poisson3d_256")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["spmv", "copy", "solve", "serve", "mmload", "precond"], default="spmv")
    ap.add_argument("--prec", default="ilu", choices=["ilu", "ssor", "icc", "amg"])
    ap.add_argument("--prec-workload", default="p2d_256", choices=sorted(PRECOND_WORKLOADS))
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--serve-requests", type=int, default=8192,
                    help="requests of the served-vs-plain comparison attached to the SpMV line (0: none)")
    ap.add_argument("--lam", type=float, default=2000.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="serve with the batch engine's default launch shape instead of the autotuner's")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-plain", action="store_true",
                    help="serve mode: run only the batch engine (for profiling its kernels)")
    ap.add_argument("--solve", default="cg_poisson2d_256", choices=sorted(SOLVE_WORKLOADS))
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2
    if args.mode == "copy":
        return copy_calibration(args.steps)
    if args.mode == "solve":
        return solve_mode(args.solve, args.repeats, 0.0 if args.no_cpu else args.cpu_seconds * 2)
    if args.mode == "serve":
        return serve_mode(args.requests, args.lam, args.seed, not args.no_tune, not args.no_plain)
    if args.mode == "precond":
        return precond_mode(args.prec, args.prec_workload, min(args.steps, 50), not args.no_cpu)
    if args.mode == "mmload":
        return mmload_mode(args.grid, args.repeats, 0.0 if args.no_cpu else args.cpu_seconds)

    import mtjgpu as M
    from mtjgpu import problems as P

    desc, sysm = workload(args.workload)
    n, m, rp, ci, va = sysm
    nnz = len(va)
    bytes_spmv = P.spmv_bytes(n, nnz)
    ctx = M.Context(0)
    A = M.CompRowMatrix(n, m, rp, ci, va, ctx=ctx)
    fmt, stored = A.format()
    kernel = {"sell": "k_sell<double,double,SPMV_MULT>", "csr": "k_spmv<double,double,SPMV_MULT>",
              "csr_xcd": "k_cb_prod + k_spmv<double,double,SPMV_MULT,PRE>"}[fmt]
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
                     "kernel": kernel, "format": fmt, "stored_entries": stored,
                     "kernel_ms": round(per_ms, 5)},
    }
    traffic = measured_traffic(args.workload)
    if traffic:
        out["roofline"]["traffic"] = traffic[0]
        out["roofline"]["traffic_source"] = traffic[1]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sysm, args.cpu_seconds)
    if args.serve_requests > 0:
        for o in (A, x, y):
            o.close()
        out["serving"] = serve_summary(ctx, args.serve_requests, args.lam, args.seed, not args.no_tune)
    print(json.dumps(out))
    return 0


TUNE_SAMPLE = 48  # systems of the autotuner's sample (a separate stream seed)


def tune_batch(ctx, lam, seed, enabled=True):
    """Launch shape of the batch engine from the autotuner, run on a sample of
    the same request distribution (seed + 1) before any timed serving.  With
    tuning off the sample still runs once, with the engine's default shape,
    so both cases start the timed stream equally warm."""
    from mtjgpu import serving as S
    sample = [S.make_system(r) for r in S.request_stream(TUNE_SAMPLE, lam=lam, seed=seed + 1)]
    t0 = time.perf_counter()
    if enabled:
        best, table = S.autotune(ctx, sample)
    else:
        best, table = S.autotune(ctx, sample, rows_per_workgroup=(S.DEFAULT_ROWS_PER_WORKGROUP,),
                                 iters_per_step=(S.DEFAULT_ITERS_PER_STEP,))
    info = {"sample_systems": TUNE_SAMPLE, "key": [float(v) for v in S.shape_key(sample)], "chosen": best,
            "table": [[r, i, round(v, 1)] for r, i, v in table], "tune_s": round(time.perf_counter() - t0, 2)}
    return best, info


def serve_summary(ctx, count, lam, seed, tune=True):
    """The metric's second half on a bounded stream: `count` requests served
    by the batch engine (launch shape from the autotuner) and by plain
    per-request FP64 Jacobi-CG."""
    from mtjgpu import serving as S
    best, tinfo = tune_batch(ctx, lam, seed, tune)
    reqs = S.request_stream(count, lam=lam, seed=seed)
    t0 = time.perf_counter()
    systems = [S.make_system(r) for r in reqs]
    print("serve: %d systems generated in %.1f s" % (count, time.perf_counter() - t0), file=sys.stderr, flush=True)
    st_b, res_b = S.serve_batched(ctx, reqs, systems, **best)
    print("serve: batched %.1f solves/s" % st_b["solves_per_s"], file=sys.stderr, flush=True)
    st_p, res_p = S.serve_plain(ctx, reqs, systems)
    print("serve: plain %.1f solves/s" % st_p["solves_per_s"], file=sys.stderr, flush=True)
    worst = max(S.true_residual(systems[k], res_b[k]["x"]) for k in range(count))
    return {"requests": count, "lambda_per_s": lam, "rtol": 1e-8, "sum_n": int(sum(s[0] for s in systems)),
            "batched_solves_per_s": round(st_b["solves_per_s"], 2), "batched_p50_ms": round(st_b["p50_ms"], 1),
            "batched_p99_ms": round(st_b["p99_ms"], 1),
            "batched_gpu_busy_frac": round(st_b["busy_frac"], 3),
            "batched_stream_GBps_span": round(st_b["stream_GBps_span"], 1),
            "batched_stream_GBps_busy": round(st_b["stream_GBps_busy"], 1),
            "plain_solves_per_s": round(st_p["solves_per_s"], 2), "plain_p50_ms": round(st_p["p50_ms"], 1),
            "plain_p99_ms": round(st_p["p99_ms"], 1),
            "speedup": round(st_b["solves_per_s"] / st_p["solves_per_s"], 3),
            "batched_worst_true_rel_residual": worst,
            "all_converged": all(res_b[k]["code"] == 0 and res_p[k]["code"] == 0 for k in range(count)),
            "autotune": tinfo}


if __name__ == "__main__":
    sys.exit(main())
