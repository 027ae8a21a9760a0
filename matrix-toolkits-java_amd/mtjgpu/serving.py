"""Serving many SPD solves on one GPU: a continuous-batching request loop over
the batched mixed-precision CG engine (mtjg_batch_*), and the plain baseline
that solves each request alone with FP64 Jacobi-CG (mtjg_solver_*).

Request stream: Poisson arrivals at rate `lam`; each request is a 2-D or 3-D
(coin flip) variable-coefficient Poisson system on a random grid whose size
n is log-uniform in [n_min, n_max], with kappa ~ lognormal(0,1) per cell,
x_true ~ U[-1,1) and b = A x_true.  Latency = completion time - arrival time.
"""
import ctypes as C
import math
import time

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from . import problems as P
from . import (CG, CompRowMatrix, DefaultIterationMonitor, DenseVector, DiagonalPreconditioner,
               IterativeSolverNotConvergedException, _ck)


class BatchServer:
    """Handle on one mtjg_batch (see include/mtjgpu.h)."""

    def __init__(self, ctx, max_slots=2048, max_workgroups=16384, arena_bytes=32 << 30,
                 rows_per_workgroup=1024, iters_per_step=16):
        self.ctx = ctx
        self._lib = ctx._lib
        h = C.c_void_p()
        _ck(self._lib.mtjg_batch_create(ctx.handle, max_slots, max_workgroups, int(arena_bytes),
                                        rows_per_workgroup, iters_per_step, C.byref(h)))
        self.handle = h

    def submit(self, n, rp, ci, va, b, rtol=1e-8, max_outer=100, inner_rtol=1e-4, max_inner=5000):
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        va = np.ascontiguousarray(va, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        rid = C.c_int64()
        _ck(self._lib.mtjg_batch_submit(self.handle, int(n), int(ci.size), rp.ctypes.data, ci.ctypes.data,
                                        va.ctypes.data, b.ctypes.data, float(rtol), int(max_outer),
                                        float(inner_rtol), int(max_inner), C.byref(rid)))
        return rid.value

    def step(self):
        a, c = C.c_int32(), C.c_int32()
        _ck(self._lib.mtjg_batch_step(self.handle, C.byref(a), C.byref(c)))
        return a.value, c.value

    def completed(self, max_ids=4096):
        ids = (C.c_int64 * max_ids)()
        cnt = C.c_int32()
        _ck(self._lib.mtjg_batch_completed(self.handle, ids, max_ids, C.byref(cnt)))
        return list(ids[:cnt.value])

    def result(self, rid, n):
        x = np.empty(n, np.float64)
        res = L.Result()
        inner = C.c_int32()
        code = self._lib.mtjg_batch_result(self.handle, int(rid), x.ctypes.data, C.byref(res), C.byref(inner))
        if code < 0:
            _ck(code)
        return x, code, res.iterations, inner.value, res.residual

    def info(self):
        q, a, u, s = C.c_int32(), C.c_int32(), C.c_int64(), C.c_int64()
        _ck(self._lib.mtjg_batch_info(self.handle, C.byref(q), C.byref(a), C.byref(u), C.byref(s)))
        return {"queued": q.value, "active": a.value, "arena_used": u.value, "arena_size": s.value}

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mtjg_batch_destroy(self.handle)
            self.handle = None


def request_stream(count, lam=2000.0, n_min=1000, n_max=262144, seed=7):
    """Arrival times and grid shapes of `count` requests (deterministic)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / lam, size=count)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    for k in range(count):
        n_target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        dim = 2 if rng.random() < 0.5 else 3
        aspect = math.exp(rng.uniform(math.log(0.5), math.log(2.0)))
        if dim == 2:
            nx = max(2, int(round(math.sqrt(n_target * aspect))))
            ny = max(2, int(round(n_target / nx)))
            shape = (nx, ny, 1)
        else:
            side = n_target ** (1.0 / 3.0)
            nx = max(2, int(round(side * aspect)))
            ny = max(2, int(round(side / math.sqrt(aspect))))
            nz = max(2, int(round(n_target / (nx * ny))))
            shape = (nx, ny, nz)
        reqs.append({"arrival": float(arrivals[k]), "dim": dim, "shape": shape, "seed": 1000 + k})
    return reqs


def make_system(req):
    """CSR (FP64) and right-hand side b = A x_true of one request."""
    nx, ny, nz = req["shape"]
    n, _, rp, ci, va = P.varcoef_poisson(req["dim"], nx, ny, nz, seed=req["seed"])
    x_true = P.uniform(n, req["seed"])
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    return n, rp, ci, va, A @ x_true


def _stats(lat, t_first, t_last, count):
    lat = np.asarray(lat)
    span = max(t_last - t_first, 1e-9)
    return {"solves": count, "solves_per_s": count / span, "span_s": span,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3)}


def serve_batched(ctx, reqs, systems, rtol=1e-8, **kw):
    """Real-time loop: submit each request when its arrival time has passed,
    step the batch, collect completions.  Returns (stats, results by index)."""
    B = BatchServer(ctx, **kw)
    try:
        order = sorted(range(len(reqs)), key=lambda k: reqs[k]["arrival"])
        id2k = {}
        results = {}
        lat = []
        nxt = 0
        t0 = time.perf_counter()
        t_last = t0
        while len(results) < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(order) and reqs[order[nxt]]["arrival"] <= now:
                k = order[nxt]
                n, rp, ci, va, b = systems[k]
                id2k[B.submit(n, rp, ci, va, b, rtol=rtol)] = k
                nxt += 1
            info = B.info()
            if info["active"] == 0 and info["queued"] == 0:
                if nxt < len(order):
                    time.sleep(max(0.0, reqs[order[nxt]]["arrival"] - (time.perf_counter() - t0)))
                continue
            B.step()
            tnow = time.perf_counter()
            for rid in B.completed():
                k = id2k.pop(rid)
                x, code, outer, inner, res = B.result(rid, systems[k][0])
                results[k] = {"x": x, "code": code, "outer": outer, "inner": inner}
                lat.append(tnow - t0 - reqs[k]["arrival"])
                t_last = tnow
        return _stats(lat, 0.0, t_last - t0, len(reqs)), results
    finally:
        B.close()


def serve_plain(ctx, reqs, systems, rtol=1e-8):
    """Each request solved alone, in arrival order, with FP64 Jacobi-CG
    through the ordinary solver path (upload, setMatrix, solve, download)."""
    order = sorted(range(len(reqs)), key=lambda k: reqs[k]["arrival"])
    results = {}
    lat = []
    t0 = time.perf_counter()
    t_last = t0
    for k in order:
        wait = reqs[k]["arrival"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        n, rp, ci, va, b = systems[k]
        A = CompRowMatrix(n, n, rp, ci, va, ctx=ctx)
        bv = DenseVector(b, ctx)
        x = DenseVector(n, ctx)
        M = DiagonalPreconditioner(n, ctx)
        M.setMatrix(A)
        S = CG(x)
        S.setPreconditioner(M)
        S.setIterationMonitor(DefaultIterationMonitor(100000, rtol))
        code = 0
        try:
            S.solve(A, bv, x)
        except IterativeSolverNotConvergedException as e:
            code = e.reason.value
        xs = x.getData()
        tnow = time.perf_counter()
        results[k] = {"x": xs, "code": code, "iterations": S.getIterationMonitor().iterations()}
        lat.append(tnow - t0 - reqs[k]["arrival"])
        t_last = tnow
        for o in (S, M, x, bv, A):
            o.close()
    return _stats(lat, 0.0, t_last - t0, len(reqs)), results


def true_residual(system, x):
    n, rp, ci, va, b = system
    A = sp.csr_matrix((va, ci, rp), shape=(n, n))
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
