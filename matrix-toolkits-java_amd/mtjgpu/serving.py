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


DEFAULT_ROWS_PER_WORKGROUP = 1024
DEFAULT_ITERS_PER_STEP = 16


class BatchServer:
    """Handle on one mtjg_batch (see include/mtjgpu.h)."""

    def __init__(self, ctx, max_slots=2048, max_workgroups=16384, arena_bytes=0,
                 rows_per_workgroup=DEFAULT_ROWS_PER_WORKGROUP, iters_per_step=DEFAULT_ITERS_PER_STEP):
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

    def stats(self):
        """Device time of the replayed graphs (ms) and steps run so far."""
        ms, st = C.c_double(), C.c_int64()
        host = (C.c_double * 8)()
        _ck(self._lib.mtjg_batch_stats(self.handle, C.byref(ms), C.byref(st), host))
        names = ("submit", "admit", "wait", "retire", "retire_copy_wait", "admit_upload_wait", "publish", "launch")
        out = {"busy_ms": ms.value, "steps": st.value}
        out.update({"host_%s_ms" % k: host[j] for j, k in enumerate(names)})
        return out

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


def batch_bytes(system, outer, inner):
    """Algorithmic HBM bytes the batch engine moves for one solved system:
    per inner FP32 CG iteration the product q = A32 p (4 m n + n on the FP32
    stencil copy of a system whose entries lie on m <= 8 diagonals: a value
    per row and diagonal and a mask byte per row; otherwise 8 nnz, the
    SELL-64 FP32 values and columns) plus 52 n (p gathered once, q written;
    d p rr q invdiag read and d rr z written; p z read and p written), per
    refinement 12 nnz + 76 n (the FP64 residual r = b - A x: values, columns,
    x and b read, r written; the restart rr = r/||r||, d = 0, z = D^-1 rr;
    p = z; x += ||r|| d), plus the final residual check."""
    n, rp, ci, va, b = system
    nnz = len(ci)
    diags = np.unique(np.asarray(ci, dtype=np.int64) - np.repeat(np.arange(n), np.diff(rp))).size
    prod32 = 4 * diags * n + n if diags <= 8 else 8 * nnz
    resid = 12 * nnz + 24 * n
    return inner * (prod32 + 52 * n) + outer * (resid + 52 * n) + resid


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
        t_step = t_collect = 0.0
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
            ts = time.perf_counter()
            B.step()
            tnow = time.perf_counter()
            t_step += tnow - ts
            for rid in B.completed():
                k = id2k.pop(rid)
                x, code, outer, inner, res = B.result(rid, systems[k][0])
                results[k] = {"x": x, "code": code, "outer": outer, "inner": inner}
                lat.append(tnow - t0 - reqs[k]["arrival"])
                t_last = tnow
            t_collect += time.perf_counter() - tnow
        st = _stats(lat, 0.0, t_last - t0, len(reqs))
        st.update(B.stats())
        st["host_step_ms"] = t_step * 1e3
        st["host_collect_ms"] = t_collect * 1e3
        nbytes = sum(batch_bytes(systems[k], r["outer"], r["inner"]) for k, r in results.items())
        st["stream_bytes"] = float(nbytes)
        st["stream_GBps_span"] = nbytes / st["span_s"] / 1e9
        st["stream_GBps_busy"] = nbytes / max(st["busy_ms"] * 1e-3, 1e-12) / 1e9
        st["busy_frac"] = st["busy_ms"] * 1e-3 / st["span_s"]
        return st, results
    finally:
        B.close()


# Launch shapes the autotuner tries: rows per workgroup (the tile each
# workgroup of a batch launch takes; multiples of 256 rows, since a
# workgroup's four wavefronts each take 64-row SELL slices, one row per lane)
# and batch iterations per replayed graph step.
TUNE_ROWS_PER_WORKGROUP = (256, 512, 1024, 2048)
TUNE_ITERS_PER_STEP = (8, 16, 32)
TUNE_MARGIN = 0.03
_tuned = {}


def shape_key(systems):
    """Autotuner key of a set of systems: the histogram of row lengths (rows
    of 1-3, 4-5, 6-7, 8-15 and 16+ stored entries, as fractions rounded to
    tenths) and the rounded log2 of the median system size."""
    edges = (1, 4, 6, 8, 16, 1 << 30)
    counts = np.zeros(len(edges) - 1)
    sizes = []
    for n, rp, _ci, _va, _b in systems:
        rl = np.diff(np.asarray(rp))
        counts += np.histogram(rl, bins=edges)[0]
        sizes.append(n)
    frac = tuple(round(float(c) / max(float(counts.sum()), 1.0), 1) for c in counts)
    return frac + (int(round(math.log2(max(np.median(sizes), 1.0)))),)


def autotune(ctx, systems, rtol=1e-8, rows_per_workgroup=TUNE_ROWS_PER_WORKGROUP,
             iters_per_step=TUNE_ITERS_PER_STEP, use_cache=True, **kw):
    """Picks the batch engine's launch shape for systems like `systems`: for
    every (rows per workgroup, iterations per step) candidate, submits all of
    them at once to a fresh batch, steps it to completion and times it; the
    candidate with the most solves per second wins, unless the engine's
    default shape is among the candidates and within TUNE_MARGIN of it (the
    sample's timings vary by a few percent from run to run, so a smaller lead
    is not a reason to leave the default).  Results are cached per
    shape_key, so a workload with the same row-length histogram and scale is
    tuned once per process.  Returns (best kwargs for BatchServer, table of
    (rows_per_workgroup, iters_per_step, solves_per_s))."""
    key = (shape_key(systems), tuple(rows_per_workgroup), tuple(iters_per_step))
    if use_cache and key in _tuned:
        return _tuned[key]
    table = []
    for rpw in rows_per_workgroup:
        for ips in iters_per_step:
            B = BatchServer(ctx, rows_per_workgroup=rpw, iters_per_step=ips, **kw)
            try:
                ids = {}
                t0 = time.perf_counter()
                for n, rp, ci, va, b in systems:
                    ids[B.submit(n, rp, ci, va, b, rtol=rtol)] = n
                while ids:
                    B.step()
                    for rid in B.completed():
                        code = B.result(rid, ids.pop(rid))[1]
                        if code != 0:
                            raise RuntimeError("autotune: a sample system ended with status %d" % code)
                dt = time.perf_counter() - t0
            finally:
                B.close()
            table.append((rpw, ips, len(systems) / dt))
    best = max(table, key=lambda r: r[2])
    default = [r for r in table if (r[0], r[1]) == (DEFAULT_ROWS_PER_WORKGROUP, DEFAULT_ITERS_PER_STEP)]
    if default and best[2] < default[0][2] * (1.0 + TUNE_MARGIN):
        best = default[0]
    out = ({"rows_per_workgroup": best[0], "iters_per_step": best[1]}, table)
    _tuned[key] = out
    return out


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
