"""GPU tests of the batched, continuously admitted mixed-precision CG engine
(mtjg_batch_*, the serving path).

Bars: every system reaches the requested FP64 relative residual
||b - A x|| / ||b|| <= 1e-8 (the outer monitor is relative to ||b|| with
x0 = 0); a system's solution is bitwise the same alone and in a mix of other
systems; requests that do not fit the memory budget wait and are solved
later; a broken system (NaN right-hand side) ends in divergence without
affecting the others.  The FP64 Jacobi-CG oracle solution is the accuracy
reference (same tolerance)."""
import numpy as np
import pytest

import mtjgpu as M
import oracle
from mtjgpu import problems as P
from mtjgpu import serving as S

pytestmark = pytest.mark.gpu


def _systems(count, seed=3, n_max=40000):
    reqs = S.request_stream(count, n_min=1000, n_max=n_max, seed=seed)
    return reqs, [S.make_system(r) for r in reqs]


def _run(ctx, systems, **kw):
    B = S.BatchServer(ctx, **kw)
    try:
        ids = {B.submit(*sysm[:4], sysm[4], rtol=1e-8): k for k, sysm in enumerate(systems)}
        out = {}
        for _ in range(100000):
            B.step()
            for rid in B.completed():
                k = ids[rid]
                out[k] = B.result(rid, systems[k][0])
            if len(out) == len(systems):
                break
        return out
    finally:
        B.close()


def test_batch_reaches_fp64_residual(ctx):
    reqs, systems = _systems(12)
    out = _run(ctx, systems, max_slots=64, max_workgroups=4096, arena_bytes=1 << 30)
    assert len(out) == len(systems)
    for k, (x, code, outer, inner, res) in out.items():
        assert code == 0, (k, code)
        assert S.true_residual(systems[k], x) <= 1e-8
        assert inner > 0 and outer >= 1


def test_same_alone_and_in_mix(ctx):
    reqs, systems = _systems(10, seed=11)
    mix = _run(ctx, systems, max_slots=64, max_workgroups=4096, arena_bytes=1 << 30)
    for k in (0, 4, 9):
        alone = _run(ctx, [systems[k]], max_slots=4, max_workgroups=4096, arena_bytes=1 << 30)
        np.testing.assert_array_equal(alone[0][0], mix[k][0])
        assert alone[0][2:4] == mix[k][2:4]


def test_matches_fp64_oracle_solution(ctx):
    reqs, systems = _systems(3, seed=5, n_max=6000)
    out = _run(ctx, systems, max_slots=8, max_workgroups=512, arena_bytes=1 << 28)
    for k, sysm in enumerate(systems):
        n, rp, ci, va, b = sysm
        O = oracle.Csr(n, n, rp, ci, va)
        code, xo, it, _ = oracle.solve("cg", O, b, np.zeros(n), oracle.Precond(oracle.PREC_DIAGONAL, O),
                                       oracle.monitor(100000, 1e-10))
        assert code == 0
        xg = out[k][0]
        assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-5


def test_admission_waits_for_memory(ctx):
    reqs, systems = _systems(8, seed=21, n_max=20000)
    need = max(S_bytes(s) for s in systems)
    # room for about two systems at a time: the rest queue and are admitted
    # as earlier ones retire
    out = _run(ctx, systems, max_slots=64, max_workgroups=4096, arena_bytes=int(2.2 * need))
    assert len(out) == len(systems)
    for k, (x, code, *_r) in out.items():
        assert code == 0
        assert S.true_residual(systems[k], x) <= 1e-8


def S_bytes(sysm):
    """Arena bytes of one admitted system: CSR, FP32 inverse diagonal, FP64
    b x r, FP32 rr d p q z, the SELL-64 copy (slice offsets, columns, FP64
    values, each slice as wide as its longest row) and the FP32 values: the
    stencil copy (a value per row and diagonal, a mask byte per row) when the
    entries lie on at most 8 diagonals, else the SELL-64 FP32 values."""
    n, rp, ci, va, b = sysm
    nnz = len(ci)
    al = lambda v: (v + 255) // 256 * 256
    lens = np.diff(rp)
    nslice = (n + 63) // 64
    sell = sum(int(lens[64 * s:64 * s + 64].max()) * 64 for s in range(nslice))
    diags = len(np.unique(ci - np.repeat(np.arange(n), lens)))
    fp32 = al(4 * diags * n) + al(n) if diags <= 8 else al(4 * sell)
    return (al(4 * (n + 1)) + al(4 * nnz) + al(8 * nnz) + al(4 * n) + 3 * al(8 * n) + 5 * al(4 * n)
            + al(4 * (nslice + 1)) + al(4 * sell) + al(8 * sell) + fp32)


def test_nan_system_is_isolated(ctx):
    reqs, systems = _systems(4, seed=31, n_max=8000)
    n, rp, ci, va, b = systems[2]
    bad = b.copy()
    bad[n // 2] = np.nan
    systems[2] = (n, rp, ci, va, bad)
    out = _run(ctx, systems, max_slots=16, max_workgroups=1024, arena_bytes=1 << 28)
    assert out[2][1] == M.Reason.Divergence.value
    for k in (0, 1, 3):
        assert out[k][1] == 0 and S.true_residual(systems[k], out[k][0]) <= 1e-8


def _plant(sysm, defects):
    n, _, rp, ci, va = sysm
    rp, ci, va = rp.copy(), ci.copy(), va.copy()
    for kind, row in defects:
        lo, hi = rp[row], rp[row + 1]
        d = lo + int(np.flatnonzero(ci[lo:hi] == row)[0])
        if kind == "column":
            ci[hi - 1] = n + 5
        elif kind == "zero":
            va[d] = 0.0
        elif kind == "missing":
            ci[d] = (row + 2) % n
        elif kind == "pointer":
            rp[row + 1] = rp[row] - 1
    return n, rp, ci, va


@pytest.mark.parametrize("grid", [20, 300])
def test_submit_checks_structure(ctx, grid):
    """Submit rejects a broken CSR structure with the error of its lowest
    failing row (checked serially for the small system, in parallel parts of
    rows for the 90,000-row one), and the batch then serves a valid system."""
    sysm = P.poisson2d(grid)
    n = sysm[0]
    lo, mid, hi = n // 7, n // 3, n - 3
    cases = [([("column", hi), ("zero", mid)], M.ZeroDiagonalException, mid),
             ([("missing", lo), ("column", hi)], M.MissingDiagonalException, lo),
             ([("column", mid), ("missing", hi)], M.IllegalArgumentException, mid),
             ([("zero", hi), ("pointer", lo)], M.IllegalArgumentException, None)]
    b = P.uniform(n, 2)
    B = S.BatchServer(ctx, max_slots=8, max_workgroups=1024)
    try:
        for defects, exc, row in cases:
            with pytest.raises(exc) as e:
                B.submit(*_plant(sysm, defects), b, rtol=1e-8)
            if row is not None:
                assert str(e.value).endswith("row %d" % row), str(e.value)
        assert B.info()["queued"] == 0
        rid = B.submit(n, sysm[2], sysm[3], sysm[4], b, rtol=1e-8)
        out = None
        for _ in range(100000):
            B.step()
            if rid in B.completed():
                out = B.result(rid, n)
                break
        assert out is not None and out[1] == 0
        assert S.true_residual((n, sysm[2], sysm[3], sysm[4], b), out[0]) <= 1e-8
    finally:
        B.close()


def test_persistent_step_matches_graph(ctx, monkeypatch):
    """MTJG_BATCH_PERSISTENT=1 runs each step as one cooperative launch with
    grid barriers instead of the captured graph of three launches per
    iteration: the same systems end with bitwise the same solutions and
    counts (every per-system sum adds its tiles' partials in tile order,
    whichever workgroup produced them)."""
    reqs, systems = _systems(10, seed=23, n_max=60000)
    outs = []
    for mode in ("0", "1"):
        monkeypatch.setenv("MTJG_BATCH_PERSISTENT", mode)
        outs.append(_run(ctx, systems))
    for k in range(len(systems)):
        g, p = outs[0][k], outs[1][k]
        assert g[1] == 0 and p[1] == 0
        assert g[2:4] == p[2:4], (k, g[2:4], p[2:4])
        np.testing.assert_array_equal(g[0], p[0])


def test_hbm_admitted_arena(ctx):
    """arena_bytes = 0: the arena starts empty and grows in chunks of device
    memory as systems are admitted, bounded by the free HBM hipMemGetInfo
    reports; the chunks are kept until the batch is destroyed."""
    reqs, systems = _systems(6, seed=51, n_max=30000)
    B = S.BatchServer(ctx, max_slots=16, max_workgroups=1024, arena_bytes=0)
    try:
        assert B.info()["arena_size"] == 0
        ids = {B.submit(*sysm[:4], sysm[4], rtol=1e-8): k for k, sysm in enumerate(systems)}
        B.step()
        info = B.info()
        assert info["arena_size"] >= 4 << 30 and info["arena_used"] >= sum(S_bytes(s) for s in systems)
        out = {}
        for _ in range(100000):
            B.step()
            for rid in B.completed():
                k = ids[rid]
                out[k] = B.result(rid, systems[k][0])
            if len(out) == len(systems):
                break
        info = B.info()
        assert info["arena_used"] == 0 and info["arena_size"] >= 4 << 30
        for k, (x, code, *_r) in out.items():
            assert code == 0 and S.true_residual(systems[k], x) <= 1e-8
    finally:
        B.close()


def test_serving_loops_small_stream(ctx):
    reqs, systems = _systems(16, seed=41, n_max=20000)
    st_b, res_b = S.serve_batched(ctx, reqs, systems, max_slots=64, max_workgroups=4096)
    st_p, res_p = S.serve_plain(ctx, reqs, systems)
    assert st_b["solves"] == st_p["solves"] == 16
    for k in range(16):
        assert res_b[k]["code"] == 0 and res_p[k]["code"] == 0
        assert S.true_residual(systems[k], res_b[k]["x"]) <= 1e-8
        assert S.true_residual(systems[k], res_p[k]["x"]) <= 1e-8


def test_autotune_picks_a_measured_shape(ctx):
    """The autotuner times every candidate launch shape on a sample, returns
    the fastest as BatchServer arguments (the default shape unless beaten by
    more than TUNE_MARGIN), caches it per row-length histogram, and the
    chosen shape serves the stream to 1e-8."""
    reqs = S.request_stream(6, n_max=20000, seed=31)
    systems = [S.make_system(r) for r in reqs]
    best, table = S.autotune(ctx, systems, rows_per_workgroup=(256, 1024), iters_per_step=(8, 16),
                             max_slots=16, max_workgroups=1024, arena_bytes=1 << 30, use_cache=False)
    assert sorted((r, i) for r, i, _ in table) == [(256, 8), (256, 16), (1024, 8), (1024, 16)]
    assert all(v > 0 for _, _, v in table)
    top = max(table, key=lambda t: t[2])
    default = [t for t in table if (t[0], t[1]) == (S.DEFAULT_ROWS_PER_WORKGROUP, S.DEFAULT_ITERS_PER_STEP)][0]
    if top[2] < default[2] * (1 + S.TUNE_MARGIN):
        assert best == {"rows_per_workgroup": default[0], "iters_per_step": default[1]}
    else:
        assert best == {"rows_per_workgroup": top[0], "iters_per_step": top[1]}
    again = S.autotune(ctx, systems, rows_per_workgroup=(256, 1024), iters_per_step=(8, 16),
                       max_slots=16, max_workgroups=1024, arena_bytes=1 << 30)
    assert S.autotune(ctx, systems, rows_per_workgroup=(256, 1024), iters_per_step=(8, 16)) is again
    st, res = S.serve_batched(ctx, reqs, systems, max_slots=16, max_workgroups=1024, arena_bytes=1 << 30,
                              **again[0])
    assert st["solves"] == len(reqs)
    for k in range(len(reqs)):
        assert res[k]["code"] == 0
        assert S.true_residual(systems[k], res[k]["x"]) < 1e-8


def test_stencil_copy_same_bits_as_sell(ctx, monkeypatch):
    """The engine's FP32 inner products read each system's stencil copy
    (values per diagonal, the stream's 2-D and 3-D Poisson systems have 5 and
    7); with MTJG_SPMV_STENCIL=0 at submission they read the SELL-64 FP32
    values instead.  Both add a row's products in column order: the
    solutions and iteration counts are bitwise equal."""
    reqs, systems = _systems(6, seed=17, n_max=30000)
    with_st = _run(ctx, systems, max_slots=16, max_workgroups=2048, arena_bytes=1 << 29)
    monkeypatch.setenv("MTJG_SPMV_STENCIL", "0")
    without = _run(ctx, systems, max_slots=16, max_workgroups=2048, arena_bytes=1 << 29)
    for k in range(len(systems)):
        np.testing.assert_array_equal(with_st[k][0], without[k][0])
        assert with_st[k][1:4] == without[k][1:4]
