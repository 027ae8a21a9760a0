"""ILU(0), ICC and SSOR grid sweeps (precond.hip: detect_grid, k_grid_sweep):
when the pattern is a natural-order 5- or 7-point grid stencil, the
triangular solves of ILU's apply and transApply (S/ILU.java:64-78,
S/UnitLowerCompRowMatrix.java:48-86, S/UpperCompRowMatrix.java:48-90),
ICC's apply (S/ICC.java:65-71) and SSOR's two sweeps (S/SSOR.java:143-194)
run as sweeps over the grid's anti-diagonals, plane groups pipelined across
workgroups.  The results must
be bitwise the oracle's, which solves row by row in MTJ's order; that covers
grids of one and several 64-row blocks, partial last plane groups, one plane
per workgroup (every plane imported across workgroups), no head start
between workgroups (imports found pending and polled), 2-D grids, and rows
with missing neighbours.  Patterns that are not such a stencil keep the
level schedule."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import _lib as L
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu


def _dropped_stencil(nx, ny, nz, drop, seed):
    """7-point (nz > 1) or 5-point pattern in natural order with each
    off-diagonal neighbour kept with probability 1 - drop (both triangles
    independently), diagonally dominant values."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    i = np.arange(n)
    x, y, z = i % nx, (i // nx) % ny, i // (nx * ny)
    rows, cols = [i], [i]
    for d, ok in ((1, x < nx - 1), (nx, y < ny - 1), (nx * ny, z < nz - 1)):
        for sign in (-1, 1):
            src = i[ok] if sign > 0 else i[ok] + d
            keep = rng.random(src.size) >= drop
            rows.append(src[keep])
            cols.append(src[keep] + sign * d)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = -rng.random(r.size) - 0.05
    v[:n] = 0.0
    m = sp.csr_matrix((v, (r, c)), shape=(n, n))
    m = m + sp.diags(np.asarray(abs(m).sum(axis=1)).ravel() + 0.5)
    m = sp.csr_matrix(m)
    m.sort_indices()
    return n, n, m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data


CASES = {
    # name: (system, grid, environment)
    "convdiff_16": (lambda: P.convdiff3d(16, seed=2), (16, 16, 16), {}),
    "convdiff_20x70x9": (lambda: P.convdiff3d(20, 70, 9, seed=3), (20, 70, 9), {}),
    "convdiff_9x130x11": (lambda: P.convdiff3d(9, 130, 11, seed=4), (9, 130, 11), {}),
    "varcoef_8x8x100": (lambda: P.varcoef_poisson(3, 8, 8, 100, seed=5), (8, 8, 100), {}),
    "varcoef_8x8x100_one_plane": (lambda: P.varcoef_poisson(3, 8, 8, 100, seed=5), (8, 8, 100),
                                  {"MTJG_ILU_GRID_PLANES": "1"}),
    "convdiff_12x40x30_no_lead": (lambda: P.convdiff3d(12, 40, 30, seed=6), (12, 40, 30),
                                  {"MTJG_ILU_GRID_PLANES": "2", "MTJG_ILU_GRID_LEAD": "0"}),
    "poisson2d_64": (lambda: P.poisson2d(64), (64, 64, 1), {}),
    "poisson2d_300x1000": (lambda: P.poisson2d(300, 1000), (300, 1000, 1), {}),
    "poisson3d_2": (lambda: P.poisson3d(2), (2, 2, 2), {}),
    "dropped_3d": (lambda: _dropped_stencil(11, 67, 13, 0.2, 7), (11, 67, 13), {}),
    "dropped_2d": (lambda: _dropped_stencil(37, 150, 1, 0.3, 8), (37, 150, 1), {}),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_grid_sweeps_bitwise_oracle(ctx, monkeypatch, name):
    make, grid, env = CASES[name]
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sysm = make()
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    ilu = M.ILU(A)
    ilu.setMatrix(A)
    assert ilu.grid() == grid
    Oi = oracle.Precond(oracle.PREC_ILU, oracle.Csr(*sysm))
    assert Oi.code == oracle.OK
    for seed in (4, 5, 6):
        bh = P.uniform(n, seed)
        b = M.DenseVector(bh, ctx)
        np.testing.assert_array_equal(ilu.apply(b, M.DenseVector(n, ctx)).getData(), Oi.apply(bh))
        np.testing.assert_array_equal(ilu.transApply(b, M.DenseVector(n, ctx)).getData(), Oi.trans_apply(bh))
    # new values on the same pattern: the skewed copies of the factor follow
    sysm2 = (sysm[0], sysm[1], sysm[2], sysm[3], np.asarray(sysm[4]) * 1.5)
    ilu.setMatrix(M.CompRowMatrix(*sysm2, ctx=ctx))
    Oi2 = oracle.Precond(oracle.PREC_ILU, oracle.Csr(*sysm2))
    bh = P.uniform(n, 9)
    b = M.DenseVector(bh, ctx)
    np.testing.assert_array_equal(ilu.apply(b, M.DenseVector(n, ctx)).getData(), Oi2.apply(bh))
    np.testing.assert_array_equal(ilu.transApply(b, M.DenseVector(n, ctx)).getData(), Oi2.trans_apply(bh))


ICC_CASES = {
    "varcoef_20x70x9": (lambda: P.varcoef_poisson(3, 20, 70, 9, seed=3), (20, 70, 9), {}),
    "varcoef_8x8x100_one_plane": (lambda: P.varcoef_poisson(3, 8, 8, 100, seed=5), (8, 8, 100),
                                  {"MTJG_ILU_GRID_PLANES": "1"}),
    "poisson3d_12x40x30_no_lead": (lambda: P.poisson3d(12, 40, 30), (12, 40, 30), {"MTJG_ILU_GRID_LEAD": "0"}),
    "poisson2d_200x400": (lambda: P.poisson2d(200, 400), (200, 400, 1), {}),
    "varcoef2d_130x77": (lambda: P.varcoef_poisson(2, 130, 77, seed=9), (130, 77, 1), {}),
}


@pytest.mark.parametrize("name", sorted(ICC_CASES))
def test_icc_grid_sweeps_bitwise_oracle(ctx, monkeypatch, name):
    """ICC's apply (R'y = b, then R x = y) as grid sweeps, bitwise the
    oracle's, also after setMatrix with new values on the same pattern."""
    make, grid, env = ICC_CASES[name]
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sysm = make()
    n = sysm[0]
    icc = M.ICC(M.CompRowMatrix(*sysm, ctx=ctx))
    icc.setMatrix(M.CompRowMatrix(*sysm, ctx=ctx))
    assert icc.grid() == grid
    Oc = oracle.Precond(oracle.PREC_ICC, oracle.Csr(*sysm))
    assert Oc.code == oracle.OK
    for seed in (4, 5):
        bh = P.uniform(n, seed)
        np.testing.assert_array_equal(icc.apply(M.DenseVector(bh, ctx), M.DenseVector(n, ctx)).getData(),
                                      Oc.apply(bh))
    sysm2 = (sysm[0], sysm[1], sysm[2], sysm[3], np.asarray(sysm[4]) * 1.5)
    icc.setMatrix(M.CompRowMatrix(*sysm2, ctx=ctx))
    Oc2 = oracle.Precond(oracle.PREC_ICC, oracle.Csr(*sysm2))
    bh = P.uniform(n, 9)
    np.testing.assert_array_equal(icc.apply(M.DenseVector(bh, ctx), M.DenseVector(n, ctx)).getData(),
                                  Oc2.apply(bh))


SSOR_CASES = ["convdiff_20x70x9", "convdiff_9x130x11", "varcoef_8x8x100_one_plane", "convdiff_12x40x30_no_lead",
              "poisson2d_300x1000", "dropped_3d", "dropped_2d"]


@pytest.mark.parametrize("wf,wr,reverse", [(1.0, 1.0, True), (1.3, 0.7, True), (1.7, 1.7, False)])
@pytest.mark.parametrize("name", SSOR_CASES)
def test_ssor_grid_sweeps_bitwise_oracle(ctx, monkeypatch, name, wf, wr, reverse):
    """SSOR's forward and (reverse) backward sweeps as grid sweeps, bitwise
    the oracle's from a zero and a nonzero incoming x, and after setMatrix
    with new values on the same pattern."""
    make, grid, env = CASES[name]
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sysm = make()
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    S = M.SSOR(A, reverse=reverse, omegaF=wf, omegaR=wr)
    S.setMatrix(A)
    assert S.grid() == grid
    Op = oracle.Precond(oracle.PREC_SSOR, oracle.Csr(*sysm), reverse=reverse, omega_f=wf, omega_r=wr)
    b = P.uniform(n, 7)
    for x0 in (np.zeros(n), P.uniform(n, 8)):
        x = M.DenseVector(x0, ctx)
        S.apply(M.DenseVector(b, ctx), x)
        assert x.getData().tobytes() == Op.apply(b, x0).tobytes()
    sysm2 = (sysm[0], sysm[1], sysm[2], sysm[3], np.asarray(sysm[4]) * 1.5)
    S.setMatrix(M.CompRowMatrix(*sysm2, ctx=ctx))
    Op2 = oracle.Precond(oracle.PREC_SSOR, oracle.Csr(*sysm2), reverse=reverse, omega_f=wf, omega_r=wr)
    x0 = P.uniform(n, 9)
    x = M.DenseVector(x0, ctx)
    S.apply(M.DenseVector(b, ctx), x)
    assert x.getData().tobytes() == Op2.apply(b, x0).tobytes()


def _not_grid_cases():
    m = 12
    T = sp.diags([np.ones(m - 1), np.ones(m), np.ones(m - 1)], [-1, 0, 1])
    s27 = sp.csr_matrix(sp.kron(sp.kron(T, T), T) + 26 * sp.eye(m ** 3))
    # 2-D Laplacian with periodic x: the x-neighbour of x = 0 wraps to the line's end
    nx, ny = 10, 12
    I1 = sp.eye(ny)
    Tx = sp.diags([-np.ones(nx - 1), 4 * np.ones(nx), -np.ones(nx - 1)], [-1, 0, 1]).tolil()
    Tx[0, nx - 1] = -1.0
    Tx[nx - 1, 0] = -1.0
    Ty = sp.diags([-np.ones(ny - 1), -np.ones(ny - 1)], [-1, 1])
    per = sp.csr_matrix(sp.kron(I1, Tx) + sp.kron(Ty, sp.eye(nx)))
    tri = sp.csr_matrix(sp.diags([-np.ones(499), 3 * np.ones(500), -np.ones(499)], [-1, 0, 1]))
    rng = np.random.default_rng(3)
    rnd = sp.random(300, 300, density=0.02, random_state=3, format="csr")
    rnd = sp.csr_matrix(rnd + rnd.T + sp.diags(rng.random(300) + 10.0))
    return {"stencil27": s27, "periodic_x": per, "tridiagonal": tri, "random": rnd,
            "tall_2d": sp.csr_matrix(sp.kron(sp.diags([-np.ones(1199), 4 * np.ones(1200), -np.ones(1199)],
                                                      [-1, 0, 1]), sp.eye(4))
                                     + sp.kron(sp.eye(1200), sp.diags([-np.ones(3), -np.ones(3)], [-1, 1])))}


@pytest.mark.parametrize("name", ["stencil27", "periodic_x", "tridiagonal", "random", "tall_2d"])
def test_not_grid_patterns_keep_levels(ctx, name):
    m = _not_grid_cases()[name]
    m.sort_indices()
    n = m.shape[0]
    sysm = (n, n, m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data)
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    ilu = M.ILU(A)
    ilu.setMatrix(A)
    assert ilu.grid() is None
    Oi = oracle.Precond(oracle.PREC_ILU, oracle.Csr(*sysm))
    bh = P.uniform(n, 4)
    np.testing.assert_array_equal(ilu.apply(M.DenseVector(bh, ctx), M.DenseVector(n, ctx)).getData(),
                                  Oi.apply(bh))


def test_poll_budget_error_path(ctx, monkeypatch):
    """A poll budget of 0, set from the host, admits no wait: every plane
    group that imports raises the error flag at its start, and the sweeps
    still run to their end.  apply then fails with MTJG_E_DEVICE (and clears
    the flag); a preconditioner with the default budget is bitwise the
    oracle's afterwards.  No kernel waits longer than its own budget."""
    sysm = P.convdiff3d(12, 40, 30, seed=6)
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    bh = P.uniform(n, 4)
    monkeypatch.setenv("MTJG_ILU_GRID_POLL_BUDGET", "0")
    for make in (lambda: M.ILU(A), lambda: M.SSOR(A)):
        pc = make()
        pc.setMatrix(A)
        assert pc.grid() == (12, 40, 30)
        with pytest.raises(M.MtjgError) as e:
            pc.apply(M.DenseVector(bh, ctx), M.DenseVector(n, ctx))
        assert e.value.code == L.E_DEVICE
    monkeypatch.delenv("MTJG_ILU_GRID_POLL_BUDGET")
    ilu = M.ILU(A)
    ilu.setMatrix(A)
    Oi = oracle.Precond(oracle.PREC_ILU, oracle.Csr(*sysm))
    np.testing.assert_array_equal(ilu.apply(M.DenseVector(bh, ctx), M.DenseVector(n, ctx)).getData(), Oi.apply(bh))


def test_grid_switch_off(ctx, monkeypatch):
    monkeypatch.setenv("MTJG_ILU_GRID", "0")
    sysm = P.convdiff3d(16, seed=2)
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    ilu = M.ILU(A)
    ilu.setMatrix(A)
    assert ilu.grid() is None


@pytest.mark.parametrize("method", ["bicgstab", "gmres", "bicg", "qmr", "cg_icc", "cg_ssor"])
def test_solvers_same_bits_as_level_schedule(ctx, monkeypatch, method):
    """BiCGstab, GMRES(30), BiCG and QMR (these two through transApply) with
    ILU on 3-D convection-diffusion 24^3, and CG with ICC or SSOR on
    variable-kappa Poisson 24^3: grid sweeps and the level schedule give the
    same iterates.  (SSOR's apply starts from x's incoming values, so inside
    CG it is not a fixed operator and CG need not converge: that case runs 30
    iterations and compares the iterates.)"""
    spd = method in ("cg_icc", "cg_ssor")
    sysm = P.varcoef_poisson(3, 24, seed=2) if spd else P.convdiff3d(24, peclet=10.0, seed=1)
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    bh = A.mult(M.DenseVector(P.uniform(n, 1), ctx), M.DenseVector(n, ctx)).getData()
    out = []
    for grid in ("1", "0"):
        monkeypatch.setenv("MTJG_ILU_GRID", grid)
        ilu = M.ICC(A) if method == "cg_icc" else M.SSOR(A) if method == "cg_ssor" else M.ILU(A)
        ilu.setMatrix(A)
        assert (ilu.grid() is not None) == (grid == "1")
        x = M.DenseVector(n, ctx)
        S = {"gmres": lambda: M.GMRES(x, 30), "bicgstab": lambda: M.BiCGstab(x), "bicg": lambda: M.BiCG(x),
             "qmr": lambda: M.QMR(x), "cg_icc": lambda: M.CG(x), "cg_ssor": lambda: M.CG(x)}[method]()
        S.setPreconditioner(ilu)
        if method == "cg_ssor":
            S.setIterationMonitor(M.DefaultIterationMonitor(30, 1e-30))
            with pytest.raises(M.IterativeSolverNotConvergedException):
                S.solve(A, M.DenseVector(bh, ctx), x)
        else:
            S.setIterationMonitor(M.DefaultIterationMonitor(10000, 1e-10))
            S.solve(A, M.DenseVector(bh, ctx), x)
        out.append((S.getIterationMonitor().iterations(), x.getData()))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
