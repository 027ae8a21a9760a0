"""Tiled triangular sweeps: one launch per sweep over tiles of R rows by W
levels, a workgroup per tile taking tiles in dependency order and waiting on
the tiles it imports from (precond.hip: tile_levels, k_sweep_tiles).  With
tiling forced (MTJG_SWEEP_TILES=1) at several tile shapes, down to tiles of
one level, ILU apply and transApply (all four triangular forms) and ICC apply
stay bitwise the oracle's (S/ILU.java:64-78, S/ICC.java:65-75,
S/UnitLowerCompRowMatrix.java:48-86, S/UpperCompRowMatrix.java:48-90), on
repeated applies (each sweep stamps the tiles' flags with a new epoch) and
after setMatrix with new values on the same pattern."""
import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu


def _random_spd(n, density, seed):
    rng = np.random.default_rng(seed)
    d = sp.random(n, n, density=density, random_state=seed, format="csr")
    d = d + d.T + sp.diags(rng.random(n) + 20.0)
    d = sp.csr_matrix(d)
    d.sort_indices()
    return n, n, d.indptr.astype(np.int32), d.indices.astype(np.int32), d.data


def _stencil27(m, seed):
    # rows of 13 strictly-lower and 13 strictly-upper entries: grouped rows
    T = sp.diags([np.ones(m - 1), np.ones(m), np.ones(m - 1)], [-1, 0, 1])
    pat = sp.kron(sp.kron(T, T), T).tocsr()
    rng = np.random.default_rng(seed)
    off = sp.triu(pat, 1).tocsr()
    off.data = -rng.random(off.nnz) - 0.1
    off = off + off.T
    d = sp.csr_matrix(off + sp.diags(np.asarray(abs(off).sum(axis=1)).ravel() + 1.0))
    d.sort_indices()
    n = d.shape[0]
    return n, n, d.indptr.astype(np.int32), d.indices.astype(np.int32), d.data


SYSTEMS = {
    "convdiff_16": lambda: P.convdiff3d(16, peclet=10.0, seed=2),
    "convdiff_40": lambda: P.convdiff3d(40, peclet=10.0, seed=5),
    "varcoef3d_24": lambda: P.varcoef_poisson(3, 24, 20, 18, seed=3),
    "poisson2d_64": lambda: P.poisson2d(64),
    "random_spd_3000": lambda: _random_spd(3000, 0.003, 7),
    "stencil27_12": lambda: _stencil27(12, 13),
}
# (rows per tile chunk, levels per tile window); None: the default heuristic
SHAPES = [(64, 1), (700, 3), (4096, 8), (None, None)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "R%s_W%s" % s)
@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_tiled_ilu_and_icc_bitwise(ctx, monkeypatch, name, shape):
    monkeypatch.setenv("MTJG_SWEEP_TILES", "1")
    monkeypatch.setenv("MTJG_ILU_GRID", "0")  # stencil patterns would take grid sweeps
    if shape[0] is not None:
        monkeypatch.setenv("MTJG_TILE_ROWS", str(shape[0]))
        monkeypatch.setenv("MTJG_TILE_LEVELS", str(shape[1]))
    sysm = SYSTEMS[name]()
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    O = oracle.Csr(*sysm)
    ilu = M.ILU(A)
    ilu.setMatrix(A)
    Oi = oracle.Precond(oracle.PREC_ILU, O)
    for seed in (4, 5, 6):
        bh = P.uniform(n, seed)
        b = M.DenseVector(bh, ctx)
        np.testing.assert_array_equal(ilu.apply(b, M.DenseVector(n, ctx)).getData(), Oi.apply(bh))
        np.testing.assert_array_equal(ilu.transApply(b, M.DenseVector(n, ctx)).getData(), Oi.trans_apply(bh))
    # new values on the same pattern: the tiled copies are refreshed
    sysm2 = (sysm[0], sysm[1], sysm[2], sysm[3], np.asarray(sysm[4]) * 1.5)
    A2 = M.CompRowMatrix(*sysm2, ctx=ctx)
    ilu.setMatrix(A2)
    Oi2 = oracle.Precond(oracle.PREC_ILU, oracle.Csr(*sysm2))
    bh = P.uniform(n, 8)
    b = M.DenseVector(bh, ctx)
    np.testing.assert_array_equal(ilu.apply(b, M.DenseVector(n, ctx)).getData(), Oi2.apply(bh))
    if name.startswith(("varcoef", "poisson", "random_spd", "stencil")):
        icc = M.ICC(A)
        icc.setMatrix(A)
        Oc = oracle.Precond(oracle.PREC_ICC, O)
        for seed in (9, 10):
            bh = P.uniform(n, seed)
            b = M.DenseVector(bh, ctx)
            np.testing.assert_array_equal(icc.apply(b, M.DenseVector(n, ctx)).getData(), Oc.apply(bh))


def test_tiled_bicgstab_matches_untiled(ctx, monkeypatch):
    """BiCGstab + ILU(0) on 40^3 convection-diffusion: the solve with tiled
    sweeps is bitwise the solve with the untiled level schedule."""
    sysm = P.convdiff3d(40, peclet=10.0, seed=5)
    n = sysm[0]
    bh = P.uniform(n, 1)
    out = []
    monkeypatch.setenv("MTJG_ILU_GRID", "0")
    for mode in ("0", "1"):
        monkeypatch.setenv("MTJG_SWEEP_TILES", mode)
        A = M.CompRowMatrix(*sysm, ctx=ctx)
        b = M.DenseVector(bh, ctx)
        x = M.DenseVector(n, ctx)
        ilu = M.ILU(A)
        ilu.setMatrix(A)
        S = M.BiCGstab(x)
        S.setPreconditioner(ilu)
        S.setIterationMonitor(M.DefaultIterationMonitor(10000, 1e-8))
        S.solve(A, b, x)
        out.append((S.getIterationMonitor().iterations(), x.getData()))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
