"""Parity at the full workload sizes, through the C ABI.

Solves (S/CG.java:58-87, S/BiCGstab.java:65-116, S/GMRES.java:127-176 with
S/DiagonalPreconditioner.java and S/ILU.java): each workload is rebuilt from
its seed, b = oracle CompRowMatrix.mult(x_true) (its SHA-256 must equal the
fixture's, so both sides solve the same system), x0 = 0, and the device
solve is compared with the oracle's solve of the same system recorded in
tests/golden/fullsize_<name>.npz by tests/golden/make_fullsize_fixtures.py:

* the same verdict (converged), iteration counts within 5% (at least 1);
* true relative residual ||b - A x|| / ||b|| within 10x of the oracle's own;
* the device solution's error ||x - x_true|| / ||x_true|| at most twice the
  oracle's own, and ||x - x_oracle|| / ||x_oracle|| over the fixture's
  16,384 sampled entries at most twice the oracle's error (both solutions
  are as accurate as the oracle's and agree to that accuracy); the printed
  line also gives the whole-vector triangle bound through x_true.

Products (S/CompRowMatrix.java:294-314): the 3-D Poisson 256^3 product is
bitwise the oracle's (every row summed in storage order, under the automatic
format and the row-block CSR kernels); the power-law matrices with 4 M rows
(1-2,000 entries per row) meet the per-row bound
|y - y_oracle| <= 1e-14 * sum_j |a_ij x_j|.

Measured figures are printed (run with -s to see them)."""
import hashlib
import importlib.util
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import mtjgpu as M
import oracle
from mtjgpu import problems as P

pytestmark = pytest.mark.gpu

_spec = importlib.util.spec_from_file_location(
    "make_fullsize_fixtures", os.path.join(os.path.dirname(__file__), "golden", "make_fullsize_fixtures.py"))
FX = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(FX)


def _report(**kw):
    print("PARITY " + json.dumps(kw), flush=True)


def _system(name):
    gen, method, prec, rtol, restart = FX.WORKLOADS[name]
    fx = FX.load(name)
    sysm = gen()
    n = sysm[0]
    assert n == fx["n"] and len(sysm[4]) == fx["nnz"]
    O = oracle.Csr(*sysm)
    x_true = P.uniform(n, 1)
    b = O.mult(x_true)
    assert hashlib.sha256(b.tobytes()).hexdigest() == fx["b_sha256"]
    return sysm, O, x_true, b, fx


def _compare(name, sysm, x_true, b, xg, it_g, fx, extra=None):
    n = sysm[0]
    A = sp.csr_matrix((sysm[4], sysm[3], sysm[2]), shape=(n, n))
    true_rel = float(np.linalg.norm(b - A @ xg) / np.linalg.norm(b))
    idx, xs = fx["idx"], fx["x_samples"]
    sampled = float(np.linalg.norm(xg[idx] - xs) / np.linalg.norm(xs))
    err_g = float(np.linalg.norm(xg - x_true))
    xt = float(np.linalg.norm(x_true))
    err_o_rel = fx["err_norm"] / xt
    bound = (err_g + fx["err_norm"]) / fx["x_norm"]
    _report(workload=name, iterations=it_g, oracle_iterations=fx["iterations"], true_rel_residual=true_rel,
            oracle_true_rel_residual=fx["true_rel_residual"], rel_diff_x_sampled=sampled,
            rel_diff_x_bound=bound, rel_err_vs_x_true=err_g / xt, oracle_rel_err_vs_x_true=err_o_rel,
            **(extra or {}))
    assert true_rel <= 10 * max(fx["true_rel_residual"], fx["rtol"] * 1e-3)
    assert err_g / xt <= 2 * err_o_rel + 1e-14
    assert sampled <= max(2 * err_o_rel, 1e-12)
    return sampled


@pytest.mark.parametrize("name", list(FX.WORKLOADS))
def test_solve_matches_oracle_fixture(ctx, name):
    gen, method, prec, rtol, restart = FX.WORKLOADS[name]
    sysm, O, x_true, bh, fx = _system(name)
    del O
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    b = M.DenseVector(bh, ctx)
    x = M.DenseVector(n, ctx)
    S = M.GMRES(x, restart) if method == "gmres" else {"cg": M.CG, "bicgstab": M.BiCGstab}[method](x)
    Mp = M.DiagonalPreconditioner(n, ctx) if prec == "diag" else M.ILU(A)
    Mp.setMatrix(A)
    S.setPreconditioner(Mp)
    S.setIterationMonitor(M.DefaultIterationMonitor(100000, rtol))
    S.solve(A, b, x)
    it_g = S.getIterationMonitor().iterations()
    xg = x.getData()
    for o in (S, Mp, x, b, A):
        o.close()
    assert fx["code"] == oracle.OK
    assert abs(it_g - fx["iterations"]) <= max(1, 0.05 * fx["iterations"]), (it_g, fx["iterations"])
    _compare(name, sysm, x_true, bh, xg, it_g, fx)


@pytest.mark.parametrize("scheme", ["mixed_cg", "ir_cg32"])
def test_mixed_precision_vs_fp64_oracle(ctx, scheme):
    """The 192^3 variable-kappa system: FP32 inner Jacobi-CG (1e-4) under FP64
    refinement to 1e-10, against the oracle's FP64 Jacobi-CG to 1e-10; as the
    solver kind MixedCG and as MTJ's IR with the FP32 CG preconditioner."""
    name = "cg_varcoef_192"
    sysm, O, x_true, bh, fx = _system(name)
    del O
    n = sysm[0]
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    b = M.DenseVector(bh, ctx)
    x = M.DenseVector(n, ctx)
    if scheme == "mixed_cg":
        S = M.MixedCG(x, inner_rtol=1e-4)
        Mp = None
    else:
        S = M.IR(x)
        Mp = M.FP32CGPreconditioner(n, inner_rtol=1e-4, ctx=ctx)
        Mp.setMatrix(A)
        S.setPreconditioner(Mp)
    S.setIterationMonitor(M.DefaultIterationMonitor(100, 1e-10))
    S.solve(A, b, x)
    outer = S.getIterationMonitor().iterations()
    inner = S.inner_iterations() if Mp is None else Mp.inner_iterations()
    xg = x.getData()
    for o in (S, Mp, x, b, A):
        if o is not None:
            o.close()
    Ah = sp.csr_matrix((sysm[4], sysm[3], sysm[2]), shape=(n, n))
    assert np.linalg.norm(bh - Ah @ xg) / np.linalg.norm(bh) <= 1e-10
    _compare(scheme + "_varcoef_192", sysm, x_true, bh, xg, outer, fx,
             {"inner_iterations": inner, "oracle_is": "FP64 Jacobi-CG to 1e-10"})


@pytest.mark.parametrize("fmt", ["auto", "csr", "xcd"])
def test_spmv_poisson3d_256_bitwise(ctx, monkeypatch, fmt):
    monkeypatch.setenv("MTJG_SPMV_FORMAT", fmt)
    sysm = P.poisson3d(256)
    n = sysm[0]
    xh = P.uniform(n, 1)
    ref = oracle.Csr(*sysm).mult(xh)
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    x = M.DenseVector(xh, ctx)
    y = M.DenseVector(n, ctx)
    yg = A.mult(x, y).getData()
    used = A.format()[0]
    for o in (y, x, A):
        o.close()
    _report(workload="spmv_poisson3d_256", format=used, bitwise=bool(np.array_equal(yg, ref)))
    np.testing.assert_array_equal(yg, ref)


@pytest.mark.parametrize("gen", ["powerlaw_4m", "powerlaw_spd_4m"])
def test_spmv_powerlaw_4m_row_bound(ctx, monkeypatch, gen):
    sysm = P.powerlaw(4_000_000, seed=0) if gen == "powerlaw_4m" else P.powerlaw_spd(4_000_000, seed=0)
    n, _, rp, ci, va = sysm
    xh = P.uniform(n, 1)
    ref = oracle.Csr(*sysm).mult(xh)
    A = M.CompRowMatrix(*sysm, ctx=ctx)
    x = M.DenseVector(xh, ctx)
    y = M.DenseVector(n, ctx)
    yg = A.mult(x, y).getData()
    used = A.format()[0]
    A.close()
    # the XCD column blocks (chosen for these scattered columns) sum the same
    # products in the same order as the row-block kernel
    monkeypatch.setenv("MTJG_SPMV_FORMAT", "csr")
    Ac = M.CompRowMatrix(*sysm, ctx=ctx)
    yc = Ac.mult(x, y).getData()
    for o in (y, x, Ac):
        o.close()
    Aabs = sp.csr_matrix((np.abs(va), ci, rp), shape=(n, n))
    scale = Aabs @ np.abs(xh)
    ratio = np.abs(yg - ref) / (1e-14 * scale + 1e-300)
    _report(workload="spmv_" + gen, nnz=int(len(va)), format=used, max_err_over_bound=float(ratio.max()),
            rows_bitwise=float(np.mean(yg == ref)), bitwise_row_block_kernel=bool(np.array_equal(yg, yc)))
    assert used == "csr_xcd"
    assert np.all(ratio <= 1.0)
    np.testing.assert_array_equal(yg, yc)
