"""The library's host ICC factorization (csrc/icc_factor.hpp, run by
ICC.setMatrix before the factor is uploaded) against the oracle's literal
O(n^2) restatement of S/ICC.java:90-136, bit for bit, on the CPU: meshes,
random SPD, indefinite (NaN pivots that MTJ spreads to later rows) and
nonsymmetric patterns, power-law rows, and the error rows.  The header is
compiled here with g++ and -ffp-contract=off, as the library is."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from mtjgpu import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "matrix-toolkits-java_amd", "csrc")


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("icc") / "libiccshim.so")
    subprocess.check_call(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC", "-I", CSRC,
                           "-o", so, os.path.join(HERE, "icc_factor_shim.cpp")])
    return C.CDLL(so)


def _host(lib, n, rp, ci, va):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.array(va, dtype=np.float64)
    diag = np.array([rp[k] + np.searchsorted(ci[rp[k]:rp[k + 1]], k) for k in range(n)], np.int32)
    bad = C.c_int32(-1)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = lib.icc_factor_shim(n, p(rp), p(ci), p(diag), p(v), C.byref(bad))
    return rc, bad.value, v


def _rnd(n, dens, seed, shift, sym=True):
    rng = np.random.default_rng(seed)
    d = sp.random(n, n, density=dens, random_state=seed, format="csr")
    d.data = rng.standard_normal(d.nnz)
    if sym:
        d = d + d.T
    d = sp.csr_matrix(d + sp.diags(rng.random(n) + shift))
    d.sort_indices()
    return n, n, d.indptr.astype(np.int32), d.indices.astype(np.int32), d.data


def _powerlaw_sym():
    n, _, rp, ci, va = P.powerlaw(2000, seed=4)
    B = sp.csr_matrix((va, ci, rp), shape=(n, n))
    S = sp.csr_matrix(B + B.T)
    S.sort_indices()
    return n, n, S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data


def _nan_pivot():
    n = 6
    d = np.eye(n) * 4.0
    d[0, 0] = 1.0
    d[0, 1] = d[1, 0] = 3.0
    d[1, 5] = d[5, 1] = 1.0
    d[3, 5] = d[5, 3] = 0.5
    m = sp.csr_matrix(d)
    return n, n, m.indptr, m.indices, m.data


CASES = {
    "p2d_64": lambda: P.poisson2d(64),
    "varcoef3d_12": lambda: P.varcoef_poisson(3, 12, 10, 9, seed=2),
    "convdiff_10": lambda: P.convdiff3d(10, peclet=10.0, seed=3),
    "powerlaw_sym_2000": _powerlaw_sym,
    "nan_pivot": _nan_pivot,
    "zero_diagonal_row1": lambda: (2, 2, np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.ones(4)),
    "zero_diagonal_row0": lambda: (2, 2, np.array([0, 2, 4]), np.array([0, 1, 0, 1]),
                                   np.array([0.0, 1.0, 1.0, 1.0])),
    "empty": lambda: (0, 0, np.array([0]), np.array([], np.int32), np.array([])),
}
for _s in range(4):
    CASES["random_spd_%d" % _s] = (lambda s=_s: _rnd(600, 0.006, s, 8.0))
    CASES["indefinite_%d" % _s] = (lambda s=_s: _rnd(300, 0.015, s, 0.5))
    CASES["nonsymmetric_%d" % _s] = (lambda s=_s: _rnd(300, 0.02, s, 1.0, sym=False))


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_factor_equals_oracle_bitwise(shim, name):
    n, _, rp, ci, va = CASES[name]()
    rc, bad, v = _host(shim, n, rp, ci, va)
    Mo = oracle.Precond(oracle.PREC_ICC, oracle.Csr(n, n, rp, ci, va))
    assert {0: oracle.OK, 1: oracle.ZERO_PIVOT, 2: oracle.ZERO_DIAGONAL}[rc] == Mo.code
    if rc:
        assert bad == Mo.bad.value
        return
    assert v.tobytes() == Mo.ilu_values()[:len(v)].tobytes()
    if name.startswith("indefinite"):
        assert np.isnan(v).any()  # the NaN-spreading path is exercised
