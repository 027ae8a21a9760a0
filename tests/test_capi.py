"""CPU checks of the C ABI: the library loads and exports every entry point
include/mtjgpu.h declares (no device calls), and the host-side generators
produce valid CompRowMatrix input."""
import ctypes as C
import os
import re

import numpy as np

from mtjgpu import _lib
from mtjgpu import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "mtjgpu.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(mtjg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {p[0] for p in _lib.PROTOTYPES}
    assert set(names) <= bound, set(names) - bound


def test_status_strings_without_device():
    lib = _lib.load()
    assert lib.mtjg_strerror(0) == b"ok"
    assert lib.mtjg_strerror(-3) == b"zero diagonal"
    assert lib.mtjg_strerror(2) == b"not converged: divergence"


def _valid(n, rp, ci, va, symmetric_pattern=False):
    assert rp[0] == 0 and rp[-1] == len(ci) == len(va)
    assert np.all(np.diff(rp) >= 1)
    for i in range(n):
        row = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0)
        assert i in row
    if symmetric_pattern:
        import scipy.sparse as sp
        m = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(n, n))
        assert (m != m.T).nnz == 0


def test_generators_valid_and_seeded():
    for gen in (lambda: P.poisson2d(11, 7), lambda: P.poisson3d(5, 4, 3),
                lambda: P.varcoef_poisson(3, 6, seed=4), lambda: P.convdiff3d(6, seed=9)):
        n, _, rp, ci, va = gen()
        _valid(n, rp, ci, va, symmetric_pattern=True)
    n, _, rp, ci, va = P.varcoef_poisson(2, 9, seed=3)
    import scipy.sparse as sp
    m = sp.csr_matrix((va, ci, rp), shape=(n, n))
    assert abs(m - m.T).max() == 0.0  # SPD generator is symmetric
    assert np.all(np.linalg.eigvalsh(m.toarray()) > 0)
    n, _, rp, ci, va = P.powerlaw(5000, seed=0)
    _valid(n, rp, ci, va)
    lens = np.diff(rp) - 1
    assert lens.min() >= 1 and lens.max() <= 2000
    a = P.powerlaw(3000, seed=5)
    b = P.powerlaw(3000, seed=5)
    for u, v in zip(a[2:], b[2:]):
        np.testing.assert_array_equal(u, v)
    np.testing.assert_array_equal(P.uniform(10, 1), P.uniform(10, 1))
    assert np.all(np.abs(P.uniform(1000, 2)) <= 1.0)
