"""The oracle's restatement of S/AMG.java (and of the reference-LAPACK LU
behind its coarsest level), pinned without running MTJ:

* java.util.HashSet<Integer> iteration order, which decides ties in the
  aggregate enlargement, against known answers of OpenJDK's HashMap
  (bucket = (h ^ h >>> 16) & (cap - 1), table growth at 3/4 load and at
  treeifyBin below 64 buckets, insertion order inside a bucket);
* aggregates of the 1-D Laplacian worked by hand, and a tie that the
  HashSet order (not ascending order) decides;
* every Galerkin operator equals I' A I (scipy), and every smoothed
  interpolation equals the closed form (I - omega D^-1 A_F) P_t of smoothed
  aggregation with weak couplings lumped into the diagonal;
* dgetf2 / dgetrs against scipy's LAPACK (same pivots, solutions to 1e-12);
* the V- and W-cycles against a numpy restatement of the two-grid and
  three-grid operators built from the hierarchy, SSOR's matrix-splitting
  closed form and scipy's LU of the coarsest matrix;
* MTJ's direct-solve-only regime (n <= min, as in the reference's
  CGAMGTest), errors, the stationary multigrid iteration, and the stateful
  start (apply begins from x's incoming values).
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

import oracle
from mtjgpu import problems as P
from test_oracle_ssor import _closed_form as ssor_closed_form

SMALL_MIN = (1.0, 1.85, 1.85, 1.0, 1, 1, 1, 2, 2.0 / 3)


def _ocsr(m):
    m = sp.csr_matrix(m)
    m.sort_indices()
    n = m.shape[0]
    return oracle.Csr(n, n, m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data)


def _level(g, k):
    rp, ci, va = g.matrix(k)
    n = g.levels()[k]
    return sp.csr_matrix((va, ci, rp), shape=(n, n))


def _interp(g, k):
    cp, ri, va = g.interpolation(k)
    lv = g.levels()
    return sp.csc_matrix((va, ri, cp), shape=(lv[k], lv[k + 1]))


# ---- HashSet order --------------------------------------------------------------

@pytest.mark.parametrize("vals,expected", [
    ([1, 2, 3], [1, 2, 3]),
    ([15, 16, 17], [16, 17, 15]),                 # buckets 15, 0, 1 of 16
    ([3, 19], [3, 19]),                           # one bucket: insertion order
    ([19, 3], [19, 3]),
    ([5, 21, 0, 1, 2, 3, 4, 6, 7, 8, 9, 10, 11, 12],  # 14 entries: 32 buckets
     [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 21]),
    ([5, 21, 0, 1, 2, 3, 4, 6, 7, 8, 9, 10],      # 12 entries: still 16 buckets
     [0, 1, 2, 3, 4, 5, 21, 6, 7, 8, 9, 10]),
    ([0, 16, 32, 48, 64, 80, 96, 112, 128],       # 9th entry in a bucket: treeifyBin resizes to 32
     [0, 32, 64, 96, 128, 16, 48, 80, 112]),
    ([65536, 1], [65536, 1]),                     # 65536 ^ (65536 >>> 16) = 65537: bucket 1
    ([1, 65536, 0], [0, 1, 65536]),
])
def test_java_hashset_order_known_answers(vals, expected):
    rc, order = oracle.java_hashset_order(vals)
    assert rc == oracle.OK
    assert list(order) == expected


def test_java_hashset_tree_bin_is_reported():
    vals = list(range(1, 26)) + [64 * k for k in range(9)]  # 64 buckets, 9 in bucket 0
    rc, _ = oracle.java_hashset_order(vals)
    assert rc == oracle.UNSUPPORTED


# ---- aggregation ----------------------------------------------------------------

def _chain(n):
    return sp.diags([np.full(n - 1, -1.0), np.full(n, 2.0), np.full(n - 1, -1.0)], [-1, 0, 1]).tocsr()


@pytest.mark.parametrize("n,expected", [
    (10, [0, 0, 1, 1, 1, 2, 2, 2, 3, 3]),
    (11, [0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3]),
    (12, [0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3]),   # node 11 joins by enlargement
])
def test_chain_aggregates_by_hand(n, expected):
    g = oracle.Amg(_ocsr(_chain(n)), *SMALL_MIN)
    assert g.code == oracle.OK
    assert list(g.aggregates(0)) == expected


def test_enlargement_tie_follows_hashset_order():
    """Node 16 is free after the initial aggregates {0, 15} and {1, 17}; its
    neighbourhood {15, 16, 17} iterates as 16, 17, 15 (buckets 0, 1, 15), so
    aggregate 1 reaches the maximum overlap first and wins the tie; ascending
    order would pick aggregate 0.  Nodes 2..14 have no off-diagonal entry and
    stay unaggregated."""
    n = 18
    d = np.diag(np.ones(n))
    for i, j in ((15, 16), (16, 17), (15, 0), (17, 1)):
        d[i, j] = d[j, i] = -1.0
    for i in (0, 1, 15, 16, 17):
        d[i, i] = 2.0
    g = oracle.Amg(_ocsr(d), *SMALL_MIN)
    assert g.code == oracle.OK
    pt = g.aggregates(0)
    assert list(pt[[0, 15]]) == [0, 0] and list(pt[[1, 17]]) == [1, 1]
    assert pt[16] == 1
    assert (pt[2:15] == -1).all()


def _systems():
    yield "poisson2d_32", P.poisson2d(32), ()
    yield "poisson2d_48_w_cycle", P.poisson2d(48), (1.2, 1.5, 1.5, 1.2, 2, 1, 2, 30, 0.5)
    yield "varcoef3d_10", P.varcoef_poisson(3, 10, seed=1), ()
    yield "poisson2d_40_unsmoothed", P.poisson2d(40), (1.0, 1.85, 1.85, 1.0, 1, 1, 1, 40, 0.0)
    yield "convdiff_9_sor", P.convdiff3d(9, peclet=10.0, seed=3), (1.2, 1.2, 1, 2, 1, 20, 2.0 / 3)


@pytest.mark.parametrize("name,sysm,params", list(_systems()))
def test_hierarchy_structure_and_galerkin(name, sysm, params):
    n, _, rp, ci, va = sysm
    g = oracle.Amg(oracle.Csr(n, n, rp, ci, va), *params)
    assert g.code == oracle.OK
    lv = g.levels()
    assert len(lv) >= 2 and lv[-1] <= (params[7] if len(params) == 9 else params[5] if params else 40)
    for k in range(len(lv) - 1):
        A = _level(g, k)
        I = _interp(g, k)
        pt = g.aggregates(k)
        # aggregates are disjoint, numbered densely, and cover every coupled node
        assert set(pt[pt >= 0]) == set(range(lv[k + 1]))
        off = (A - sp.diags(A.diagonal())).tocsr()
        off.eliminate_zeros()
        coupled = np.diff(off.indptr) > 0
        assert (pt[coupled] >= 0).all()
        ref = (I.T @ A @ I).toarray()
        Ac = _level(g, k + 1).toarray()
        assert np.abs(Ac - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("name,sysm,params", list(_systems()))
def test_interpolation_closed_form(name, sysm, params):
    n, _, rp, ci, va = sysm
    g = oracle.Amg(oracle.Csr(n, n, rp, ci, va), *params)
    omega = params[-1] if params else 2.0 / 3
    for k in range(len(g.levels()) - 1):
        A = _level(g, k).toarray()
        nk, c = g.levels()[k], g.levels()[k + 1]
        pt = g.aggregates(k)
        Pt = np.zeros((nk, c))
        Pt[pt >= 0, pt[pt >= 0]] = 1.0
        eps = 0.08 * 0.5 ** k
        d = np.diag(A)
        strong = np.abs(A) >= eps * np.sqrt(np.outer(d, d))
        keep = (strong | (A == 0)) & (pt[None, :] >= 0)
        AF = np.where(keep, A, 0.0)
        weak = np.where(~keep & (pt[None, :] >= 0), A, 0.0).sum(axis=1)
        AF[np.arange(nk), np.arange(nk)] -= weak
        ref = Pt - (omega / d)[:, None] * (AF @ Pt)
        ref[pt < 0, :] = 0.0
        I = _interp(g, k).toarray()
        assert np.abs(I - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())


# ---- LAPACK -----------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(5))
def test_dgetf2_dgetrs_against_scipy(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 40))
    a = rng.standard_normal((n, n))
    info, lu, piv = oracle.dgetf2(a)
    assert info == 0
    slu, spiv = sl.lu_factor(a)
    assert list(piv - 1) == list(spiv)
    assert np.abs(lu - slu).max() <= 1e-12 * np.abs(slu).max()
    b = rng.standard_normal(n)
    for trans in (False, True):
        x = oracle.dgetrs(lu, piv, b, trans)
        ref = sl.lu_solve((slu, spiv), b, trans=int(trans))
        np.testing.assert_allclose(x, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_dgetf2_singular():
    a = np.array([[1.0, 2.0, 0.0], [2.0, 4.0, 0.0], [0.0, 0.0, 0.0]])
    info, _, _ = oracle.dgetf2(a)
    assert info == 2


# ---- cycles -------------------------------------------------------------------------

def _ssor(A, b, x, wf, wr, reverse):
    return ssor_closed_form(A, b, x, wf, wr, reverse)


def _cycle_ref(g, params, k, f, u, transpose):
    """Numpy restatement of S/AMG.java:297-342 from the hierarchy."""
    wpf, wpr, wqf, wqr, nu1, nu2, gamma, _, _ = params[0]
    reverse = params[1]
    lv = g.levels()
    if k == len(lv) - 1:
        Ac = _level(g, k).toarray()
        lu = sl.lu_factor(Ac)
        return sl.lu_solve(lu, f, trans=int(transpose))
    A = _level(g, k)
    I = _interp(g, k)
    for _ in range(nu1):
        u = _ssor(A, f, u, wpf, wpr, reverse)
    r = f - A @ u
    fc = I.T @ r
    uc = np.zeros(lv[k + 1])
    for _ in range(gamma):
        uc = _cycle_ref(g, params, k + 1, fc, uc, transpose)
    u = u + I @ uc
    for _ in range(nu2):
        u = _ssor(A, f, u, wqf, wqr, reverse)
    return u


def _full_params(params):
    if not params:
        return (1.0, 1.85, 1.85, 1.0, 1, 1, 1, 40, 2.0 / 3), True
    if len(params) == 7:
        w1, w2, nu1, nu2, gm, mn, om = params
        return (w1, w1, w2, w2, nu1, nu2, gm, mn, om), False
    return tuple(params), True


@pytest.mark.parametrize("name,sysm,params", list(_systems()))
def test_cycle_against_closed_form(name, sysm, params):
    n, _, rp, ci, va = sysm
    g = oracle.Amg(oracle.Csr(n, n, rp, ci, va), *params)
    fp = _full_params(params)
    b = P.uniform(n, 5)
    x0 = P.uniform(n, 6)
    for transpose in (False, True):
        rc, x = g.apply(b, x0, transpose=transpose)
        assert rc == oracle.OK
        ref = _cycle_ref(g, fp, 0, b, x0.copy(), transpose)
        assert np.abs(x - ref).max() <= 1e-11 * np.abs(ref).max()


def test_direct_solve_only_when_small():
    """n <= min: no aggregation, apply solves A x = b by LU (the regime of the
    reference's CGAMGTest systems)."""
    rng = np.random.default_rng(3)
    n = 30
    d = rng.random((n, n)) * (rng.random((n, n)) < 0.2) + np.eye(n) * 5
    g = oracle.Amg(_ocsr(d))
    assert g.levels() == [n]
    b = rng.standard_normal(n)
    rc, x = g.apply(b, rng.standard_normal(n))
    assert rc == oracle.OK
    np.testing.assert_allclose(x, np.linalg.solve(d, b), rtol=1e-12, atol=1e-12)
    rc, x = g.apply(b, None, transpose=True)
    np.testing.assert_allclose(x, np.linalg.solve(d.T, b), rtol=1e-12, atol=1e-12)


def test_errors():
    m = sp.csr_matrix(np.array([[1.0, 1.0], [1.0, 0.0]]))
    m.eliminate_zeros()
    g = oracle.Amg(_ocsr(m), 1.0, 1.85, 1.85, 1.0, 1, 1, 1, 1, 2.0 / 3)
    assert g.code == oracle.MISSING_DIAGONAL and g.bad.value == 1 and g.bad_level.value == 0
    # a relaxation outside [0, 2] fails once SSOR levels are built ...
    sysm = P.poisson2d(10)
    g = oracle.Amg(oracle.Csr(*sysm), 1.0, 2.5, 1.85, 1.0, 1, 1, 1, 40, 2.0 / 3)
    assert g.code == oracle.BAD_ARG
    # ... and is never checked when the matrix is small enough for LU alone
    g = oracle.Amg(oracle.Csr(*P.poisson2d(6)), 1.0, 2.5, 1.85, 1.0, 1, 1, 1, 40, 2.0 / 3)
    assert g.code == oracle.OK and g.levels() == [36]
    # a singular coarsest matrix fails at apply (MatrixSingularException)
    s = np.array([[1.0, 1.0], [1.0, 1.0]])
    g = oracle.Amg(_ocsr(s))
    assert g.code == oracle.OK
    x0 = np.array([7.0, 8.0])
    rc, x = g.apply(np.ones(2), x0)
    assert rc == oracle.SINGULAR and list(x) == [7.0, 8.0]


def test_multigrid_iteration_converges():
    sysm = P.poisson2d(64)
    n = sysm[0]
    O = oracle.Csr(*sysm)
    g = oracle.Amg(O)
    b = O.mult(P.uniform(n, 1))
    x = np.zeros(n)
    for it in range(1, 40):
        _, x = g.apply(b, x)
        if np.linalg.norm(b - O.mult(x)) < 1e-8 * np.linalg.norm(b):
            break
    assert it <= 16


def test_apply_starts_from_x():
    """apply(b, x) cycles from x's incoming values (S/AMG.java:207-214), so
    inside MTJ's CG (which passes the previous z) it is not a fixed operator:
    CG with it stagnates on the 2-D Poisson matrix, in MTJ as here, while a
    zero start each time is an effective preconditioner."""
    sysm = P.poisson2d(32)
    n = sysm[0]
    O = oracle.Csr(*sysm)
    g = oracle.Amg(O)
    b = P.uniform(n, 2)
    _, x1 = g.apply(b, np.zeros(n))
    _, x2 = g.apply(b, np.ones(n))
    assert np.abs(x1 - x2).max() > 1e-3
    bb = O.mult(P.uniform(n, 1))
    code, _, _, _ = oracle.solve("cg", O, bb, np.zeros(n), g.precond(), oracle.monitor(500, 1e-8))
    assert code == oracle.ITERATIONS


@pytest.mark.parametrize("seed", range(5))
def test_cg_with_amg_reference_style(seed):
    """CGAMGTest's systems (n < 50 <= min + 10): LU on the whole matrix."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 41))
    d = np.zeros((n, n))
    for i in range(n):
        cols = rng.choice(n, size=min(n, 5), replace=False)
        d[i, cols] = rng.random(len(cols))
    d = d + d.T + np.eye(n) * 100.0
    O = _ocsr(d)
    g = oracle.Amg(O)
    xt = rng.random(n)
    code, x, it, _ = oracle.solve("cg", O, d @ xt, rng.random(n), g.precond(), oracle.monitor(100000, 1e-5))
    assert code == oracle.OK
    np.testing.assert_allclose(x, xt, atol=1e-4, rtol=0)
