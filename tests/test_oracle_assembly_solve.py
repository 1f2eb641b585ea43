"""Pins the oracle's element matrices, right-hand sides, Dirichlet elimination
and Jacobi-CG against closed forms and invariants.

Mirrors testsuite/feelvf/test_forms.cpp:135-139 (energy of 1 = measure),
testsuite/feelalg/test_matrix.cpp:219 (energy(1,1) = k*measure),
python/pyfeelpp/tests/test_forms.py:8-53 (mass(1,1)=measure, Dirichlet g=x
reproduces u=x), plus P1 closed forms, quadratic-field energies, rigid-body
modes and scipy.sparse direct solves.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import dof_positions


def space(orc, dim, order, ncomp=1, n=4, jitter=0.15, seed=11):
    if dim == 2:
        coords, elems = orc.mesh_square(n, jitter=jitter, seed=seed)
    else:
        coords, elems = orc.mesh_kuhn(n, jitter=jitter, seed=seed)
    e2d, ndof = orc.doftable(dim, order, ncomp, elems)
    ia, ja = orc.graph(e2d, ndof)
    return coords, elems, e2d, ndof, ia, ja


def dense(ia, ja, val):
    n = len(ia) - 1
    return sp.csr_matrix((val, ja, ia), shape=(n, n)).toarray()


@pytest.mark.parametrize("dim", [2, 3])
def test_p1_single_element_closed_form(orc, dim):
    rng = np.random.default_rng(4)
    coords = rng.random((dim + 1, dim))
    elems = np.arange(dim + 1, dtype=np.int32)[None, :]
    e2d, ndof = orc.doftable(dim, 1, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    K = dense(ia, ja, orc.assemble_scalar(dim, 1, 2, coords, elems, e2d, ia, ja, diff=1.0))
    M = dense(ia, ja, orc.assemble_scalar(dim, 1, 2, coords, elems, e2d, ia, ja, diff=0.0, mass=1.0))
    T = np.hstack([np.ones((dim + 1, 1)), coords])
    vol = abs(np.linalg.det(T)) / (2 if dim == 2 else 6)
    G = np.linalg.inv(T)[1:, :].T          # rows: grad lambda_i
    assert np.abs(K - vol * G @ G.T).max() < 1e-12 * np.abs(K).max()
    Mex = vol * (np.ones((dim + 1, dim + 1)) + np.eye(dim + 1)) / (12 if dim == 2 else 20)
    assert np.abs(M - Mex).max() < 1e-14


@pytest.mark.parametrize("dim,order,qorder", [(2, 1, 2), (2, 2, 4), (3, 1, 2), (3, 2, 4), (3, 3, 6)])
def test_mass_energy_is_weighted_measure(orc, dim, order, qorder):
    coords, elems, e2d, ndof, ia, ja = space(orc, dim, order, n=3)
    k = orc.lognormal(len(elems), 0.0, 1.0, 7)
    val = orc.assemble_scalar(dim, order, qorder, coords, elems, e2d, ia, ja, diff=0.0, mass=1.0, kmass=k)
    one = np.ones(ndof)
    X = coords[elems]
    vols = np.abs(np.linalg.det(X[:, 1:, :] - X[:, :1, :])) / (2 if dim == 2 else 6)
    assert abs(one @ orc.spmv(ia, ja, val, one) - (k * vols).sum()) < 1e-12 * (k * vols).sum()


@pytest.mark.parametrize("dim,order,qorder", [(2, 1, 2), (3, 1, 2), (3, 2, 4), (3, 3, 6)])
def test_stiffness_row_sums_vanish_and_symmetric(orc, dim, order, qorder):
    coords, elems, e2d, ndof, ia, ja = space(orc, dim, order, n=3)
    k = orc.lognormal(len(elems), 0.0, 1.0, 8)
    val = orc.assemble_scalar(dim, order, qorder, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k)
    A = dense(ia, ja, val)
    assert np.abs(A.sum(axis=1)).max() < 1e-12 * np.abs(A).max()
    assert np.abs(A - A.T).max() < 1e-14 * np.abs(A).max()


@pytest.mark.parametrize("order,qorder", [(2, 4), (3, 6), (4, 6)])
def test_stiffness_energy_of_quadratic(orc, order, qorder):
    # u = x^2 - y z is interpolated exactly by P2/P3: u^T K u = int |grad u|^2
    coords, elems, e2d, ndof, ia, ja = space(orc, 3, order, n=3, jitter=0.15)
    val = orc.assemble_scalar(3, order, qorder, coords, elems, e2d, ia, ja, diff=1.0)
    pos, _ = dof_positions(coords, elems, e2d, orc.lagrange_nodes(3, order), ndof)
    u = pos[:, 0] ** 2 - pos[:, 1] * pos[:, 2]
    exact = 4.0 / 3.0 + 2.0 / 3.0        # int 4x^2 + z^2 + y^2 over the unit cube
    assert abs(u @ orc.spmv(ia, ja, val, u) - exact) < 1e-11


def test_elasticity_rigid_modes_and_linear_energy(orc):
    coords, elems = orc.mesh_kuhn(2, 1, 1, lx=1.0, ly=0.25, lz=0.25, jitter=0.1, seed=3)
    for order, qorder in [(1, 2), (2, 4), (3, 6), (4, 8)]:
        e2d, ndof = orc.doftable(3, order, 3, elems)
        ia, ja = orc.graph(e2d, ndof)
        E = np.full(len(elems), 10.0)
        nu = np.full(len(elems), 0.3)
        val = orc.assemble_elasticity(order, qorder, coords, elems, e2d, E, nu, ia, ja)
        A = dense(ia, ja, val)
        assert np.abs(A - A.T).max() < 1e-13 * np.abs(A).max()
        pos, _ = dof_positions(coords, elems, e2d, orc.lagrange_nodes(3, order), ndof, ncomp=3)
        pos = pos[0::3]
        modes = []
        for c in range(3):
            t = np.zeros((ndof // 3, 3)); t[:, c] = 1; modes.append(t.reshape(-1))
        for (a, b) in [(0, 1), (1, 2), (0, 2)]:
            r = np.zeros((ndof // 3, 3)); r[:, a] = -pos[:, b]; r[:, b] = pos[:, a]; modes.append(r.reshape(-1))
        for m in modes:
            assert np.abs(A @ m).max() < 1e-11 * np.abs(A).max()
        # linear displacement u = G x: energy = vol*(lam tr(eps)^2 + 2 mu eps:eps)
        Gm = np.array([[0.1, 0.2, -0.3], [0.05, -0.2, 0.1], [0.3, 0.0, 0.25]])
        u = (pos @ Gm.T).reshape(-1)
        eps = 0.5 * (Gm + Gm.T)
        lam = 10 * 0.3 / (1.3 * 0.4); mu = 10 / 2.6
        exact = 0.25 * 0.25 * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
        assert abs(u @ A @ u - exact) < 1e-12


def test_rhs_constant_source_sums_to_measure(orc):
    for dim, order, q in [(2, 1, 2), (3, 2, 2), (3, 3, 3)]:
        coords, elems, e2d, ndof, ia, ja = space(orc, dim, order, n=3)
        b = orc.assemble_rhs(dim, order, 1, q, coords, elems, e2d, ndof, orc.FK_CONST, [2.5])
        assert abs(b.sum() - 2.5) < 1e-13


def test_rhs_sinsin_p1_converges(orc):
    # sum_i b_i = int f = 4/pi^2 up to quadrature error O(h^2)
    errs = []
    for n in [8, 16, 32]:
        coords, elems, e2d, ndof, ia, ja = space(orc, 2, 1, n=n, jitter=0.0)
        b = orc.assemble_rhs(2, 1, 1, 2, coords, elems, e2d, ndof, orc.FK_SINSIN, [1.0])
        errs.append(abs(b.sum() - 4 / np.pi ** 2))
    assert errs[1] < errs[0] / 3 and errs[2] < errs[1] / 3


@pytest.mark.parametrize("dim,order,qorder", [(2, 1, 2), (2, 2, 2), (3, 1, 0), (3, 2, 2), (3, 4, 6)])
def test_dirichlet_patch_test(orc, dim, order, qorder):
    # pyfeelpp tests/test_forms.py:44-53: Laplace with u = x on the boundary gives u = x
    coords, elems, e2d, ndof, ia, ja = space(orc, dim, order, n=4, jitter=0.2)
    val = orc.assemble_scalar(dim, order, qorder, coords, elems, e2d, ia, ja, diff=1.0)
    b = np.zeros(ndof)
    isb = orc.boundary_dofs(dim, order, 1, coords, elems, e2d, ndof)
    pos, _ = dof_positions(coords, elems, e2d, orc.lagrange_nodes(dim, order), ndof)
    g = pos[:, 0].copy()
    for mode in (0, 1):
        v, bb = val.copy(), b.copy()
        orc.dirichlet(mode, ia, ja, v, bb, isb, g)
        x, its, res, reason = orc.pcg(ia, ja, v, bb, rtol=1e-13, maxit=2000) if mode == 1 else \
            (spla.spsolve(sp.csr_matrix((v, ja, ia)), bb), 0, 0, 2)
        assert reason > 0
        assert np.abs(x - g).max() < 1e-10


def test_boundary_dof_counts(orc):
    coords, elems = orc.mesh_kuhn(4)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    isb = orc.boundary_dofs(3, 2, 1, coords, elems, e2d, ndof)
    # 8n+1 lattice on each axis: boundary nodes of the (2n+1)^3 P2 lattice minus face-centre holes
    pos, _ = dof_positions(coords, elems, e2d, orc.lagrange_nodes(3, 2), ndof)
    onb = np.any((pos < 1e-12) | (pos > 1 - 1e-12), axis=1)
    assert (isb.astype(bool) == onb).all()
    # clamped end x = 0 for a vector space
    e2dv, ndv = orc.doftable(3, 1, 3, elems)
    isbv = orc.boundary_dofs(3, 1, 3, coords, elems, e2dv, ndv, sel=1)
    assert isbv.sum() == 3 * 25


def test_pcg_matches_direct_solve(orc):
    coords, elems, e2d, ndof, ia, ja = space(orc, 2, 1, n=16, jitter=0.2, seed=1)
    val = orc.assemble_scalar(2, 1, 2, coords, elems, e2d, ia, ja, diff=1.0, mass=1.0)
    b = orc.assemble_rhs(2, 1, 1, 2, coords, elems, e2d, ndof, orc.FK_SINSIN, [1.0])
    isb = orc.boundary_dofs(2, 1, 1, coords, elems, e2d, ndof)
    orc.dirichlet(1, ia, ja, val, b, isb)
    xd = spla.spsolve(sp.csr_matrix((val, ja, ia)).tocsc(), b)
    x, its, res, reason = orc.pcg(ia, ja, val, b, rtol=1e-12, maxit=1000)
    assert reason == 2 and 0 < its < 200
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10


def test_pcg_status_semantics(orc):
    coords, elems, e2d, ndof, ia, ja = space(orc, 2, 1, n=8)
    val = orc.assemble_scalar(2, 1, 2, coords, elems, e2d, ia, ja, diff=1.0, mass=1.0)
    b = np.ones(ndof)
    x, its, res, reason = orc.pcg(ia, ja, val, b, rtol=1e-12, maxit=1)
    assert its == 1 and reason == -3        # KSP_DIVERGED_ITS
    x, its, res, reason = orc.pcg(ia, ja, val, np.zeros(ndof), rtol=1e-12)
    assert its == 0 and reason > 0          # zero rhs converges immediately
    # an indefinite operator is reported through the status, not raised:
    # A = [[1,2],[2,1]] has eigenvalues 3 and -1; the second step has p'Ap < 0
    # (KSP_DIVERGED_INDEFINITE_MAT)
    ia2 = np.array([0, 2, 4], np.int32); ja2 = np.array([0, 1, 0, 1], np.int32)
    x, its, res, reason = orc.pcg(ia2, ja2, np.array([1.0, 2.0, 2.0, 1.0]), np.array([1.0, 0.0]),
                                  rtol=1e-12, maxit=50)
    assert reason == -10 and its == 2
    # A = [[1,1/2],[1/2,-1]]: the Jacobi preconditioner diag(1,-1) is indefinite;
    # r1 = (0,-1/2), z1 = (0,1/2), so (r,z) turns from 1 to -1/4 at the top of
    # the second step (KSP_DIVERGED_INDEFINITE_PC)
    x, its, res, reason = orc.pcg(ia2, ja2, np.array([1.0, 0.5, 0.5, -1.0]), np.array([1.0, 0.0]),
                                  rtol=1e-12, maxit=50)
    assert reason == -8 and its == 2
    assert np.allclose(x, [1.0, 0.0])             # x1 = x0 + p0 (a = 1), left where CG stopped


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_baseline_matches(orc, threads):
    # the threaded CPU baseline assembles the same values (row ownership keeps
    # the per-row addition order) and solves to the same solution
    coords, elems = orc.mesh_kuhn(5, jitter=0.15, seed=11)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    k = orc.lognormal(len(elems), 0.0, 1.0, 5)
    v1 = orc.assemble_scalar(3, 2, 4, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k, mass=0.5)
    vt = orc.assemble_scalar_mt(3, 2, 4, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k, mass=0.5, threads=threads)
    assert np.array_equal(v1, vt)
    b = orc.assemble_rhs(3, 2, 1, 4, coords, elems, e2d, ndof, orc.FK_CONST, [1.0])
    x1, its1, _, r1 = orc.pcg(ia, ja, v1, b, rtol=1e-12, maxit=5000)
    xt, itst, _, rt = orc.pcg_mt(ia, ja, vt, b, rtol=1e-12, maxit=5000, threads=threads)
    assert r1 > 0 and rt > 0 and abs(its1 - itst) <= 2
    assert np.linalg.norm(x1 - xt) <= 1e-10 * np.linalg.norm(x1)


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_oracle_norms_closed_forms(orc, dim, order):
    # pins of the norm restatement by what the mathematics fixes:
    # ||1||_L2 = sqrt(measure); the interpolant of an affine function has zero
    # L2 and H1-seminorm error; |grad(a.x)|^2 integrates to |a|^2 measure
    if dim == 2:
        coords, elems = orc.mesh_square(6, jitter=0.2, seed=3)
    else:
        coords, elems = orc.mesh_kuhn(3, jitter=0.15, seed=3)
    e2d, ndof = orc.doftable(dim, order, 1, elems)
    nodes = orc.lagrange_nodes(dim, order)
    pos, mism = dof_positions(coords, elems, e2d, nodes, ndof)
    assert mism < 1e-12
    zero = np.zeros(ndof)
    one = [(orc.FN_LINEAR, [1.0, 0, 0, 0])]
    assert abs(orc.norm(dim, order, 1, 2 * order, coords, elems, e2d, zero, one) - 1.0) < 1e-13
    a = np.array([0.3, -1.2, 0.7, 2.0][:dim + 1])
    lin = [(orc.FN_LINEAR, list(a))]
    u = a[0] + pos @ a[1:]
    assert orc.norm(dim, order, 1, 2 * order, coords, elems, e2d, u, lin) < 1e-13
    assert orc.norm(dim, order, 1, 2 * order, coords, elems, e2d, u, lin, semi=True) < 1e-12
    h1 = orc.norm(dim, order, 1, 2 * order, coords, elems, e2d, u, None, semi=True)
    assert abs(h1 - np.linalg.norm(a[1:])) < 1e-12


def test_oracle_norm_convergence(orc):
    # interpolation error of sin(pi x) sin(pi y) in L2 decays like h^(k+1)
    errs = []
    for n in (8, 16):
        coords, elems = orc.mesh_square(n, jitter=0.0, seed=0)
        e2d, ndof = orc.doftable(2, 2, 1, elems)
        pos, _ = dof_positions(coords, elems, e2d, orc.lagrange_nodes(2, 2), ndof)
        u = np.sin(np.pi * pos[:, 0]) * np.sin(np.pi * pos[:, 1])
        errs.append(orc.norm(2, 2, 1, 6, coords, elems, e2d, u, [(orc.FN_SINSIN, [1.0, 0, 0, 0])]))
    rate = np.log2(errs[0] / errs[1])
    assert 2.8 < rate < 3.2, rate
