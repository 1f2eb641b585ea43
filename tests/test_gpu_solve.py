"""GPU Jacobi-CG through the C-ABI vs the FP64 oracle: FP64 solutions within
1e-10 relative error of the oracle's tightly converged solve; FP32-in-FP64
refinement within 1e-8; PETSc status semantics."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import feelpp_amd as fa
from helpers import rel_fro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return fa.Context(0)


def problem(ctx, orc, dim, order, n, qo, rhs_kind, rhs_q, jitter, seed, mass=0.0, lognormal=True):
    if dim == 2:
        coords, elems = orc.mesh_square(n, jitter, seed)
    else:
        coords, elems = orc.mesh_kuhn(n, jitter=jitter, seed=seed)
    e2d, ndof = orc.doftable(dim, order, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    k = orc.lognormal(len(elems), 0.0, 1.0, seed + 1) if lognormal else np.ones(len(elems))
    val = orc.assemble_scalar(dim, order, qo, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k, mass=mass)
    b = orc.assemble_rhs(dim, order, 1, rhs_q, coords, elems, e2d, ndof, rhs_kind, [1.0])
    isb = orc.boundary_dofs(dim, order, 1, coords, elems, e2d, ndof)
    orc.dirichlet(1, ia, ja, val, b, isb)
    space = fa.Space(ctx, coords, elems, order, 1, e2d, ndof)
    g = fa.Graph(space, ia, ja)
    A = fa.Matrix(g, fa.FP64_FP32)
    kd = ctx.vector(len(elems), k)
    A.assemble(fa.DIFFUSION | (fa.MASS if mass else 0), qo, kdiff=kd, mass=mass)
    bb = ctx.vector(ndof)
    fa.assemble_linear(space, bb, rhs_kind, rhs_q, (1.0,))
    rows = fa.Rows(ctx, ndof, np.nonzero(isb)[0])
    A.zero_rows(rows, None, bb, mode=fa.ELIMINATION_SYMMETRIC)
    return dict(ia=ia, ja=ja, val=val, b=b, A=A, bb=bb, ndof=ndof, keep=(space, g, kd, rows))


def test_c1_solve_matches_oracle_and_direct(ctx, orc):
    P = problem(ctx, orc, 2, 1, 128, 2, orc.FK_SINSIN, 2, 0.2, 1, mass=1.0, lognormal=False)
    xo, its_o, res_o, reason_o = orc.pcg(P["ia"], P["ja"], P["val"], P["b"], rtol=1e-12, maxit=5000)
    x = ctx.vector(P["ndof"])
    st = P["A"].solve(P["bb"], x, rtol=1e-12, maxit=5000)
    print("oracle its", its_o, "gpu", st)
    assert st["converged"] and st["reason"] == 2
    assert abs(st["its"] - its_o) <= 2
    xd = spla.spsolve(sp.csr_matrix((P["val"], P["ja"], P["ia"])).tocsc(), P["b"])
    assert rel_fro(x.get(), xo) < 1e-10
    assert rel_fro(x.get(), xd) < 1e-10


@pytest.mark.parametrize("order,n,qo", [(2, 8, 4), (3, 4, 6)])
def test_p2_p3_poisson_fp64(ctx, orc, order, n, qo):
    P = problem(ctx, orc, 3, order, n, qo, orc.FK_CONST, order, 0.15, 9)
    xo, its_o, _, _ = orc.pcg(P["ia"], P["ja"], P["val"], P["b"], rtol=1e-13, maxit=10000)
    x = ctx.vector(P["ndof"])
    st = P["A"].solve(P["bb"], x, rtol=1e-13, maxit=10000)
    assert st["converged"]
    assert abs(st["its"] - its_o) <= 2
    assert rel_fro(x.get(), xo) < 1e-10


@pytest.mark.parametrize("delta", [0.1, 1e-2, 1e-4])
def test_mixed_precision_reliable_updates(ctx, orc, delta):
    # FP32 CG with FP64 reliable updates: the FP64 true residual meets the
    # tolerance, the solution matches the oracle to 1e-8, and the iteration
    # count stays near the FP64 one (the Krylov recurrence is not restarted)
    P = problem(ctx, orc, 3, 2, 8, 4, orc.FK_CONST, 2, 0.15, 21)
    xo, its_o, _, _ = orc.pcg(P["ia"], P["ja"], P["val"], P["b"], rtol=1e-10, maxit=10000)
    xt, _, _, _ = orc.pcg(P["ia"], P["ja"], P["val"], P["b"], rtol=1e-14, maxit=10000)
    x = ctx.vector(P["ndof"])
    st = P["A"].solve(P["bb"], x, rtol=1e-10, maxit=20000, precision=fa.FP32, inner_rtol=delta)
    print(st, its_o)
    assert st["converged"] and st["outer_its"] >= 2
    assert st["residual"] <= 1e-10 * st["rnorm0"]
    assert rel_fro(x.get(), xt) < 1e-8
    assert st["its"] <= 1.5 * its_o
    assert st["fp64_its"] == 0   # well conditioned: FP32 all the way


def test_status_semantics(ctx, orc):
    P = problem(ctx, orc, 2, 1, 16, 2, orc.FK_CONST, 1, 0.1, 3)
    x = ctx.vector(P["ndof"])
    st = P["A"].solve(P["bb"], x, rtol=1e-12, maxit=1)
    assert st["its"] == 1 and st["reason"] == -3 and not st["converged"]
    z = ctx.vector(P["ndof"])
    x.zero()
    st = P["A"].solve(z, x, rtol=1e-12)
    assert st["its"] == 0 and st["converged"]
    # x holds the initial guess (UIRNorm: the test is relative to ||B r_0|| of that guess)
    x.zero()
    st1 = P["A"].solve(P["bb"], x, rtol=1e-12, maxit=1000)
    st2 = P["A"].solve(P["bb"], x, rtol=1e-3, maxit=1000)
    assert st1["converged"] and st2["converged"]
    assert st2["rnorm0"] < 1e-10 * st1["rnorm0"]


def test_solve_deterministic(ctx, orc):
    P = problem(ctx, orc, 3, 2, 5, 4, orc.FK_CONST, 2, 0.15, 4)
    x1, x2 = ctx.vector(P["ndof"]), ctx.vector(P["ndof"])
    s1 = P["A"].solve(P["bb"], x1, rtol=1e-10)
    s2 = P["A"].solve(P["bb"], x2, rtol=1e-10)
    assert s1 == s2 and np.array_equal(x1.get(), x2.get())


@pytest.mark.parametrize("order", [2, 3])
def test_clamped_beam_elasticity(ctx, orc, order):
    # left face clamped, body force (0,0,-1); FP64 and mixed solves vs the oracle
    coords, elems = orc.mesh_kuhn(8, 2, 2, lx=1.0, ly=0.25, lz=0.25, jitter=0.1, seed=17)
    e2d, ndof = orc.doftable(3, order, 3, elems)
    ia, ja = orc.graph(e2d, ndof)
    E = orc.uniform(len(elems), 1.0, 100.0, 1, log=True)
    nu = orc.uniform(len(elems), 0.2, 0.35, 2)
    qo = 2 * order
    val = orc.assemble_elasticity(order, qo, coords, elems, e2d, E, nu, ia, ja)
    b = orc.assemble_rhs(3, order, 3, qo, coords, elems, e2d, ndof, orc.FK_VECTOR, [0.0, 0.0, -1.0])
    isb = orc.boundary_dofs(3, order, 3, coords, elems, e2d, ndof, sel=1)
    orc.dirichlet(1, ia, ja, val, b, isb)
    xo, its_o, _, reason = orc.pcg(ia, ja, val, b, rtol=1e-13, maxit=50000)
    assert reason > 0
    space = fa.Space(ctx, coords, elems, order, 3, e2d, ndof)
    g = fa.Graph(space, ia, ja)
    A = fa.Matrix(g, fa.FP64_FP32)
    A.assemble(fa.ELASTICITY, qo, young=ctx.vector(len(elems), E), poisson=ctx.vector(len(elems), nu))
    bb = ctx.vector(ndof)
    fa.assemble_linear(space, bb, fa.SRC_VECTOR, qo, (0.0, 0.0, -1.0))
    assert rel_fro(bb.get(), orc.assemble_rhs(3, order, 3, qo, coords, elems, e2d, ndof, orc.FK_VECTOR,
                                              [0.0, 0.0, -1.0])) < 1e-12
    rows = fa.Rows(ctx, ndof, np.nonzero(isb)[0])
    A.zero_rows(rows, None, bb, mode=fa.ELIMINATION_SYMMETRIC)
    x = ctx.vector(ndof)
    st = A.solve(bb, x, rtol=1e-13, maxit=50000)
    assert st["converged"] and rel_fro(x.get(), xo) < 1e-10
    x.zero()
    st = A.solve(bb, x, rtol=1e-10, maxit=50000, precision=fa.FP32)
    print(st)
    assert st["converged"] and rel_fro(x.get(), xo) < 1e-8
    # kappa(D^-1 A) of the clamped beam is ~1e5-1e6: the FP32 solve hands over to FP64
    assert 0 < st["fp64_its"] < st["its"]
