"""GPU assembly through the C-ABI vs the FP64 oracle.

Bar: CSR pattern bit-exact (the device scatters into the host graph, which
tests/test_host_parity.py pins to the oracle bit for bit), FP64 values within
1e-12 relative Frobenius norm, FP32 values within 1e-6."""
import numpy as np
import pytest

import feelpp_amd as fa
from helpers import dof_positions, rel_fro

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module")
def ctx():
    return fa.Context(0)


def setup(ctx, orc, dim, order, n, jitter=0.15, seed=5, ncomp=1):
    if dim == 2:
        coords, elems = orc.mesh_square(n, jitter, seed)
    else:
        coords, elems = orc.mesh_kuhn(n, jitter=jitter, seed=seed)
    e2d, ndof = orc.doftable(dim, order, ncomp, elems)
    ia, ja = orc.graph(e2d, ndof)
    sp = fa.Space(ctx, coords, elems, order, ncomp, e2d, ndof)
    g = fa.Graph(sp, ia, ja)
    return coords, elems, e2d, ndof, ia, ja, sp, g


@pytest.mark.parametrize("dim,order,n,qo", [(2, 1, 16, 2), (2, 2, 8, 4), (3, 1, 5, 2), (3, 2, 5, 4), (3, 3, 3, 6),
                                            (3, 4, 2, 8)])
def test_diffusion_mass_fp64_vs_oracle(ctx, orc, dim, order, n, qo):
    coords, elems, e2d, ndof, ia, ja, sp, g = setup(ctx, orc, dim, order, n)
    k = orc.lognormal(len(elems), 0.0, 1.0, 11)
    m = orc.uniform(len(elems), 0.5, 2.0, 12)
    ref = orc.assemble_scalar(dim, order, qo, coords, elems, e2d, ia, ja, diff=1.3, kdiff=k, mass=0.7, kmass=m)
    A = fa.Matrix(g, fa.FP64_FP32)
    kd, km = ctx.vector(len(elems), k), ctx.vector(len(elems), m)
    A.assemble(fa.MASS | fa.DIFFUSION, qo, diff=1.3, kdiff=kd, mass=0.7, kmass=km)
    v = A.values()
    print(f"dim={dim} P{order} nnz={len(ja)} relfro={rel_fro(v, ref):.2e}")
    assert rel_fro(v, ref) < TOL64
    v32 = A.values(fa.FP32).astype(np.float64)
    assert rel_fro(v32, ref) < 1e-6


@pytest.mark.parametrize("order,n,qo", [(3, 3, 6), (4, 2, 8)])
def test_scalar_element_matrix_paths(ctx, orc, order, n, qo):
    # P3/P4 tetrahedra: fused contraction, matrix-core GEMM, and the timed
    # choice all match the oracle; replace and accumulate both honoured
    coords, elems, e2d, ndof, ia, ja, sp, g = setup(ctx, orc, 3, order, n)
    k = orc.lognormal(len(elems), 0.0, 1.0, 11)
    m = orc.uniform(len(elems), 0.5, 2.0, 12)
    ref = orc.assemble_scalar(3, order, qo, coords, elems, e2d, ia, ja, diff=1.3, kdiff=k, mass=0.7, kmass=m)
    kd, km = ctx.vector(len(elems), k), ctx.vector(len(elems), m)
    out = {}
    for mat in (fa.MATRICES_FUSED, fa.MATRICES_MFMA, fa.MATRICES_AUTO):
        A = fa.Matrix(g, fa.FP64_FP32)
        A.assemble(fa.MASS | fa.DIFFUSION, qo, diff=1.3, kdiff=kd, mass=0.7, kmass=km, matrices=mat)
        assert rel_fro(A.values(), ref) < TOL64
        assert rel_fro(A.values(fa.FP32).astype(np.float64), ref) < 1e-6
        out[mat] = A.values()
        A.assemble(fa.MASS | fa.DIFFUSION, qo, diff=1.3, kdiff=kd, mass=0.7, kmass=km, matrices=mat, add=True)
        assert rel_fro(A.values(), 2.0 * ref) < TOL64
    assert rel_fro(out[fa.MATRICES_MFMA], out[fa.MATRICES_FUSED]) < 1e-14
    coords2, elems2, *_, g2 = setup(ctx, orc, 3, 2, 2)
    with pytest.raises(fa.FeelHipError):
        fa.Matrix(g2).assemble(fa.DIFFUSION, 4, matrices=fa.MATRICES_MFMA)


def test_c1_full_size_stiffness_plus_mass(ctx, orc):
    # 128x128 P1 triangles, jitter 0.2h, k = 1, quadrature order 2
    coords, elems, e2d, ndof, ia, ja, sp, g = setup(ctx, orc, 2, 1, 128, jitter=0.2, seed=1)
    ref = orc.assemble_scalar(2, 1, 2, coords, elems, e2d, ia, ja, diff=1.0, mass=1.0)
    A = fa.Matrix(g)
    A.assemble(fa.MASS | fa.DIFFUSION, 2)
    assert rel_fro(A.values(), ref) < TOL64


def test_assemble_replace_add_and_determinism(ctx, orc):
    coords, elems, e2d, ndof, ia, ja, sp, g = setup(ctx, orc, 3, 2, 4)
    k = ctx.vector(len(elems), orc.lognormal(len(elems), 0.0, 1.0, 3))
    A = fa.Matrix(g)
    A.assemble(fa.DIFFUSION, 4, kdiff=k)
    v1 = A.values()
    A.assemble(fa.DIFFUSION, 4, kdiff=k)
    assert np.array_equal(v1, A.values())            # operator= replaces, run-to-run bitwise identical
    A.assemble(fa.MASS, 4, mass=2.0, add=True)      # operator+=
    ref = orc.assemble_scalar(3, 2, 4, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k.get(), mass=2.0)
    assert rel_fro(A.values(), ref) < TOL64


def test_add_matrix_host_path(ctx, orc):
    coords, elems, e2d, ndof, ia, ja, sp, g = setup(ctx, orc, 2, 1, 3)
    A = fa.Matrix(g)
    rows = e2d[0]
    blk = np.arange(9.0).reshape(3, 3)
    A.add_matrix(rows, rows, blk)
    A.add_matrix(rows, rows, blk)
    v = A.values()
    for i, r in enumerate(rows):
        for j, c in enumerate(rows):
            p = ia[r] + np.searchsorted(ja[ia[r]:ia[r + 1]], c)
            assert v[p] == 2 * blk[i, j]
    with pytest.raises(fa.FeelHipError):
        A.add_matrix([0], [ndof - 1], [[1.0]])       # outside the pattern


@pytest.mark.parametrize("dim,order,kind,qo,params", [
    (2, 1, fa.SRC_SINSIN, 2, (1.0,)), (2, 2, fa.SRC_CONST, 2, (3.0,)),
    (3, 2, fa.SRC_CONST, 2, (1.0,)), (3, 3, fa.SRC_SINSIN, 6, (2.0,))])
def test_rhs_vs_oracle(ctx, orc, dim, order, kind, qo, params):
    coords, elems, e2d, ndof, ia, ja, sp, g = setup(ctx, orc, dim, order, 6 if dim == 3 else 12)
    ref = orc.assemble_rhs(dim, order, 1, qo, coords, elems, e2d, ndof, kind, list(params))
    b = ctx.vector(ndof)
    fa.assemble_linear(sp, b, kind, qo, params)
    assert rel_fro(b.get(), ref) < TOL64
    fa.assemble_linear(sp, b, kind, qo, params, add=True)
    assert rel_fro(b.get(), 2 * ref) < TOL64


@pytest.mark.parametrize("mode", [fa.ELIMINATION, fa.ELIMINATION_SYMMETRIC])
def test_dirichlet_vs_oracle(ctx, orc, mode):
    coords, elems, e2d, ndof, ia, ja, sp, g = setup(ctx, orc, 3, 2, 4)
    val = orc.assemble_scalar(3, 2, 4, coords, elems, e2d, ia, ja, diff=1.0)
    b = orc.assemble_rhs(3, 2, 1, 2, coords, elems, e2d, ndof, orc.FK_CONST, [1.0])
    isb = orc.boundary_dofs(3, 2, 1, coords, elems, e2d, ndof)
    gval = np.sin(np.arange(ndof))
    orc.dirichlet(mode, ia, ja, val, b, isb, gval)
    A = fa.Matrix(g, fa.FP64_FP32)
    A.assemble(fa.DIFFUSION, 4)
    bb = ctx.vector(ndof)
    fa.assemble_linear(sp, bb, fa.SRC_CONST, 2, (1.0,))
    rows = fa.Rows(ctx, ndof, np.nonzero(isb)[0])
    A.zero_rows(rows, ctx.vector(ndof, gval), bb, mode=mode)
    assert rel_fro(A.values(), val) < TOL64
    assert rel_fro(bb.get(), b) < TOL64
    assert np.array_equal(A.values(fa.FP32) == 0, A.values() == 0)   # FP32 copy eliminated alike
    # the same row set on a second graph of the same pattern (its rows to touch
    # are listed per graph) and again on the first: the same bits each time;
    # duplicated row ids change nothing
    g2 = fa.Graph(sp, ia, ja)
    dup = np.concatenate([np.nonzero(isb)[0], np.nonzero(isb)[0][::3]])
    for graph, rs in ((g2, rows), (g, rows), (g2, fa.Rows(ctx, ndof, dup))):
        A2 = fa.Matrix(graph, fa.FP64_FP32)
        A2.assemble(fa.DIFFUSION, 4)
        b2 = ctx.vector(ndof)
        fa.assemble_linear(sp, b2, fa.SRC_CONST, 2, (1.0,))
        A2.zero_rows(rs, ctx.vector(ndof, gval), b2, mode=mode)
        assert np.array_equal(A2.values(), A.values())
        assert np.array_equal(b2.get(), bb.get())


def test_full_size_p2_properties(ctx):
    # 64^3 Kuhn P2 (1 572 864 tets, 2 146 689 dofs, 60 859 905 nnz): size-independent checks
    n = 64
    coords, elems = fa.mesh_kuhn_box(n, jitter=0.15, seed=2024)
    e2d, ndof = fa.doftable(3, 2, 1, elems, coords.shape[0])
    ia, ja = fa.graph(e2d, ndof)
    assert ndof == 2146689 and len(ja) == 60859905
    sp = fa.Space(ctx, coords, elems, 2, 1, e2d, ndof)
    g = fa.Graph(sp, ia, ja)
    A = fa.Matrix(g)
    A.assemble(fa.DIFFUSION, 4)
    x, y = ctx.vector(ndof, np.ones(ndof)), ctx.vector(ndof)
    A.mult(x, y)
    v = A.values()
    assert np.abs(y.get()).max() < 1e-12 * np.abs(v).max()           # constants in the kernel
    pos, _ = dof_positions(coords, elems, e2d, fa.lagrange_nodes(3, 2), ndof)
    u = pos[:, 0] ** 2 - pos[:, 1] * pos[:, 2]
    x.set(u)
    A.mult(x, y)
    assert abs(u @ y.get() - 2.0) < 1e-9                              # int |grad u|^2 = 4/3 + 2/3
    rng = np.random.default_rng(0)
    a, c = rng.standard_normal(ndof), rng.standard_normal(ndof)
    x.set(a)
    A.mult(x, y)
    Aa = y.get()
    x.set(c)
    A.mult(x, y)
    assert abs(c @ Aa - a @ y.get()) < 1e-10 * abs(c @ Aa) + 1e-12    # symmetric


def beam(ctx, orc, order, qo):
    # 1 x 0.25 x 0.25 beam, E log-uniform in [1,100], nu uniform in [0.2,0.35]
    coords, elems = orc.mesh_kuhn(4, 2, 2, lx=1.0, ly=0.25, lz=0.25, jitter=0.1, seed=13)
    e2d, ndof = orc.doftable(3, order, 3, elems)
    ia, ja = orc.graph(e2d, ndof)
    E = orc.uniform(len(elems), 1.0, 100.0, 5, log=True)
    nu = orc.uniform(len(elems), 0.2, 0.35, 6)
    ref = orc.assemble_elasticity(order, qo, coords, elems, e2d, E, nu, ia, ja)
    sp = fa.Space(ctx, coords, elems, order, 3, e2d, ndof)
    g = fa.Graph(sp, ia, ja)
    return g, ctx.vector(len(elems), E), ctx.vector(len(elems), nu), ref


@pytest.mark.parametrize("order,qo,matrices", [(1, 2, fa.MATRICES_AUTO), (2, 4, fa.MATRICES_AUTO),
                                               (3, 6, fa.MATRICES_AUTO), (4, 8, fa.MATRICES_AUTO),
                                               (3, 6, fa.MATRICES_FUSED), (4, 8, fa.MATRICES_FUSED),
                                               (3, 6, fa.MATRICES_MFMA), (4, 8, fa.MATRICES_MFMA)])
def test_elasticity_fp64_vs_oracle(ctx, orc, order, qo, matrices):
    # AUTO times both paths on the first P3/P4 assembly; FUSED contracts inside
    # the scatter; MFMA materialises the element matrices on the matrix cores
    g, E, nu, ref = beam(ctx, orc, order, qo)
    A = fa.Matrix(g, fa.FP64_FP32)
    A.assemble(fa.ELASTICITY, qo, young=E, poisson=nu, matrices=matrices)
    assert rel_fro(A.values(), ref) < TOL64
    assert rel_fro(A.values(fa.FP32).astype(np.float64), ref) < 1e-6
    A.assemble(fa.ELASTICITY, qo, young=E, poisson=nu, matrices=matrices, add=True)
    assert rel_fro(A.values(), 2.0 * ref) < TOL64


def test_elasticity_paths_agree_and_reject_misuse(ctx, orc):
    g, E, nu, ref = beam(ctx, orc, 4, 8)
    out = []
    for m in (fa.MATRICES_FUSED, fa.MATRICES_MFMA):
        A = fa.Matrix(g, fa.FP64)
        A.assemble(fa.ELASTICITY, 8, young=E, poisson=nu, matrices=m)
        out.append(A.values())
        A.assemble(fa.ELASTICITY, 8, young=E, poisson=nu, matrices=m)   # workspace reused
        assert np.array_equal(A.values(), out[-1])
    # same entries, different summation order inside an entry: round-off apart
    assert rel_fro(out[1], out[0]) < 1e-14
    # the first AUTO assembly times both (the matrix-core run last), later ones
    # take the faster: every result is one of the two paths' bits
    A = fa.Matrix(g, fa.FP64)
    for _ in range(2):
        A.assemble(fa.ELASTICITY, 8, young=E, poisson=nu)
        v = A.values()
        assert np.array_equal(v, out[0]) or np.array_equal(v, out[1])
    g2, E2, nu2, _ = beam(ctx, orc, 2, 4)
    with pytest.raises(fa.FeelHipError):
        fa.Matrix(g2).assemble(fa.ELASTICITY, 4, young=E2, poisson=nu2, matrices=fa.MATRICES_MFMA)
    with pytest.raises(fa.FeelHipError):
        fa.Matrix(g).assemble(fa.ELASTICITY, 8, young=E, poisson=nu, matrices=7)


def test_unsupported_requests_fail_loudly(ctx, orc):
    coords, elems, e2d, ndof, ia, ja, sp, g = setup(ctx, orc, 3, 1, 2)
    A = fa.Matrix(g)
    with pytest.raises(fa.FeelHipError):
        A.assemble(fa.ELASTICITY, 2)                       # scalar space
    with pytest.raises(fa.FeelHipError):
        A.assemble(8, 2)                                   # unknown expression kind
    with pytest.raises(fa.FeelHipError):
        A.assemble(fa.DIFFUSION, 99)                       # no such quadrature
    with pytest.raises(fa.FeelHipError):
        fa.Space(ctx, coords, elems, 5, 1, e2d, ndof)      # P5 not provided
