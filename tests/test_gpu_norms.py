"""Post-solve evaluation on the GPU: normL2 / H1 seminorm of u - f vs the
oracle (feelvf/norml2.hpp:37-57), and the check pyfeelpp's test_forms.py makes
after a Dirichlet patch solve (python/pyfeelpp/tests/test_forms.py:44-53)."""
import numpy as np
import pytest

import feelpp_amd as fa
from helpers import dof_positions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return fa.Context(0)


def mesh(orc, dim):
    if dim == 2:
        return orc.mesh_square(7, jitter=0.2, seed=11)
    return orc.mesh_kuhn(3, jitter=0.15, seed=11)


@pytest.mark.parametrize("dim,order,ncomp", [(2, 1, 1), (2, 2, 1), (3, 1, 1), (3, 2, 1), (3, 3, 1), (3, 4, 1),
                                             (3, 2, 3)])
@pytest.mark.parametrize("semi", [False, True])
def test_norms_match_oracle(ctx, orc, dim, order, ncomp, semi):
    coords, elems = mesh(orc, dim)
    e2d, ndof = orc.doftable(dim, order, ncomp, elems)
    rng = np.random.default_rng(order + 10 * dim + ncomp)
    u = rng.standard_normal(ndof)
    funcs = [(orc.FN_SINSIN, [1.3, 0, 0, 0]), (orc.FN_LINEAR, [0.2, 1.0, -0.5, 0.3]), (orc.FN_ZERO, [0, 0, 0, 0])]
    funcs = funcs[:ncomp] if ncomp > 1 else funcs[:1]
    space = fa.Space(ctx, coords, elems, order, ncomp, e2d, ndof)
    uv = ctx.vector(ndof, u)
    qo = 2 * order + 2
    fn = fa.norm_semi_h1 if semi else fa.norm_l2
    got = fn(space, uv, funcs, qo)
    ref = orc.norm(dim, order, ncomp, qo, coords, elems, e2d, u, funcs, semi=semi)
    assert abs(got - ref) <= 1e-12 * ref


@pytest.mark.parametrize("dim", [2, 3])
def test_patch_solve_then_norm(ctx, orc, dim):
    # a = grad-grad, l = 0, on(boundary, g = x): the solution is x, so
    # normL2(u - x) vanishes up to the solver tolerance
    coords, elems = mesh(orc, dim)
    order = 2
    e2d, ndof = orc.doftable(dim, order, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    space = fa.Space(ctx, coords, elems, order, 1, e2d, ndof)
    A = fa.Matrix(fa.Graph(space, ia, ja), fa.FP64)
    A.assemble(fa.DIFFUSION, 2 * order - 2)
    b = ctx.vector(ndof)
    isb = orc.boundary_dofs(dim, order, 1, coords, elems, e2d, ndof, sel=0)
    pos, _ = dof_positions(coords, elems, e2d, orc.lagrange_nodes(dim, order), ndof)
    g = ctx.vector(ndof, pos[:, 0])
    A.zero_rows(fa.Rows(ctx, ndof, np.nonzero(isb)[0]), g, b, mode=fa.ELIMINATION_SYMMETRIC)
    x = ctx.vector(ndof)
    st = A.solve(b, x, rtol=1e-13, maxit=10000)
    assert st["converged"]
    lin = [(fa.FN_LINEAR, [0.0, 1.0, 0.0, 0.0])]
    assert fa.norm_l2(space, x, lin) < 1e-10
    assert fa.norm_semi_h1(space, x, lin) < 1e-9
    assert abs(fa.norm_semi_h1(space, x) - 1.0) < 1e-9   # |grad x| = 1 on the unit domain
