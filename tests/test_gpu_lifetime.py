"""Handle lifetime: reference-counted destroy in any order (the reference shares
matrices and graphs through shared_ptr, feelalg/matrixsparse.hpp:113)."""
import numpy as np
import pytest

import feelpp_amd as fa

pytestmark = pytest.mark.gpu


def build(orc):
    ctx = fa.Context(0)
    coords, elems = orc.mesh_kuhn(2, jitter=0.1, seed=3)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    space = fa.Space(ctx, coords, elems, 2, 1, e2d, ndof)
    g = fa.Graph(space)
    A = fa.Matrix(g, fa.FP64)
    A.assemble(fa.DIFFUSION, 2)
    x = ctx.vector(ndof, np.ones(ndof))
    y = ctx.vector(ndof)
    return ctx, space, g, A, x, y


@pytest.mark.parametrize("order", ["parents_first", "children_first"])
def test_destroy_order(orc, order):
    ctx, space, g, A, x, y = build(orc)
    A.mult(x, y)
    ref = y.get()
    assert np.abs(ref).max() < 1e-12   # stiffness annihilates constants
    handles = [ctx, space, g, A, x, y]
    if order == "children_first":
        handles.reverse()
    for i, h in enumerate(handles):
        h.close()
        # what is still open keeps working after its parents were released
        rest = handles[i + 1:]
        if A in rest and x in rest and y in rest:
            A.mult(x, y)
            assert np.array_equal(y.get(), ref)
