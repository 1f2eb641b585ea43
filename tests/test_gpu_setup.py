"""Device-built DofTable and sparsity pattern (GPU, C-ABI).

feel_hip_space_create_from_mesh numbers the dofs on the device and
feel_hip_graph_create_from_space builds the COUPLED pattern there; both must
be bit-identical to the oracle's first-touch DofTable and Stencil graph
(tests/test_oracle_space_graph.py pins those), on Feel++'s 8-triangle fixture,
jittered Kuhn boxes and meshes with randomly relabelled vertices (which makes
elements see their edges and faces in every orientation).  A space and graph
built on the device must then give the same assembled values and solution,
bit for bit, as the ones uploaded from the host."""
import numpy as np
import pytest

import feelpp_amd as fa
from helpers import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return fa.Context(0)


def eight_elements():
    g = golden("eight_elements_mesh.json")
    coords = np.array([n[1:3] for n in g["nodes"]])
    elems = np.array([e["nodes"] for e in g["elements"] if e["type"] == 2], np.int32) - 1
    return coords, elems


def relabel(coords, elems, seed):
    """the same mesh with vertex ids permuted at random"""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(coords.shape[0])          # old -> new
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    return coords[inv], perm[elems].astype(np.int32)


def check(ctx, orc, coords, elems, order, ncomp):
    dim = coords.shape[1]
    sp = fa.Space.from_mesh(ctx, coords, elems, order, ncomp)
    e2d, ndof = orc.doftable(dim, order, ncomp, elems)
    assert sp.ndof == ndof
    assert np.array_equal(sp.elem2dof, e2d)
    g = fa.Graph.from_space(sp)
    ia, ja = orc.graph(e2d, ndof)
    assert np.array_equal(g.ia, ia) and np.array_equal(g.ja, ja)
    return sp, g


@pytest.mark.parametrize("order", [1, 2])
def test_eight_elements(ctx, orc, order):
    coords, elems = eight_elements()
    check(ctx, orc, coords, elems, order, 1)
    check(ctx, orc, coords, elems, order, 2)


@pytest.mark.parametrize("order", [1, 2])
def test_square(ctx, orc, order):
    coords, elems = orc.mesh_square(9, jitter=0.2, seed=3)
    check(ctx, orc, coords, elems, order, 1)
    check(ctx, orc, *relabel(coords, elems, 11), order, 2)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_kuhn(ctx, orc, order):
    coords, elems = orc.mesh_kuhn(3, 4, 2, jitter=0.1, seed=5)
    check(ctx, orc, coords, elems, order, 1)
    rc, re = relabel(coords, elems, 17 + order)
    check(ctx, orc, rc, re, order, 1)
    check(ctx, orc, rc, re, order, 3)


def test_shuffled_elements(ctx, orc):
    # element order drives first touch: a shuffled element list renumbers
    coords, elems = orc.mesh_kuhn(4, jitter=0.1, seed=2)
    rng = np.random.default_rng(4)
    check(ctx, orc, coords, np.ascontiguousarray(elems[rng.permutation(len(elems))]), 3, 1)


def test_device_built_setup_assembles_and_solves_identically(ctx, orc):
    coords, elems = orc.mesh_kuhn(6, jitter=0.15, seed=9)
    host_sp = fa.Space(ctx, coords, elems, 2)
    host_g = fa.Graph(host_sp)
    dev_sp = fa.Space.from_mesh(ctx, coords, elems, 2)
    dev_g = fa.Graph.from_space(dev_sp)
    k = ctx.vector(len(elems), orc.lognormal(len(elems), 0.0, 1.0, 3))
    out = []
    for sp, g in ((host_sp, host_g), (dev_sp, dev_g)):
        A = fa.Matrix(g, fa.FP64)
        A.assemble(fa.DIFFUSION, 2, kdiff=k)
        b = ctx.vector(sp.ndof)
        fa.assemble_linear(sp, b, fa.SRC_CONST, 2, (1.0,))
        rows = fa.Rows(ctx, sp.ndof, np.nonzero(fa.boundary_dofs(3, 2, 1, coords, elems, sp.elem2dof, sp.ndof))[0])
        A.zero_rows(rows, None, b, mode=fa.ELIMINATION_SYMMETRIC)
        x = ctx.vector(sp.ndof)
        st = A.solve(b, x, rtol=1e-10, maxit=20000)
        assert st["converged"], st
        out.append((A.values(), b.get(), x.get()))
    for u, v in zip(*out):
        assert np.array_equal(u, v)
