"""Pins the oracle's dof numbering (DofTable first touch) and CSR pattern (Stencil).

Hand-derived first-touch numbering on Feel++'s 8-triangle fixture
(testsuite/feeldiscr/eight_elements_mesh.msh), geometric conformity of the
numbering (one global dof = one physical node), and brute-force dense
neighbour sets / stencil row counts for the pattern.
"""
import numpy as np
import pytest

from helpers import dof_positions, golden


def eight_elements():
    g = golden("eight_elements_mesh.json")
    coords = np.array([n[1:3] for n in g["nodes"]])
    tris = [e["nodes"] for e in g["elements"] if e["type"] == 2]
    # point ids are the file's node ids (importergmsh.hpp:2050), 1-based here
    elems = np.array(tris, np.int32) - 1
    return coords, elems


def test_eight_elements_p1_first_touch(orc):
    coords, elems = eight_elements()
    e2d, ndof = orc.doftable(2, 1, 1, elems)
    assert ndof == 9
    expect = [[0, 1, 2], [1, 3, 4], [2, 4, 5], [4, 6, 7], [1, 4, 2], [3, 6, 4], [4, 7, 5], [6, 8, 7]]
    assert e2d.tolist() == expect


def test_eight_elements_p2_first_touch(orc):
    coords, elems = eight_elements()
    e2d, ndof = orc.doftable(2, 2, 1, elems)
    assert ndof == 25   # 9 vertices + 16 edges
    expect = [[0, 1, 2, 3, 4, 5], [1, 6, 7, 8, 9, 10], [2, 7, 11, 12, 13, 14], [7, 15, 16, 17, 18, 19],
              [1, 7, 2, 14, 3, 9], [6, 15, 7, 19, 8, 20], [7, 16, 11, 21, 12, 18],
              [15, 22, 16, 23, 17, 24]]
    assert e2d.tolist() == expect


def test_eight_elements_vector_interleave(orc):
    # vector spaces: local index nloc*c + l, global = base + c (doftable.hpp:1213-1300)
    coords, elems = eight_elements()
    e2d1, nd1 = orc.doftable(2, 1, 1, elems)
    e2d2, nd2 = orc.doftable(2, 1, 2, elems)
    assert nd2 == 2 * nd1
    assert (e2d2[:, :3] == 2 * e2d1).all() and (e2d2[:, 3:] == 2 * e2d1 + 1).all()


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_kuhn_numbering_is_conforming(orc, order):
    coords, elems = orc.mesh_kuhn(3, 2, 2, jitter=0.15, seed=5)
    e2d, ndof = orc.doftable(3, order, 1, elems)
    pos, mism = dof_positions(coords, elems, e2d, orc.lagrange_nodes(3, order), ndof)
    assert mism < 1e-14
    # distinct dofs sit at distinct points
    key = np.round(pos, 12)
    assert len(np.unique(key, axis=0)) == ndof
    nv, ne = len(coords), len(elems)
    edges = {tuple(sorted((a, b))) for t in elems for a, b in
             [(t[1], t[2]), (t[2], t[0]), (t[0], t[1]), (t[0], t[3]), (t[1], t[3]), (t[2], t[3])]}
    faces = {tuple(sorted(f)) for t in elems for f in
             [(t[1], t[2], t[3]), (t[0], t[2], t[3]), (t[0], t[1], t[3]), (t[0], t[1], t[2])]}
    expect = nv + (order - 1) * len(edges) + (order - 1) * (order - 2) // 2 * len(faces) + \
        (order - 1) * (order - 2) * (order - 3) // 6 * len(elems)
    assert ndof == expect


def test_first_touch_order_is_monotone(orc):
    # new dofs appear in element order: max dof of elements 0..e is non-decreasing
    coords, elems = orc.mesh_kuhn(2, seed=1)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    running = np.maximum.accumulate(e2d.max(axis=1))
    assert (np.diff(running) >= 0).all() and running[-1] == ndof - 1
    assert e2d[0].tolist() == list(range(10))


def dense_pattern(e2d, ndof):
    M = np.zeros((ndof, ndof), bool)
    for row in e2d:
        M[np.ix_(row, row)] = True
    np.fill_diagonal(M, True)
    return M


@pytest.mark.parametrize("dim,order,ncomp", [(2, 1, 1), (2, 2, 1), (3, 1, 1), (3, 2, 1), (3, 3, 1), (3, 4, 1), (3, 1, 3)])
def test_graph_matches_dense_neighbours(orc, dim, order, ncomp):
    if dim == 2:
        coords, elems = orc.mesh_square(4, jitter=0.2, seed=3)
    else:
        coords, elems = orc.mesh_kuhn(2, 2, 1, seed=3)
    e2d, ndof = orc.doftable(dim, order, ncomp, elems)
    ia, ja = orc.graph(e2d, ndof)
    M = dense_pattern(e2d, ndof)
    for r in range(ndof):
        cols = ja[ia[r]:ia[r + 1]]
        assert (np.diff(cols) > 0).all()
        assert cols.tolist() == np.nonzero(M[r])[0].tolist()


def test_p1_square_nnz(orc):
    n = 8
    coords, elems = orc.mesh_square(n)
    e2d, ndof = orc.doftable(2, 1, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    nedges = 3 * n * n + 2 * n   # structured triangulation
    assert ndof == (n + 1) ** 2 and len(ja) == ndof + 2 * nedges


@pytest.mark.parametrize("n", [3, 4])
def test_p2_kuhn_stencils(orc, n):
    # per-row stencils for P2 on Kuhn cubes: 65 (interior vertex), 27 (axis or
    # body-diagonal edge), 19 (face-diagonal edge); nnz = 230n^3+138n^2+24n+1
    coords, elems = orc.mesh_kuhn(n)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    assert len(ja) == 230 * n ** 3 + 138 * n ** 2 + 24 * n + 1
    pos, _ = dof_positions(coords, elems, e2d, orc.lagrange_nodes(3, 2), ndof)
    interior = np.all((pos > 1e-9) & (pos < 1 - 1e-9), axis=1)
    counts = np.diff(ia)
    frac = np.abs(pos * n - np.round(pos * n))
    is_vertex = np.all(frac < 1e-9, axis=1)
    half = np.abs(frac - 0.5) < 1e-9
    nhalf = half.sum(axis=1)
    assert set(counts[interior & is_vertex]) == {65}
    assert set(counts[interior & (nhalf == 1)]) == {27}   # axis edges
    assert set(counts[interior & (nhalf == 3)]) == {27}   # body diagonals
    assert set(counts[interior & (nhalf == 2)]) == {19}   # face diagonals
