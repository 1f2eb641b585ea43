"""Pins the oracle's quadrature tables, node sets and Lagrange basis.

Quadrature exactness mirrors testsuite/feelpoly/test_im.cpp:579-720 (closed-form
simplex monomial integrals, 5e-13); node sets are checked against Feel++'s own
fixtures data/operatorlagrangep1/simplex-3d-lagrange-order{2,3-warpblend}.msh;
the basis against the nodal property (testsuite/feelpoly/testlagrange.hpp).
"""
from fractions import Fraction
from itertools import product
from math import comb, factorial

import numpy as np
import pytest

from helpers import barycentric_of_ref, golden


def exact_ref_monomial(exps):
    """Integral of prod x_i^a_i over the [-1,1] reference simplex, exactly.

    x_i = 2*lam_i - 1 on the unit simplex (Jacobian 2^d); Dirichlet's formula
    int lam^k = prod k_i! / (sum k + d)!."""
    d = len(exps)
    total = Fraction(0)
    for ks in product(*[range(a + 1) for a in exps]):
        coef = Fraction(1)
        for a, k in zip(exps, ks):
            coef *= comb(a, k) * (2 ** k) * ((-1) ** (a - k))
        num = 1
        for k in ks:
            num *= factorial(k)
        total += coef * Fraction(num, factorial(sum(ks) + d))
    return float(total * 2 ** d)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", list(range(0, 9)))
def test_quadrature_exactness(orc, dim, order):
    pts, w = orc.quadrature(dim, order)
    assert abs(w.sum() - (2.0 if dim == 2 else 4.0 / 3.0)) < 1e-13
    for exps in product(range(order + 1), repeat=dim):
        if sum(exps) > order:
            continue
        q = np.sum(w * np.prod(pts ** np.array(exps), axis=1))
        ex = exact_ref_monomial(exps)
        assert abs(q - ex) < 5e-13 * max(1.0, abs(ex)), (dim, order, exps, q, ex)


def test_quadrature_keast_negative_weight(orc):
    # imsimplex.cpp:1922: the order-4 tetrahedron rule has a negative centroid weight
    pts, w = orc.quadrature(3, 4)
    assert len(w) == 11 and w[0] < 0 and np.allclose(pts[0], -0.5)


@pytest.mark.parametrize("fixture,dim,order,tol", [
    ("nodes_tet_p2.json", 3, 2, 1e-15),
    ("nodes_tet_p3_warpblend.json", 3, 3, 5e-15),
    ("nodes_tri_p2.json", 2, 2, 1e-15),
])
def test_nodes_match_reference_fixture(orc, fixture, dim, order, tol):
    g = golden(fixture)
    ref = np.array([n[1:1 + dim] for n in g["nodes"]])
    nodes = orc.lagrange_nodes(dim, order)
    assert nodes.shape == ref.shape
    assert np.abs(nodes - ref).max() < tol


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_basis_nodal_property(orc, dim, order):
    nodes = orc.lagrange_nodes(dim, order)
    phi, dphi = orc.lagrange_eval(dim, order, nodes)
    assert np.abs(phi - np.eye(len(nodes))).max() < 1e-13


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_basis_partition_of_unity(orc, dim, order):
    rng = np.random.default_rng(1)
    lam = rng.dirichlet(np.ones(dim + 1), size=20)
    pts = 2 * lam[:, 1:] - 1
    phi, dphi = orc.lagrange_eval(dim, order, pts)
    assert np.abs(phi.sum(axis=0) - 1).max() < 1e-13
    assert np.abs(dphi.sum(axis=0)).max() < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_basis_p1_is_barycentric(orc, dim):
    rng = np.random.default_rng(2)
    pts = 2 * rng.dirichlet(np.ones(dim + 1), size=10)[:, 1:] - 1
    phi, dphi = orc.lagrange_eval(dim, 1, pts)
    lam = barycentric_of_ref(pts)
    assert np.abs(phi - lam.T).max() < 1e-14
    g = np.zeros((dim + 1, dim))
    g[0, :] = -0.5
    for i in range(dim):
        g[i + 1, i] = 0.5
    assert np.abs(dphi - g[:, None, :]).max() < 1e-14


def test_basis_reproduces_polynomials(orc):
    # P3 interpolation is exact on cubics: sum_i p(x_i) phi_i(y) = p(y)
    nodes = orc.lagrange_nodes(3, 3)
    rng = np.random.default_rng(3)
    pts = 2 * rng.dirichlet(np.ones(4), size=15)[:, 1:] - 1
    p = lambda x: x[:, 0] ** 3 - 2 * x[:, 0] * x[:, 1] * x[:, 2] + x[:, 2] ** 2 - 0.5
    phi, _ = orc.lagrange_eval(3, 3, pts)
    assert np.abs(p(nodes) @ phi - p(pts)).max() < 1e-12


def test_basis_p4_reproduces_quartics(orc):
    nodes = orc.lagrange_nodes(3, 4)
    rng = np.random.default_rng(5)
    pts = 2 * rng.dirichlet(np.ones(4), size=15)[:, 1:] - 1
    p = lambda x: x[:, 0] ** 4 - 3 * x[:, 1] ** 2 * x[:, 2] * x[:, 0] + x[:, 2] ** 3 - x[:, 1]
    phi, _ = orc.lagrange_eval(3, 4, pts)
    assert np.abs(p(nodes) @ phi - p(pts)).max() < 1e-11


def test_p4_node_set_structure(orc):
    """P4 warp-blend nodes: GLL points on edges (the code path the P3 fixture pins),
    three points per face on the face medians, weight t at the near vertex,
    interior centroid.  t follows from warpblend.hpp's warp factor, which for N=4
    is wf(r) = (8/3)(sqrt(3/7) - 1/2) r; the closed form below restates it."""
    nodes = orc.lagrange_nodes(3, 4)
    V = np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    E = [(1, 2), (2, 0), (0, 1), (0, 3), (1, 3), (2, 3)]
    F = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    g = np.sqrt(3 / 7)
    for e, (a, b) in enumerate(E):
        for m, s in enumerate([-g, 0.0, g]):
            assert np.allclose(nodes[4 + 3 * e + m], V[a] * (1 - s) / 2 + V[b] * (1 + s) / 2, atol=1e-15)
    wf = lambda r: 8.0 / 3.0 * (g - 0.5) * r
    alpha = 0.1001
    dy = np.sqrt(3) / 2 * 2 * (0.5 * wf(0.25) * (1 + alpha ** 2 / 16))
    t = (np.sqrt(3) * (0.5 / np.sqrt(3) + dy) + 1) / 3
    for f, vs in enumerate(F):
        for l in range(3):
            w = np.full(3, (1 - t) / 2)
            w[l] = t
            assert np.allclose(nodes[22 + 3 * f + l], w @ V[list(vs)], atol=1e-14)
    assert np.allclose(nodes[34], [-0.5, -0.5, -0.5])


def _bary_tri(pts):
    """Barycentric coordinates on the [-1,1] reference triangle (-1,-1), (1,-1), (-1,1)."""
    l1, l2 = (pts[:, 0] + 1) / 2, (pts[:, 1] + 1) / 2
    return np.stack([1 - l1 - l2, l1, l2], axis=1)


@pytest.mark.parametrize("name", ["nodes_tri_p4_equispaced.json", "nodes_tri_p4_fekete.json"])
def test_p4_triangle_fixtures_interior_order(name):
    # Feel++'s own P4 triangle node sets: vertices, 3 points per edge, then the
    # three interior points, the l-th nearest vertex l (the lattice order that
    # the 3D face points inherit through pointsBySubEntity, warpblend.hpp:100-118)
    g = golden(name)
    pts = np.array([n[1:3] for n in g["nodes"]])
    assert len(pts) == 15
    lam = _bary_tri(pts)
    assert np.allclose(lam[:3], np.eye(3), atol=1e-15)
    inner = lam[12:15]
    assert list(inner.argmax(axis=1)) == [0, 1, 2]
    for l in range(3):
        other = np.delete(inner[l], l)
        assert abs(other[0] - other[1]) < 1e-14          # on the median through vertex l
    if "equispaced" in name:
        assert np.allclose(inner.max(axis=1), 0.5, atol=1e-15)


def test_p4_face_points_follow_the_lattice_order(orc):
    # every face's three interior points in the order of the fixtures above:
    # point l on the median through the face's l-th vertex, nearest to it, and
    # off the opposite vertex of the tetrahedron
    nodes = orc.lagrange_nodes(3, 4)
    lam = np.hstack([1 - ((nodes + 1) / 2).sum(axis=1, keepdims=True), (nodes + 1) / 2])
    F = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
    for f, vs in enumerate(F):
        opp = ({0, 1, 2, 3} - set(vs)).pop()
        for l in range(3):
            w = lam[22 + 3 * f + l]
            assert abs(w[opp]) < 1e-15
            assert w.argmax() == vs[l]
            others = [w[v] for m, v in enumerate(vs) if m != l]
            assert abs(others[0] - others[1]) < 1e-15


@pytest.mark.parametrize("t", [0.5, 0.6])
def test_p4_solution_independent_of_face_node_position(orc, t):
    # The P4 space does not depend on where the face nodes sit, so the Galerkin
    # solution is the same function whatever the face weight; its values at the
    # vertices and edge nodes (shared by both node sets) must agree.  This pins
    # the P4 face-dof numbering (a non-conforming face numbering makes the two
    # solutions differ) and the basis built on the face nodes, leaving only the
    # node position itself as a convention.
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    coords, elems = orc.mesh_kuhn(2, jitter=0.15, seed=41)
    e2d, ndof = orc.doftable(3, 4, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    k = orc.lognormal(len(elems), 0.0, 1.0, 42)
    isb = orc.boundary_dofs(3, 4, 1, coords, elems, e2d, ndof)

    def solve():
        val = orc.assemble_scalar(3, 4, 6, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k, mass=0.5)
        b = orc.assemble_rhs(3, 4, 1, 8, coords, elems, e2d, ndof, orc.FK_SINSIN, [1.0])
        orc.dirichlet(1, ia, ja, val, b, isb)
        return spla.spsolve(sp.csr_matrix((val, ja, ia), shape=(ndof, ndof)).tocsc(), b)

    x_ref = solve()
    try:
        orc.set_p4_face_weight(t)
        x_t = solve()
    finally:
        orc.set_p4_face_weight(0.0)
    ve = np.unique(e2d[:, :22])                     # vertex and edge dofs (local 0..21)
    fa = np.setdiff1d(np.unique(e2d[:, 22:34]), np.nonzero(isb)[0])   # interior face dofs
    assert np.abs(x_t[ve] - x_ref[ve]).max() < 1e-10 * np.abs(x_ref).max()
    assert np.abs(x_t[fa] - x_ref[fa]).max() > 1e-6 * np.abs(x_ref).max()   # the bases do differ


def test_quadrature_table_copies_agree():
    # the oracle and the product each compile a copy of the quadrature tables
    # transcribed from imsimplex.cpp by gen_im_tables.py; apart from the line
    # naming the copy they must be the same data
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    a = open(os.path.join(root, "oracle", "im_simplex_tables.inc")).read().splitlines()
    b = open(os.path.join(root, "feelpp_amd", "csrc", "host", "im_simplex_tables.inc")).read().splitlines()
    assert len(a) == len(b) > 100
    diff = [i for i, (x, y) in enumerate(zip(a, b)) if x != y]
    assert len(diff) <= 1 and all("Generated by" in a[i] for i in diff), diff
