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
