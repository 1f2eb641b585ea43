"""Host side of the product vs the oracle (CPU): meshes, Feel++ DofTable
numbering, Stencil CSR pattern and boundary-dof sets must be bit-exact;
node sets, quadrature and reference tensors must agree to round-off."""
import os

import numpy as np
import pytest

import feelpp_amd as fa


@pytest.mark.parametrize("n,jit,seed", [(4, 0.0, 0), (16, 0.2, 7), (128, 0.2, 1)])
def test_square_mesh_bit_exact(orc, n, jit, seed):
    c1, e1 = fa.mesh_unit_square(n, jit, seed)
    c2, e2 = orc.mesh_square(n, jit, seed)
    assert (e1 == e2).all() and np.array_equal(c1, c2)


@pytest.mark.parametrize("dims", [(3, 3, 3, 1, 1, 1), (5, 2, 3, 1.0, 0.25, 0.25), (12, 12, 12, 1, 1, 1)])
def test_kuhn_mesh_bit_exact(orc, dims):
    nx, ny, nz, lx, ly, lz = dims
    c1, e1 = fa.mesh_kuhn_box(nx, ny, nz, lx, ly, lz, 0.15, 42)
    c2, e2 = orc.mesh_kuhn(nx, ny, nz, lx, ly, lz, 0.15, 42)
    assert (e1 == e2).all() and np.array_equal(c1, c2)


def test_random_fields_bit_exact(orc):
    assert np.array_equal(fa.lognormal(1001, 0.0, 1.0, 9), orc.lognormal(1001, 0.0, 1.0, 9))
    assert np.array_equal(fa.uniform(777, 1.0, 100.0, 3, log=True), orc.uniform(777, 1.0, 100.0, 3, log=True))


CASES = [(2, 1, 1), (2, 2, 1), (2, 1, 2), (3, 1, 1), (3, 2, 1), (3, 3, 1), (3, 4, 1), (3, 1, 3), (3, 2, 3), (3, 3, 3),
         (3, 4, 3)]


def mesh_for(orc, dim, n=5):
    return orc.mesh_square(n, 0.2, 3) if dim == 2 else orc.mesh_kuhn(n, n - 1, n - 2, jitter=0.15, seed=3)


@pytest.mark.parametrize("dim,order,ncomp", CASES)
def test_doftable_graph_boundary_bit_exact(orc, dim, order, ncomp):
    coords, elems = mesh_for(orc, dim)
    e2d_o, nd_o = orc.doftable(dim, order, ncomp, elems)
    e2d_p, nd_p = fa.doftable(dim, order, ncomp, elems, coords.shape[0])
    assert nd_o == nd_p and np.array_equal(e2d_o, e2d_p)
    ia_o, ja_o = orc.graph(e2d_o, nd_o)
    ia_p, ja_p = fa.graph(e2d_p, nd_p)
    assert np.array_equal(ia_o, ia_p) and np.array_equal(ja_o, ja_p)
    for sel in (0, 1):
        b_o = orc.boundary_dofs(dim, order, ncomp, coords, elems, e2d_o, nd_o, sel)
        b_p = fa.boundary_dofs(dim, order, ncomp, coords, elems, e2d_p, nd_p, sel)
        assert np.array_equal(b_o, b_p)


def test_p2_kuhn_graph_large_bit_exact(orc):
    # a 12^3 P2 space (10 368 tets, ~15k dofs): pattern from both implementations
    coords, elems = orc.mesh_kuhn(12, jitter=0.15, seed=2)
    e2d_o, nd_o = orc.doftable(3, 2, 1, elems)
    e2d_p, nd_p = fa.doftable(3, 2, 1, elems, coords.shape[0])
    assert np.array_equal(e2d_o, e2d_p)
    ia_o, ja_o = orc.graph(e2d_o, nd_o)
    ia_p, ja_p = fa.graph(e2d_p, nd_p)
    assert np.array_equal(ia_o, ia_p) and np.array_equal(ja_o, ja_p)
    n = 12
    assert len(ja_p) == 230 * n ** 3 + 138 * n ** 2 + 24 * n + 1


def test_c1_full_size_pattern_bit_exact(orc):
    # the 128x128 P1 triangle space (32 768 elements, 16 641 dofs, 115 457 nnz)
    coords, elems = orc.mesh_square(128, 0.2, 1)
    e2d_o, nd_o = orc.doftable(2, 1, 1, elems)
    e2d_p, nd_p = fa.doftable(2, 1, 1, elems, coords.shape[0])
    ia_o, ja_o = orc.graph(e2d_o, nd_o)
    ia_p, ja_p = fa.graph(e2d_p, nd_p)
    assert nd_p == 16641 and len(ja_p) == 115457
    assert np.array_equal(e2d_o, e2d_p) and np.array_equal(ia_o, ia_p) and np.array_equal(ja_o, ja_p)


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_reference_element_agrees(orc, dim, order):
    assert np.abs(fa.lagrange_nodes(dim, order) - orc.lagrange_nodes(dim, order)).max() < 1e-15
    for q in range(0, 9):
        p1, w1 = fa.quadrature(dim, q)
        p2, w2 = orc.quadrature(dim, q)
        assert np.array_equal(p1, p2) and np.array_equal(w1, w2)
    # reference tensors vs the oracle's basis tabulated at the same quadrature
    qo = 2 * order
    R, M, m = fa.reference_tensors(dim, order, qo)
    pts, w = orc.quadrature(dim, qo)
    phi, dphi = orc.lagrange_eval(dim, order, pts)
    Rr = np.einsum("q,iqa,jqb->abij", w, dphi, dphi).reshape(dim * dim, phi.shape[0], phi.shape[0])
    assert np.abs(R - Rr).max() < 1e-13 * max(1, np.abs(Rr).max())
    assert np.abs(M - np.einsum("q,iq,jq->ij", w, phi, phi)).max() < 1e-14
    assert np.abs(m - phi @ w).max() < 1e-14


def test_msh22_roundtrip(tmp_path):
    coords, elems = fa.mesh_kuhn_box(3, 2, 2, jitter=0.1, seed=4)
    p = os.path.join(str(tmp_path), "m.msh")
    fa.write_msh22(p, coords, elems)
    c2, e2 = fa.read_msh22(p, 3)
    assert np.array_equal(elems, e2) and np.abs(coords - c2).max() < 1e-15


def test_host_rejects_bad_input():
    with pytest.raises(fa.FeelHipError):
        fa.doftable(3, 2, 1, np.array([[0, 1, 2, 99]], np.int32), 4)
    with pytest.raises(fa.FeelHipError):
        fa.doftable(3, 9, 1, np.array([[0, 1, 2, 3]], np.int32), 4)


def test_cg_condition_estimate(orc):
    # Lanczos estimate from the coefficients of a Jacobi-CG run converges to
    # kappa(D^{-1/2} A D^{-1/2}) computed by a dense eigensolver
    coords, elems = orc.mesh_kuhn(3, jitter=0.2, seed=5)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    k = orc.lognormal(len(elems), 0.0, 1.0, 9)
    val = orc.assemble_scalar(3, 2, 2, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k, mass=0.3)
    A = np.zeros((ndof, ndof))
    for r in range(ndof):
        A[r, ja[ia[r]:ia[r + 1]]] = val[ia[r]:ia[r + 1]]
    d = np.diag(A)
    S = A / np.sqrt(np.outer(d, d))
    ev = np.linalg.eigvalsh(S)
    kappa = ev[-1] / ev[0]
    b = np.sin(np.arange(ndof) * 0.37) + 0.1
    x, r = np.zeros(ndof), b.copy()
    z = r / d
    p = z.copy()
    gam, pap = [z @ r], []
    for _ in range(ndof):
        w = A @ p
        pap.append(p @ w)
        a = gam[-1] / pap[-1]
        x += a * p
        r -= a * w
        z = r / d
        gam.append(z @ r)
        if gam[-1] <= 1e-28 * gam[0]:
            break
        p = z + gam[-1] / gam[-2] * p
    est = fa.cg_condition_estimate(gam, pap)
    assert abs(est - kappa) <= 1e-6 * kappa
    # few iterations: a lower bound
    assert fa.cg_condition_estimate(gam[:6], pap[:5]) <= kappa * (1 + 1e-12)
    with pytest.raises(RuntimeError):
        fa.cg_condition_estimate([1.0, -1.0, 1.0], [1.0, 1.0])


@pytest.mark.parametrize("dim,order,ncomp", [(2, 1, 1), (2, 2, 1), (3, 2, 1), (3, 3, 3)])
def test_msh22_solution_export(orc, tmp_path, dim, order, ncomp):
    # Feel++'s Gmsh exporter layout: the mesh, then $ElementNodeData with one
    # line per element "id nvert v..." (values at the element's vertices)
    if dim == 2:
        coords, elems = fa.mesh_unit_square(4, jitter=0.1, seed=2)
    else:
        coords, elems = fa.mesh_kuhn_box(2, jitter=0.1, seed=2)
    e2d, ndof = fa.doftable(dim, order, ncomp, elems, coords.shape[0])
    vals = np.cos(np.arange(ndof) * 0.7)
    path = str(tmp_path / "sol.msh")
    fa.write_msh22_solution(path, coords, elems, order, e2d, vals, ncomp=ncomp, name="u", time=0.5, step=3)
    c2, e2 = fa.read_msh22(path, dim)             # the mesh part still reads back
    assert np.array_equal(e2, elems) and np.allclose(c2, coords, rtol=0, atol=1e-15)
    text = open(path).read()
    block = text[text.index("$ElementNodeData"):text.index("$EndElementNodeData")].split("\n")[1:-1]
    assert block[0] == "1" and block[1] == '"u"' and float(block[3]) == 0.5
    assert int(block[5]) == 3 and int(block[6]) == ncomp and int(block[7]) == len(elems)
    nloc = e2d.shape[1] // ncomp
    for e, line in enumerate(block[8:]):
        f = line.split()
        assert int(f[0]) == e + 1 and int(f[1]) == dim + 1
        got = np.array([float(t) for t in f[2:]]).reshape(dim + 1, ncomp)
        want = np.array([[vals[e2d[e, c * nloc + v]] for c in range(ncomp)] for v in range(dim + 1)])
        assert np.array_equal(got, want)
