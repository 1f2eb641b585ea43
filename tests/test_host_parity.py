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


MSH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "msh")


def _msh_sections(path):
    """Nodes (id -> xyz) and elements (type, node ids) of an MSH 2.2 text file, read plainly."""
    lines = open(path).read().split("\n")
    nodes, elems = {}, []
    i = lines.index("$Nodes")
    for ln in lines[i + 2:i + 2 + int(lines[i + 1])]:
        t = ln.split()
        nodes[int(t[0])] = [float(v) for v in t[1:4]]
    i = lines.index("$Elements")
    for ln in lines[i + 2:i + 2 + int(lines[i + 1])]:
        t = [int(v) for v in ln.split()]
        elems.append((t[1], t[3 + t[2]:]))
    return nodes, elems


def test_msh22_reads_feelpp_eight_elements_mesh():
    # Feel++'s own 2D test mesh (points, lines and triangles, no tags): point id =
    # file id (importergmsh.hpp:2040-2066), only the triangles become elements,
    # element id = their count in file order (addVolume/addFace: mesh->elements().size())
    path = os.path.join(MSH, "eight_elements_mesh.msh")
    coords, elems = fa.read_msh22(path, 2)
    nodes, raw = _msh_sections(path)
    assert coords.shape == (9, 2) and elems.shape == (8, 3)
    assert np.array_equal(coords, np.array([nodes[i][:2] for i in sorted(nodes)]))
    assert elems.tolist() == [[v - 1 for v in vs] for t, vs in raw if t == 2]
    # the numbering built on it is the hand-derived first-touch table
    e2d, ndof = fa.doftable(2, 1, 1, elems, 9)
    assert ndof == 9
    assert e2d.tolist() == [[0, 1, 2], [1, 3, 4], [2, 4, 5], [4, 6, 7], [1, 4, 2], [3, 6, 4], [4, 7, 5], [6, 8, 7]]
    e2d2, ndof2 = fa.doftable(2, 2, 1, elems, 9)
    assert ndof2 == 25 and e2d2[4].tolist() == [1, 7, 2, 14, 3, 9]


@pytest.mark.parametrize("name,verts", [("one-elt-real-rotx.msh", [2, 3, 0, 1]), ("tetra.msh", [0, 2, 3, 1])])
def test_msh22_reads_feelpp_tetrahedra(name, verts, orc):
    # one-tetrahedron meshes of Feel++ with $PhysicalNames, tagged faces and
    # edges, and vertex orders that are not ascending: the element keeps the
    # file's vertex order (GmshOrdering is the identity on P1 simplices), so the
    # edge orientations and the P2/P3 numbering follow it
    path = os.path.join(MSH, name)
    coords, elems = fa.read_msh22(path, 3)
    nodes, raw = _msh_sections(path)
    assert np.array_equal(coords, np.array([nodes[i] for i in sorted(nodes)]))
    assert elems.tolist() == [verts]
    for order in (2, 3):
        e2d, ndof = fa.doftable(3, order, 1, elems, 4)
        e2o, ndo = orc.doftable(3, order, 1, elems)
        assert ndof == ndo and np.array_equal(e2d, e2o)


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
    # the solver's threshold test agrees with the estimate away from the limit
    for n in (4, 6, 12, len(pap)):
        e = fa.cg_condition_estimate(gam[:n + 1], pap[:n])
        for lim in (e / 3, e / 1.05, e * 1.05, e * 3):
            assert fa.cg_condition_exceeds(gam[:n + 1], pap[:n], lim) == (e > lim), (n, e, lim)


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


def test_p2_stiffness_tensors_exact_at_degree_two():
    # The element-by-element FP32 operator of P2 diffusion integrates
    # grad phi_i . F grad phi_j with the 4-point degree-2 rule: the integrand is
    # quadratic, so that rule gives the same reference tensors as the rules the
    # assembly uses (order 4, 11-point Keast with its negative weight; order 6)
    R2, _, _ = fa.reference_tensors(3, 2, 2)
    p, w = fa.quadrature(3, 2)
    assert len(w) == 4
    for qo in (4, 6):
        Rq, _, _ = fa.reference_tensors(3, 2, qo)
        assert np.max(np.abs(R2 - Rq)) < 1e-13 * np.max(np.abs(Rq))
    # stiffness rows sum to zero (constants in the kernel) for every a, b
    assert np.max(np.abs(R2.sum(axis=2))) < 1e-13
    # the operator's closed form: nodal P2 basis phi_v = l_v (2 l_v - 1),
    # phi_(a,b) = 4 l_a l_b on the [-1,1] tetrahedron with the edge order
    # (1,2) (0,2) (0,1) (0,3) (1,3) (2,3), at the symmetric 4-point rule
    # (barycentric (a, b, b, b), weights 1/3) gives the reference tensors
    alpha, beta = (5 + 3 * np.sqrt(5)) / 20, (5 - np.sqrt(5)) / 20
    edges = [(1, 2), (0, 2), (0, 1), (0, 3), (1, 3), (2, 3)]
    glam = np.array([[-0.5, -0.5, -0.5], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]])
    R = np.zeros((3, 3, 10, 10))
    for q in range(4):
        lam = np.full(4, beta)
        lam[q] = alpha
        G = np.zeros((10, 3))
        for v in range(4):
            G[v] = (4 * lam[v] - 1) * glam[v]
        for k, (a, b) in enumerate(edges):
            G[4 + k] = 4 * (lam[a] * glam[b] + lam[b] * glam[a])
        R += (1.0 / 3.0) * np.einsum("ia,jb->abij", G, G)
    R4, _, _ = fa.reference_tensors(3, 2, 4)
    assert np.max(np.abs(R.reshape(9, 10, 10) - R4)) < 1e-13 * np.max(np.abs(R4))
    # and the nodes are the vertices and those edge midpoints
    nodes = fa.lagrange_nodes(3, 2)
    ref = np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float)
    assert np.allclose(nodes[:4], ref)
    for k, (a, b) in enumerate(edges):
        assert np.allclose(nodes[4 + k], 0.5 * (ref[a] + ref[b]))
