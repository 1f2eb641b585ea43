"""The Feel++-shaped C++ surface (form2/form1/integrate/on/solve) end to end:
the quickstart drivers run on the GPU and their solutions are compared with
the oracle's FP64 assembly and tightly converged Jacobi-CG."""
import json
import os
import subprocess

import numpy as np
import pytest

import feelpp_amd as fa

from helpers import rel_fro

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "feelpp_amd", "bin")


def run(app, args, tmp_path):
    out = os.path.join(str(tmp_path), "run")
    r = subprocess.run([os.path.join(BIN, app)] + [str(a) for a in args] + ["--out", out], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    info = json.loads(r.stdout.strip().splitlines()[-1])
    return info, np.fromfile(out + ".sol", dtype=np.float64)


def oracle_poisson(orc, dim, n, order, qo, rhs_q, kappa_seed, source, mass, jitter, seed):
    coords, elems = orc.mesh_square(n, jitter, seed) if dim == 2 else orc.mesh_kuhn(n, jitter=jitter, seed=seed)
    e2d, ndof = orc.doftable(dim, order, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    k = orc.lognormal(len(elems), 0.0, 1.0, kappa_seed) if kappa_seed is not None else None
    val = orc.assemble_scalar(dim, order, qo, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k, mass=mass)
    b = orc.assemble_rhs(dim, order, 1, rhs_q, coords, elems, e2d, ndof, source, [1.0])
    isb = orc.boundary_dofs(dim, order, 1, coords, elems, e2d, ndof)
    orc.dirichlet(1, ia, ja, val, b, isb)
    x, its, _, reason = orc.pcg(ia, ja, val, b, rtol=1e-14, maxit=20000)
    assert reason > 0
    return x


def test_qs_laplacian_p2_lognormal(orc, tmp_path):
    info, x = run("feelpp_qs_laplacian", ["--dim", 3, "--n", 6, "--order", 2, "--seed", 5, "--jitter", 0.15,
                                          "--rtol", 1e-12], tmp_path)
    assert info["form"] == "diffusion" and info["converged"] == 1
    xo = oracle_poisson(orc, 3, 6, 2, 2, 2, 6, orc.FK_CONST, 0.0, 0.15, 5)
    assert rel_fro(x, xo) < 1e-10


def test_qs_laplacian_c1_like(orc, tmp_path):
    info, x = run("feelpp_qs_laplacian", ["--dim", 2, "--n", 32, "--order", 1, "--seed", 1, "--jitter", 0.2,
                                          "--kappa", "one", "--source", "sinsin", "--mass", 1.0, "--rtol", 1e-12],
                  tmp_path)
    assert info["form"] == "mass+diffusion"
    xo = oracle_poisson(orc, 2, 32, 1, 2, 2, None, orc.FK_SINSIN, 1.0, 0.2, 1)
    assert rel_fro(x, xo) < 1e-10
    # post-solve evaluation vs the oracle's norms of its own solution, and the
    # O(h^2) L2 error against the exact solution sin(pi x) sin(pi y) / (2 pi^2 + 1)
    coords, elems = orc.mesh_square(32, jitter=0.2, seed=1)
    e2d, _ = orc.doftable(2, 1, 1, elems)
    assert abs(info["norm_l2"] - orc.norm(2, 1, 1, 2, coords, elems, e2d, xo)) < 1e-9 * info["norm_l2"]
    assert abs(info["norm_semi_h1"] - orc.norm(2, 1, 1, 2, coords, elems, e2d, xo, semi=True)) < 1e-9 * info["norm_semi_h1"]
    exact = [(orc.FN_SINSIN, [1.0 / (2 * np.pi ** 2 + 1.0), 0, 0, 0])]
    assert abs(info["l2_error"] - orc.norm(2, 1, 1, 4, coords, elems, e2d, xo, exact)) < 1e-6 * info["l2_error"]
    assert info["l2_error"] < 5e-3 * info["norm_l2"]


def test_qs_laplacian_mixed_precision(orc, tmp_path):
    info, x = run("feelpp_qs_laplacian", ["--dim", 3, "--n", 6, "--order", 2, "--seed", 5, "--jitter", 0.15,
                                          "--rtol", 1e-11, "--precision", "fp32"], tmp_path)
    assert info["converged"] == 1 and info["outer"] >= 2
    xo = oracle_poisson(orc, 3, 6, 2, 2, 2, 6, orc.FK_CONST, 0.0, 0.15, 5)
    assert rel_fro(x, xo) < 1e-8


@pytest.mark.parametrize("order", [1, 2])
def test_qs_elasticity_beam(orc, tmp_path, order):
    info, x = run("feelpp_qs_elasticity", ["--nx", 8, "--ny", 2, "--nz", 2, "--order", order, "--seed", 7,
                                           "--rtol", 1e-12], tmp_path)
    assert info["form"] == "elasticity" and info["converged"] == 1
    coords, elems = orc.mesh_kuhn(8, 2, 2, lx=1.0, ly=0.25, lz=0.25, jitter=0.0, seed=7)
    e2d, ndof = orc.doftable(3, order, 3, elems)
    ia, ja = orc.graph(e2d, ndof)
    E = orc.uniform(len(elems), 1.0, 100.0, 8, log=True)
    nu = orc.uniform(len(elems), 0.2, 0.35, 9)
    val = orc.assemble_elasticity(order, 2 * order, coords, elems, e2d, E, nu, ia, ja)
    b = orc.assemble_rhs(3, order, 3, 2 * order, coords, elems, e2d, ndof, orc.FK_VECTOR, [0.0, 0.0, -1.0])
    isb = orc.boundary_dofs(3, order, 3, coords, elems, e2d, ndof, sel=1)
    orc.dirichlet(1, ia, ja, val, b, isb)
    xo, _, _, reason = orc.pcg(ia, ja, val, b, rtol=1e-14, maxit=100000)
    assert rel_fro(x, xo) < 1e-10


def test_qs_laplacian_gmsh_export(orc, tmp_path):
    # --export writes the mesh and the solution's vertex values as Feel++'s Gmsh
    # exporter lays them out; they match the raw solution at the vertex dofs
    path = str(tmp_path / "u.msh")
    info, x = run("feelpp_qs_laplacian", ["--dim", 3, "--n", 4, "--order", 2, "--seed", 5, "--jitter", 0.15,
                                          "--rtol", 1e-12, "--export", path], tmp_path)
    coords, elems = fa.read_msh22(path, 3)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    assert ndof == x.shape[0]
    text = open(path).read()
    block = text[text.index("$ElementNodeData"):text.index("$EndElementNodeData")].split("\n")[1:-1]
    assert block[1] == '"u"' and int(block[7]) == len(elems)
    for e, line in enumerate(block[8:]):
        vals = np.array([float(t) for t in line.split()[2:]])
        assert np.array_equal(vals, x[e2d[e, :4]])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_heat_toolbox_backward_euler(orc, tmp_path, precision):
    # BDF1 heat steps as the heat toolbox assembles them, against the same
    # steps on the oracle's matrices with a direct solver
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    dim, n, order, dt, steps, q = 2, 12, 2, 0.02, 5, 3.0
    info, x = run("feelpp_heat", ["--dim", dim, "--n", n, "--order", order, "--seed", 4, "--jitter", 0.1, "--dt", dt,
                                  "--steps", steps, "--q", q, "--rtol", 1e-12, "--precision", precision], tmp_path)
    assert info["converged"] == 1 and info["form"] == "mass+diffusion"
    coords, elems = orc.mesh_square(n, jitter=0.1, seed=4)
    e2d, ndof = orc.doftable(dim, order, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    va = orc.assemble_scalar(dim, order, 2 * order, coords, elems, e2d, ia, ja, diff=1.0, mass=1.0 / dt)
    vm = orc.assemble_scalar(dim, order, 2 * order, coords, elems, e2d, ia, ja, diff=0.0, mass=1.0)
    F = orc.assemble_rhs(dim, order, 1, order, coords, elems, e2d, ndof, orc.FK_CONST, [q])
    isb = orc.boundary_dofs(dim, order, 1, coords, elems, e2d, ndof)
    orc.dirichlet(1, ia, ja, va, np.zeros(ndof), isb)
    A = sp.csr_matrix((va, ja, ia), shape=(ndof, ndof)).tocsc()
    M = sp.csr_matrix((vm, ja, ia), shape=(ndof, ndof))
    T = np.zeros(ndof)
    for _ in range(steps):
        b = F + M @ T / dt
        b[isb.astype(bool)] = 0.0
        T = spla.spsolve(A, b)
    assert rel_fro(x, T) < (1e-9 if precision == "fp64" else 1e-8)
    assert abs(info["norm_l2"] - orc.norm(dim, order, 1, 2 * order, coords, elems, e2d, T)) < 1e-8 * info["norm_l2"]
