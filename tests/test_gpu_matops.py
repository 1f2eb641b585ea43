"""The rest of the MatrixSparse / Vector interface on the device, through the
C-ABI, against numpy / scipy on the oracle's CSR values: scale, addMatrix(a, X)
on the same and on a containing pattern, copy, diagonal, transpose, l1 / linf /
Frobenius norms, energy, block insertion, entry reads, vector reductions,
setConstant, add(s), setVector / addVector with repeated indices, y += A x."""
import numpy as np
import pytest
import scipy.sparse as sp

import feelpp_amd as fa
from helpers import rel_fro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return fa.Context(0)


@pytest.fixture(scope="module")
def p2(ctx, orc):
    """P2 Kuhn 4^3 Poisson matrix after non-symmetric Dirichlet elimination
    (on.type = elimination: rows zeroed, columns kept), so A != A^T."""
    coords, elems = orc.mesh_kuhn(4, jitter=0.15, seed=31)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    k = orc.lognormal(len(elems), 0.0, 1.0, 32)
    val = orc.assemble_scalar(3, 2, 4, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k, mass=0.3)
    b = orc.assemble_rhs(3, 2, 1, 2, coords, elems, e2d, ndof, orc.FK_CONST, [1.0])
    isb = orc.boundary_dofs(3, 2, 1, coords, elems, e2d, ndof)
    orc.dirichlet(0, ia, ja, val, b, isb, None, 2.5)
    space = fa.Space(ctx, coords, elems, 2, 1, e2d, ndof)
    g = fa.Graph(space, ia, ja)
    A = fa.Matrix(g, fa.FP64_FP32)
    A.set_values(val)
    S = sp.csr_matrix((val, ja, ia), shape=(ndof, ndof))
    return dict(space=space, g=g, A=A, S=S, val=val, ia=ia, ja=ja, ndof=ndof, e2d=e2d)


def test_scale_copy_diagonal(ctx, p2):
    A, val = p2["A"], p2["val"]
    B = fa.Matrix(p2["g"], fa.FP64_FP32)
    B.copy_from(A)
    assert np.array_equal(B.values(), val)
    B.scale(-1.75)
    assert np.array_equal(B.values(), -1.75 * val)
    assert np.array_equal(B.values(fa.FP32), (-1.75 * val).astype(np.float32))
    assert np.array_equal(A.diagonal(), p2["S"].diagonal())


def test_axpy_same_pattern(ctx, p2):
    A, val = p2["A"], p2["val"]
    Y = fa.Matrix(p2["g"], fa.FP64_FP32)
    Y.copy_from(A)
    Y.axpy(0.5, A)
    assert np.array_equal(Y.values(), val + 0.5 * val)
    assert np.array_equal(Y.values(fa.FP32), (val + 0.5 * val).astype(np.float32))


def test_axpy_into_a_containing_pattern(ctx, orc):
    # X on the exact P1 pattern, Y on that pattern plus extra entries: SUBSET
    # adds X into Y's graph; the reverse direction is refused and leaves X as is
    coords, elems = orc.mesh_square(6, 0.1, 3)
    e2d, ndof = orc.doftable(2, 1, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    S1 = sp.csr_matrix((np.ones(len(ja)), ja, ia), shape=(ndof, ndof))
    extra = sp.csr_matrix((np.ones(ndof - 7), (np.arange(ndof - 7), np.arange(7, ndof))), shape=(ndof, ndof))
    S2 = (S1 + extra + extra.T).tocsr()
    S2.sort_indices()
    space = fa.Space(ctx, coords, elems, 1, 1, e2d, ndof)
    g1 = fa.Graph(space, ia, ja)
    g2 = fa.Graph(space, S2.indptr.astype(np.int32), S2.indices.astype(np.int32))
    X = fa.Matrix(g1)
    X.assemble(fa.DIFFUSION | fa.MASS, 2)
    Y = fa.Matrix(g2)
    rng = np.random.default_rng(4)
    y0 = rng.standard_normal(g2.nnz)
    Y.set_values(y0)
    Y.axpy(-2.0, X, fa.SUBSET_NONZERO_PATTERN)
    Xs = sp.csr_matrix((X.values(), ja, ia), shape=(ndof, ndof))
    ref = (sp.csr_matrix((y0, S2.indices, S2.indptr), shape=(ndof, ndof)) - 2.0 * Xs).tocsr()
    ref.sort_indices()
    got = sp.csr_matrix((Y.values(), S2.indices, S2.indptr), shape=(ndof, ndof))
    assert abs(got - ref).max() == 0.0
    x0 = X.values()
    with pytest.raises(fa.FeelHipError):
        X.axpy(1.0, Y, fa.DIFFERENT_NONZERO_PATTERN)
    with pytest.raises(fa.FeelHipError):
        Y.axpy(1.0, X, fa.SAME_NONZERO_PATTERN)
    assert np.array_equal(X.values(), x0)
    # a structurally non-symmetric graph has no transpose map
    S3 = (S1 + extra).tocsr()
    S3.sort_indices()
    g3 = fa.Graph(space, S3.indptr.astype(np.int32), S3.indices.astype(np.int32))
    Z, Zt = fa.Matrix(g3), fa.Matrix(g3)
    with pytest.raises(fa.FeelHipError):
        Z.transpose_into(Zt)
    with pytest.raises(fa.FeelHipError):
        Z.norm(fa.NORM_1)
    assert Z.norm(fa.NORM_INF) >= 0.0


def test_transpose_and_norms(ctx, p2):
    A, S = p2["A"], p2["S"]
    assert abs(S - S.T).max() > 0                      # elimination left the matrix non-symmetric
    At = fa.Matrix(p2["g"], fa.FP64_FP32)
    A.transpose_into(At)
    T = S.T.tocsr()
    T.sort_indices()
    assert np.array_equal(At.values(), T.data)
    assert np.array_equal(At.values(fa.FP32), T.data.astype(np.float32))
    absS = abs(S)
    linf = absS.sum(axis=1).max()
    l1 = absS.sum(axis=0).max()
    fro = np.sqrt((S.data ** 2).sum())
    assert abs(A.norm(fa.NORM_INF) - linf) <= 1e-14 * linf
    assert abs(A.norm(fa.NORM_1) - l1) <= 1e-14 * l1
    assert abs(A.norm(fa.NORM_FROBENIUS) - fro) <= 1e-14 * fro
    assert A.norm(fa.NORM_1) != A.norm(fa.NORM_INF)
    assert A.norm(fa.NORM_1) == At.norm(fa.NORM_INF)   # column sums of A are row sums of A^T, same order
    with pytest.raises(fa.FeelHipError):
        A.norm(7)


def test_energy(ctx, p2):
    A, S, n = p2["A"], p2["S"], p2["ndof"]
    rng = np.random.default_rng(8)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    du, dv = ctx.vector(n, u), ctx.vector(n, v)
    e, et = v @ (S @ u), v @ (S.T @ u)
    assert abs(A.energy(dv, du) - e) <= 1e-13 * abs(v) @ (abs(S) @ abs(u))
    assert abs(A.energy(dv, du, transpose=True) - et) <= 1e-13 * abs(v) @ (abs(S.T) @ abs(u))
    assert abs(e - et) > 1e-6 * abs(e)


def test_set_matrix_and_get_entries(ctx, p2):
    A, ia, ja = p2["A"], p2["ia"], p2["ja"]
    B = fa.Matrix(p2["g"], fa.FP64_FP32)
    B.copy_from(A)
    rows = p2["e2d"][3]
    blk = np.arange(100.0).reshape(10, 10) - 50.0
    B.set_matrix(rows, rows, blk)
    got = B.get_entries(np.repeat(rows, 10), np.tile(rows, 10))
    assert np.array_equal(got, blk.ravel())
    # repeated entries: INSERT keeps the last value
    B.set_matrix([rows[0], rows[0]], [rows[1]], [[1.0], [2.0]])
    assert B.get_entries([rows[0]], [rows[1]])[0] == 2.0
    # untouched entries unchanged; an entry outside the pattern reads 0 and cannot be set
    far = int(ja[ia[0]:ia[1]].max()) + 1
    while far in set(ja[ia[0]:ia[1]]):
        far += 1
    assert B.get_entries([0], [far])[0] == 0.0
    with pytest.raises(fa.FeelHipError):
        B.set_matrix([0], [far], [[1.0]])
    mask = np.ones(len(ja), bool)
    for r in rows:
        for c in list(rows):
            mask[ia[r] + np.searchsorted(ja[ia[r]:ia[r + 1]], c)] = False
    assert np.array_equal(B.values()[mask], A.values()[mask])


def test_vector_operations(ctx, p2):
    n, S = p2["ndof"], p2["S"]
    rng = np.random.default_rng(12)
    a = rng.standard_normal(n)
    v = ctx.vector(n, a)
    assert abs(v.sum() - a.sum()) <= 1e-13 * np.abs(a).sum()
    assert v.min() == a.min() and v.max() == a.max() and v.linfty_norm() == np.abs(a).max()
    assert abs(v.l1_norm() - np.abs(a).sum()) <= 1e-13 * np.abs(a).sum()
    assert abs(v.l2_norm() - np.linalg.norm(a)) <= 1e-13 * np.linalg.norm(a)
    v.shift(2.5)
    assert np.array_equal(v.get(), a + 2.5)
    v.set_constant(-1.0)
    assert np.array_equal(v.get(), -np.ones(n))
    # setVector: the last of repeated indices wins; addVector: repeats add in order
    idx = np.array([5, 9, 5, n - 1, 0], np.int32)
    vals = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    v.set_indexed(idx, vals)
    ref = -np.ones(n)
    for i, x in zip(idx, vals):
        ref[i] = x
    assert np.array_equal(v.get(), ref)
    v.add_indexed(idx, vals)
    for i, x in zip(idx, vals):
        ref[i] += x
    assert np.array_equal(v.get(), ref)
    with pytest.raises(fa.FeelHipError):
        v.add_indexed([n], [1.0])
    # y += A x
    x = rng.standard_normal(n)
    y = ctx.vector(n, a)
    y.mult_add(p2["A"], ctx.vector(n, x))
    assert rel_fro(y.get(), a + S @ x) < 1e-14
