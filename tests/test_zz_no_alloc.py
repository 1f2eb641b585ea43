"""feel_hip_mat_add_matrix_noalloc (MatrixSparse::addMatrix staged in a static
device array) against feel_hip_mat_add_matrix, bitwise, through the C-ABI:
two calls in a row (the second smaller), the capacity limit and the pattern
check."""
import numpy as np
import pytest

import feelpp_amd as fa

pytestmark = pytest.mark.gpu

CAPACITY = 65536


def noalloc():
    return fa.lib().feel_hip_mat_add_matrix_noalloc


def add_both(A, B, rows, cols, data):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    data = np.ascontiguousarray(data, dtype=np.float64)
    rc = noalloc()(A.h, rows.shape[0], rows, cols.shape[0], cols, data)
    assert rc == 0, fa.lib().feel_hip_last_error()
    B.add_matrix(rows, cols, data.reshape(rows.shape[0], cols.shape[0]))


def same(A, B):
    for prec in (fa.FP64, fa.FP32):
        a, b = A.values(prec), B.values(prec)
        assert np.array_equal(a, b), prec


def test_add_matrix_noalloc_matches_add_matrix(orc):
    ctx = fa.Context(0)
    coords, elems = orc.mesh_kuhn(3, jitter=0.1, seed=3)
    e2d, ndof = orc.doftable(3, 2, 1, elems)
    ia, ja = orc.graph(e2d, ndof)
    space = fa.Space(ctx, coords, elems, 2, 1, e2d, ndof)
    g = fa.Graph(space, ia, ja)
    A, B = fa.Matrix(g, fa.FP64_FP32), fa.Matrix(g, fa.FP64_FP32)
    A.zero()
    B.zero()
    rng = np.random.default_rng(11)
    # 1: a long row repeated up to the capacity (ordered ADD of repeated entries)
    r = int(np.argmax(np.diff(ia)))
    cols = ja[ia[r]:ia[r + 1]]
    reps = CAPACITY // cols.shape[0]
    rows = np.full(reps, r, dtype=np.int32)
    add_both(A, B, rows, cols, rng.standard_normal(reps * cols.shape[0]))
    same(A, B)
    # 2: a smaller call, one element matrix
    add_both(A, B, e2d[5], e2d[5], rng.standard_normal(e2d.shape[1] ** 2))
    same(A, B)
    assert np.count_nonzero(A.values(fa.FP64)) > 0
    before = A.values(fa.FP64)
    f = noalloc()
    # beyond the capacity: rejected, nothing changed
    big = np.full(reps + 1, r, dtype=np.int32)
    data = np.ones((reps + 1) * cols.shape[0])
    assert f(A.h, big.shape[0], big, cols.shape[0], np.ascontiguousarray(cols), data) == -1
    assert b"capacity" in fa.lib().feel_hip_last_error()
    # an entry outside the graph: EPATTERN, nothing changed
    far = np.array([0], dtype=np.int32)
    outside = np.array([ndof - 1], dtype=np.int32)
    assert ndof - 1 not in set(ja[ia[0]:ia[1]].tolist())
    assert f(A.h, 1, far, 1, outside, np.ones(1)) == -5
    # a row out of range: EINVAL
    assert f(A.h, 1, np.array([ndof], dtype=np.int32), 1, far, np.ones(1)) == -1
    assert np.array_equal(A.values(fa.FP64), before)
    # after the rejected calls the staging array still serves a valid call
    add_both(A, B, e2d[7][:4], e2d[7], rng.standard_normal(4 * e2d.shape[1]))
    same(A, B)
