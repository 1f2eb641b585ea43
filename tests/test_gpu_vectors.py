"""Vector interface on the device (feelalg/vector.hpp: add(a, v) :474,
scale :576, operator= :271) and on()'s rhs rows (feelvf/on.hpp:457-701)."""
import numpy as np
import pytest

import feelpp_amd as fa

pytestmark = pytest.mark.gpu


def test_vector_operations():
    ctx = fa.Context(0)
    n = 1000
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    x, y = ctx.vector(n, a), ctx.vector(n, b)
    y.axpy(2.5, x)
    assert np.array_equal(y.get(), 2.5 * a + b)
    y.scale(-0.5)
    assert np.array_equal(y.get(), -0.5 * (2.5 * a + b))
    z = ctx.vector(n)
    z.copy_from(x)
    assert np.array_equal(z.get(), a)
    rows = np.array([0, 7, 999, 500], dtype=np.int32)
    R = fa.Rows(ctx, n, rows)
    z.set_rows(R)
    want = a.copy()
    want[rows] = 0.0
    assert np.array_equal(z.get(), want)
    z.set_rows(R, y)
    want[rows] = y.get()[rows]
    assert np.array_equal(z.get(), want)
    with pytest.raises(RuntimeError):
        x.axpy(1.0, ctx.vector(n + 1))
