"""Affine-parameter Poisson families on unit-cube Kuhn meshes, resident on the
device, for the batched and serving drivers.

A family is (order k, n cubes per side): the four term matrices
A_t = integrate(elements, chi_t * inner(gradt(u), grad(v))) over the four
axis-aligned subdomains {x <> 1/2} x {y <> 1/2} (chi_t the indicator of
subdomain t, a P0 field), the load l = integrate(elements, id(v)) and the
boundary rows of on(boundaryfaces(mesh), g = 0).  An instance is a parameter
mu in R^4: A(mu) = sum_t mu_t A_t.
"""
import time

import numpy as np

from . import DIFFUSION, FP64, SRC_CONST, Graph, Matrix, Rows, Space, assemble_linear, boundary_dofs, mesh_kuhn_box


class Family:
    def __init__(self, ctx, order, n, seed=0, jitter=0.15):
        t0 = time.perf_counter()
        self.order, self.n = order, n
        coords, elems = mesh_kuhn_box(n, jitter=jitter, seed=seed)
        self.space = Space.from_mesh(ctx, coords, elems, order)
        self.graph = Graph.from_space(self.space)
        e2d, ndof, ja = self.space.elem2dof, self.space.ndof, self.graph.ja
        cen = coords[elems].mean(axis=1)
        self.sub = (cen[:, 0] > 0.5).astype(np.int64) + 2 * (cen[:, 1] > 0.5).astype(np.int64)
        self.qorder = max(2 * order - 2, 1)   # exact for inner(gradt, grad) on affine P_k
        self.terms = []
        for t in range(4):
            A = Matrix(self.graph, FP64)
            A.assemble(DIFFUSION, self.qorder, kdiff=ctx.vector(len(elems), (self.sub == t) * 1.0))
            self.terms.append(A)
        self.rhs = ctx.vector(ndof)
        assemble_linear(self.space, self.rhs, SRC_CONST, order, (1.0,))
        isb = boundary_dofs(3, order, 1, coords, elems, e2d, ndof)
        self.rows = Rows(ctx, ndof, np.nonzero(isb)[0])
        self.ndof, self.ne, self.nnz = ndof, len(elems), len(ja)
        ctx.synchronize()
        self.setup_s = time.perf_counter() - t0
