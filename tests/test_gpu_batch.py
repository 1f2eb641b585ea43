"""Batched and continuously batched solves of affine families (GPU, C-ABI).

Each instance's solution must match the plain FP64 device solve of the same
A(mu) (itself pinned to the oracle by tests/test_gpu_solve.py) within 1e-8,
and be bitwise identical whether solved alone or in a mix of instances."""
import numpy as np
import pytest

import feelpp_amd as fa
from feelpp_amd.families import Family
from helpers import rel_fro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return fa.Context(0)


def family(ctx, n, order, seed=3):
    F = Family(ctx, order, n, seed=seed)
    return dict(obj=F, sp=F.space, g=F.graph, terms=F.terms, b=F.rhs, rows=F.rows, sub=F.sub, ndof=F.ndof,
                order=order, ne=F.ne, qorder=F.qorder)


def reference(ctx, F, mu):
    A = fa.Matrix(F["g"], fa.FP64)
    order = F["order"]
    A.assemble(fa.DIFFUSION, F["qorder"], kdiff=ctx.vector(F["ne"], np.asarray(mu)[F["sub"]]))
    b = ctx.vector(F["ndof"])
    fa.assemble_linear(F["sp"], b, fa.SRC_CONST, order, (1.0,))
    A.zero_rows(F["rows"], None, b, mode=fa.ELIMINATION_SYMMETRIC)
    x = ctx.vector(F["ndof"])
    st = A.solve(b, x, rtol=1e-13, maxit=100000)
    assert st["converged"]
    return x.get()


def run_all(B, ndofs, chunk=16, limit=20000):
    """Step until every request finished; fetch each solution when it completes
    (which releases its buffers for waiting requests).  ndofs: request -> size."""
    out = {}
    for _ in range(limit):
        act, _ = B.step(chunk)
        while True:
            got = B.poll()
            if not got:
                break
            rid, sd = got
            out[rid] = (sd, B.solution(rid, ndofs[rid]))
        if act == 0 and len(out) == len(ndofs):
            break
    return out


@pytest.mark.parametrize("precision", [fa.FP32, fa.FP64])
def test_batch_matches_single_solves(ctx, precision):
    fams = [family(ctx, 6, 1), family(ctx, 4, 2), family(ctx, 3, 3)]
    B = fa.Batch(ctx, rtol=1e-10, precision=precision)
    ids = [B.add_family(F["terms"], F["b"], F["rows"]) for F in fams]
    rng = np.random.default_rng(0)
    reqs = {}
    for k in range(9):
        fi = k % 3
        mu = np.exp(rng.uniform(np.log(0.1), np.log(10), 4))
        reqs[B.submit(ids[fi], mu)] = (fi, mu)
    done = run_all(B, {r: fams[fi]["ndof"] for r, (fi, _) in reqs.items()})
    assert set(done) == set(reqs)
    for rid, (fi, mu) in reqs.items():
        sd, x = done[rid]
        assert sd["converged"], sd
        assert rel_fro(x, reference(ctx, fams[fi], mu)) < 1e-8


def test_alone_equals_in_the_mix(ctx):
    F1, F2 = family(ctx, 5, 2), family(ctx, 6, 1)
    mu = np.array([0.3, 2.0, 7.5, 0.9])
    B1 = fa.Batch(ctx, rtol=1e-10)
    f1 = B1.add_family(F1["terms"], F1["b"], F1["rows"])
    r = B1.submit(f1, mu)
    alone = run_all(B1, {r: F1["ndof"]})[r][1]
    B2 = fa.Batch(ctx, rtol=1e-10)
    g1 = B2.add_family(F1["terms"], F1["b"], F1["rows"])
    g2 = B2.add_family(F2["terms"], F2["b"], F2["rows"])
    rng = np.random.default_rng(1)
    reqs = {B2.submit(g2, np.exp(rng.uniform(-2, 2, 4))): F2["ndof"] for _ in range(3)}
    target = B2.submit(g1, mu)
    reqs[target] = F1["ndof"]
    reqs.update({B2.submit(g1, np.exp(rng.uniform(-2, 2, 4))): F1["ndof"] for _ in range(2)})
    mixed = run_all(B2, reqs, chunk=7)[target][1]
    assert np.array_equal(alone, mixed)


def test_admission_by_memory_budget(ctx):
    F = family(ctx, 5, 2)
    probe = fa.Batch(ctx, rtol=1e-9)
    fp = probe.add_family(F["terms"], F["b"], F["rows"])
    probe.submit(fp, np.ones(4))
    probe.step(0)
    one = probe.device_bytes()
    B = fa.Batch(ctx, rtol=1e-9, hbm_budget=int(2.5 * one))
    f = B.add_family(F["terms"], F["b"], F["rows"])
    reqs = {B.submit(f, np.full(4, 1.0 + k)): F["ndof"] for k in range(6)}
    B.step(0)
    assert B.device_bytes() <= 2.5 * one          # only two admitted at first
    done = run_all(B, reqs)
    assert set(done) == set(reqs) and all(d[0]["converged"] for d in done.values())
    assert B.device_bytes() <= 2.5 * one


def two_term(ctx, F):
    """A(mu) = mu_0 A_{sub<2} + mu_1 A_{sub>=2} on the family's graph."""
    terms = []
    for m in (F["sub"] < 2, F["sub"] >= 2):
        A = fa.Matrix(F["g"], fa.FP64)
        A.assemble(fa.DIFFUSION, F["qorder"], kdiff=ctx.vector(F["ne"], m * 1.0))
        terms.append(A)
    return terms


@pytest.mark.parametrize("group", [True, False])
def test_groups_of_mixed_term_counts(ctx, group):
    # eleven instances of a 4-term family (two groups of the grouped SpMV) and
    # three of a 2-term family on the same graph, whose terms are padded to
    # four in the grouped pass; every solution against the FP64 solve, and a
    # 2-term instance bitwise equal to the same request solved alone
    F = family(ctx, 5, 2)
    t2 = two_term(ctx, F)
    B = fa.Batch(ctx, rtol=1e-10, group_spmv=group)
    f4 = B.add_family(F["terms"], F["b"], F["rows"])
    f2 = B.add_family(t2, F["b"], F["rows"])
    rng = np.random.default_rng(7)
    reqs = {}
    for k in range(14):
        if k % 5 == 2:
            mu = np.exp(rng.uniform(-2, 2, 2))
            reqs[B.submit(f2, mu)] = np.repeat(mu, 2)
        else:
            mu = np.exp(rng.uniform(-2, 2, 4))
            reqs[B.submit(f4, mu)] = mu
    done = run_all(B, {r: F["ndof"] for r in reqs}, chunk=9)
    for rid, mu4 in reqs.items():
        sd, x = done[rid]
        assert sd["converged"], sd
        assert rel_fro(x, reference(ctx, F, mu4)) < 1e-8
    rid2 = next(r for r, mu4 in reqs.items() if mu4[0] == mu4[1] and mu4[2] == mu4[3])
    A = fa.Batch(ctx, rtol=1e-10, group_spmv=group)
    a2 = A.add_family(t2, F["b"], F["rows"])
    r = A.submit(a2, reqs[rid2][::2])
    alone = run_all(A, {r: F["ndof"]}, chunk=9)[r][1]
    assert np.array_equal(alone, done[rid2][1])
