"""Batched and continuously batched solves of affine families (GPU, C-ABI).

Each instance's solution must match the FP64 oracle's tightly converged solve
of the same A(mu) (oracle assembly with k = mu[subdomain], MatZeroRowsColumns
elimination, Jacobi-CG to 1e-13) within 1e-8, and be bitwise identical whether
solved alone or in a mix of instances, including instances admitted while
others are on the GPU."""
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


def oracle_problem(orc, F, mu):
    """The oracle's own numbering, pattern, A(mu), rhs and Dirichlet elimination of
    a family member; the device DofTable of the family must equal the oracle's."""
    sp, order = F["sp"], F["order"]
    coords, elems = sp.coords, sp.elems
    e2d, ndof = orc.doftable(3, order, 1, elems)
    assert ndof == F["ndof"] and np.array_equal(e2d, sp.elem2dof)
    ia, ja = orc.graph(e2d, ndof)
    val = orc.assemble_scalar(3, order, F["qorder"], coords, elems, e2d, ia, ja, diff=1.0,
                              kdiff=np.asarray(mu, dtype=np.float64)[F["sub"]])
    b = orc.assemble_rhs(3, order, 1, order, coords, elems, e2d, ndof, orc.FK_CONST, [1.0])
    isb = orc.boundary_dofs(3, order, 1, coords, elems, e2d, ndof)
    orc.dirichlet(1, ia, ja, val, b, isb)
    return ia, ja, val, b


def oracle_solution(orc, F, mu):
    ia, ja, val, b = oracle_problem(orc, F, mu)
    x, its, _, reason = orc.pcg(ia, ja, val, b, rtol=1e-13, maxit=100000)
    assert reason > 0
    return x


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
def test_batch_vs_oracle(ctx, orc, precision):
    # C4's family shapes at oracle-sized n: orders P1-P4 in one batch, every
    # solution against the oracle's FP64 solve of the same A(mu)
    fams = [family(ctx, 6, 1), family(ctx, 4, 2), family(ctx, 3, 3), family(ctx, 2, 4)]
    B = fa.Batch(ctx, rtol=1e-10, precision=precision)
    ids = [B.add_family(F["terms"], F["b"], F["rows"]) for F in fams]
    rng = np.random.default_rng(0)
    reqs = {}
    for k in range(12):
        fi = k % len(fams)
        mu = np.exp(rng.uniform(np.log(0.1), np.log(10), 4))
        reqs[B.submit(ids[fi], mu)] = (fi, mu)
    done = run_all(B, {r: fams[fi]["ndof"] for r, (fi, _) in reqs.items()})
    assert set(done) == set(reqs)
    for rid, (fi, mu) in reqs.items():
        sd, x = done[rid]
        assert sd["converged"], sd
        assert sd["fp64_its"] == (sd["its"] if precision == fa.FP64 else 0)
        err = rel_fro(x, oracle_solution(orc, fams[fi], mu))
        print(f"P{fams[fi]['order']} n={fams[fi]['obj'].n} mu={np.round(mu, 3)} its={sd['its']} err={err:.2e}")
        assert err < 1e-8


def test_batch_matches_single_solves(ctx):
    # the batched engine and the single-solve path agree on the same A(mu)
    fams = [family(ctx, 6, 1), family(ctx, 3, 3)]
    B = fa.Batch(ctx, rtol=1e-10)
    ids = [B.add_family(F["terms"], F["b"], F["rows"]) for F in fams]
    rng = np.random.default_rng(5)
    reqs = {}
    for k in range(4):
        mu = np.exp(rng.uniform(np.log(0.1), np.log(10), 4))
        reqs[B.submit(ids[k % 2], mu)] = (k % 2, mu)
    done = run_all(B, {r: fams[fi]["ndof"] for r, (fi, _) in reqs.items()})
    for rid, (fi, mu) in reqs.items():
        assert rel_fro(done[rid][1], reference(ctx, fams[fi], mu)) < 1e-8


def test_requests_admitted_mid_flight(ctx, orc):
    # continuous batching: requests of mixed order and size submitted between
    # step() calls while earlier ones are on the GPU, with an HBM budget that
    # keeps some of them queued; every result matches the oracle and is bitwise
    # the same as the request solved alone in a fresh batch
    fams = [family(ctx, 5, 2), family(ctx, 6, 1), family(ctx, 3, 3), family(ctx, 2, 4)]
    probe = fa.Batch(ctx, rtol=1e-10)
    pids = [probe.add_family(F["terms"], F["b"], F["rows"]) for F in fams]
    for i in pids:
        probe.submit(i, np.ones(4))
    probe.step(0)
    budget = int(0.7 * probe.device_bytes())          # not all four fit at once
    del probe
    B = fa.Batch(ctx, rtol=1e-10, hbm_budget=budget)
    ids = [B.add_family(F["terms"], F["b"], F["rows"]) for F in fams]
    rng = np.random.default_rng(11)
    schedule = {0: [0, 2], 2: [1, 3], 3: [0], 5: [2, 1]}   # step index -> families submitted before it
    reqs, done, inflight_at_submit = {}, {}, []
    for k in range(20000):
        for fi in schedule.get(k, []):
            if reqs:
                inflight_at_submit.append(len(reqs) - len(done))
            mu = np.exp(rng.uniform(np.log(0.1), np.log(10), 4))
            reqs[B.submit(ids[fi], mu)] = (fi, mu)
        act, _ = B.step(4)
        assert B.device_bytes() <= budget
        while True:
            got = B.poll()
            if not got:
                break
            rid, sd = got
            done[rid] = (sd, B.solution(rid, fams[reqs[rid][0]]["ndof"]))
        if act == 0 and k > max(schedule):
            break
    assert set(done) == set(reqs) and len(reqs) == 7
    assert min(inflight_at_submit) >= 1                # every later submission met requests in flight
    for rid, (fi, mu) in reqs.items():
        sd, x = done[rid]
        assert sd["converged"], sd
        assert rel_fro(x, oracle_solution(orc, fams[fi], mu)) < 1e-8
        A = fa.Batch(ctx, rtol=1e-10)
        a = A.add_family(fams[fi]["terms"], fams[fi]["b"], fams[fi]["rows"])
        r = A.submit(a, mu)
        alone = run_all(A, {r: fams[fi]["ndof"]}, chunk=4)[r][1]
        assert np.array_equal(alone, x), f"request {rid} differs alone and in the mix"


@pytest.mark.parametrize("precision", [fa.FP32, fa.FP64])
def test_breakdown_returns_its_iterate(ctx, orc, precision):
    # an indefinite A(mu) (negative parameters) stops CG with a negative reason;
    # the returned vector is that request's last iterate, with its true FP64
    # residual ||D^-1 (b - A x)||, even when its buffers held an earlier solution
    F = family(ctx, 4, 2)
    good, bad = np.array([1.0, 2.0, 0.5, 1.5]), np.array([1.0, -3.0, 1.0, -3.0])
    B = fa.Batch(ctx, rtol=1e-10, precision=precision)
    f = B.add_family(F["terms"], F["b"], F["rows"])
    r0 = B.submit(f, good)
    x0 = run_all(B, {r0: F["ndof"]})[r0][1]
    r1 = B.submit(f, bad)                               # reuses r0's instance buffers
    sd, x1 = run_all(B, {r1: F["ndof"]})[r1]
    ia, ja, val, b = oracle_problem(orc, F, bad)
    _, _, _, reason_o = orc.pcg(ia, ja, val, b, rtol=1e-10, maxit=100000)
    print(sd, "oracle reason", reason_o)
    assert reason_o < 0 and sd["reason"] < 0 and not sd["converged"]
    assert np.all(np.isfinite(x1)) and not np.array_equal(x1, x0)
    d = np.array([val[ia[i]:ia[i + 1]][ja[ia[i]:ia[i + 1]] == i][0] for i in range(len(b))])
    res = np.linalg.norm((b - orc.spmv(ia, ja, val, x1)) / d)
    assert abs(res - sd["residual"]) <= 1e-9 * max(res, np.linalg.norm(b / d))


def test_poll_after_solution_taken(ctx):
    # a request whose solution is fetched before it was polled is not reported again
    F = family(ctx, 4, 1)
    B = fa.Batch(ctx, rtol=1e-10)
    f = B.add_family(F["terms"], F["b"], F["rows"])
    r = B.submit(f, np.ones(4))
    for _ in range(1000):
        act, nd = B.step(16)
        if nd:
            break
    B.solution(r, F["ndof"])
    assert B.poll() is None


def test_full_size_p4_family_member(ctx):
    # the largest C4 problem at full size (P4, Kuhn 32^3: 2 146 689 dofs): the
    # batched solution's true FP64 residual, evaluated through the single-solve
    # assembly and the CSR product, meets the tolerance; the reported residual
    # is that residual
    F = family(ctx, 32, 4, seed=432)
    mu = np.array([0.15, 7.0, 1.3, 0.4])
    B = fa.Batch(ctx, rtol=1e-8)
    f = B.add_family(F["terms"], F["b"], F["rows"])
    r = B.submit(f, mu)
    sd, x = run_all(B, {r: F["ndof"]}, chunk=32)[r]
    assert sd["converged"] and F["ndof"] == 2146689
    A = fa.Matrix(F["g"], fa.FP64)
    A.assemble(fa.DIFFUSION, F["qorder"], kdiff=ctx.vector(F["ne"], mu[F["sub"]]))
    b = ctx.vector(F["ndof"])
    fa.assemble_linear(F["sp"], b, fa.SRC_CONST, 4, (1.0,))
    A.zero_rows(F["rows"], None, b, mode=fa.ELIMINATION_SYMMETRIC)
    y = ctx.vector(F["ndof"])
    A.mult(ctx.vector(F["ndof"], x), y)
    dinv = 1.0 / A.diagonal()
    res = np.linalg.norm(dinv * (b.get() - y.get()))
    res0 = np.linalg.norm(dinv * b.get())
    print(sd, res / res0)
    assert res <= 1e-8 * res0 * (1 + 1e-6)
    assert abs(res - sd["residual"]) <= 1e-6 * res


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


@pytest.mark.parametrize("precision", [fa.FP32, fa.FP64])
@pytest.mark.parametrize("group", [True, False])
def test_groups_of_mixed_term_counts(ctx, group, precision):
    # eleven instances of a 4-term family (two groups of the grouped SpMV) and
    # three of a 2-term family on the same graph, whose terms are padded to
    # four in the grouped pass; every solution against the FP64 solve, and a
    # 2-term instance bitwise equal to the same request solved alone
    F = family(ctx, 5, 2)
    t2 = two_term(ctx, F)
    B = fa.Batch(ctx, rtol=1e-10, group_spmv=group, precision=precision)
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
    A = fa.Batch(ctx, rtol=1e-10, group_spmv=group, precision=precision)
    a2 = A.add_family(t2, F["b"], F["rows"])
    r = A.submit(a2, reqs[rid2][::2])
    alone = run_all(A, {r: F["ndof"]}, chunk=9)[r][1]
    assert np.array_equal(alone, done[rid2][1])


def test_backfill_admission(ctx):
    # A budget for one large instance and a few small ones: in strict order the
    # small requests wait behind the second large one; with backfill they pass
    # it, until it has waited `backfill` steps.  Results do not depend on the
    # admission order.
    big, small = family(ctx, 6, 2), family(ctx, 2, 1)

    def one_instance(F):
        P = fa.Batch(ctx, rtol=1e-9, group_spmv=False)
        f = P.add_family(F["terms"], F["b"], F["rows"])
        base = P.device_bytes()
        P.submit(f, np.ones(4))
        P.step(0)
        return P.device_bytes() - base

    Bb, Bs = one_instance(big), one_instance(small)
    assert Bb > 4 * Bs
    budget = Bb + 2 * Bs + Bs // 2
    mus = [np.full(4, 1.0 + k) for k in range(4)]

    def batch(backfill):
        B = fa.Batch(ctx, rtol=1e-9, group_spmv=False, hbm_budget=budget, backfill=backfill)
        fb = B.add_family(big["terms"], big["b"], big["rows"])
        fs = B.add_family(small["terms"], small["b"], small["rows"])
        return B, fb, fs

    results = {}
    for backfill in (0, 2):
        B, fb, fs = batch(backfill)
        base = B.device_bytes()
        ids = [B.submit(fb, mus[0]), B.submit(fb, mus[1]), B.submit(fs, mus[2]), B.submit(fs, mus[3])]
        B.step(0)
        held = B.device_bytes() - base
        assert held == (Bb if backfill == 0 else Bb + 2 * Bs), (backfill, held)
        ndofs = dict(zip(ids, [big["ndof"]] * 2 + [small["ndof"]] * 2))
        done = run_all(B, ndofs)
        assert all(d[0]["converged"] for d in done.values())
        results[backfill] = [done[i][1] for i in ids]
    for a, b in zip(results[0], results[2]):
        assert np.array_equal(a, b)
    # no starvation: after `backfill` steps nothing passes the waiting request
    B, fb, fs = batch(1)
    base = B.device_bytes()
    ids = [B.submit(fb, mus[0]), B.submit(fb, mus[1]), B.submit(fs, mus[2])]
    B.step(0)
    assert B.device_bytes() - base == Bb + Bs
    ids.append(B.submit(fs, mus[3]))
    B.step(0)
    assert B.device_bytes() - base == Bb + Bs
    ndofs = dict(zip(ids, [big["ndof"]] * 2 + [small["ndof"]] * 2))
    done = run_all(B, ndofs)
    assert set(done) == set(ids) and all(d[0]["converged"] for d in done.values())
    for i, x in zip(ids, results[0]):
        assert np.array_equal(done[i][1], x)
