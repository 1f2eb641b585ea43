#!/usr/bin/env python3
"""Benchmark: P2-tet Poisson assemble+solve/s on one MI355X.

Workload (one step): on a 64^3 Kuhn tetrahedral mesh of the unit cube
(1 572 864 tets, 2 146 689 P2 dofs, 60 859 905 nonzeros, vertex jitter 0.15h),
assemble inner(k grad u, grad v) with a lognormal(0,1) per-element k at
quadrature order 4, assemble the source f = 1, eliminate homogeneous Dirichlet
rows/columns on the whole boundary, and solve with Jacobi-CG to relative
residual 1e-10.  Mesh, dof table, graph and coefficient are resident in HBM
before the timed region; every step re-assembles and re-solves from x = 0.

Two solver forms are timed: plain FP64 Jacobi-CG, and FP32 Jacobi-CG inside
FP64 iterative refinement (FP64 residual and stopping test).  The reported
value is the faster form whose solution is within 1e-8 relative error of the
FP64 reference solution.

Prints one JSON line (rank 0).  --gpus N>1 is not supported: this build runs
on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "P2-tet Poisson assemble+solve/s at 1e-8 rel err vs FP64; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_problem(fa, ctx, n, order, jitter, seed):
    """Mesh on the host; DofTable, sparsity pattern and SELL image on the device.
    host: the arrays the CPU baseline works on (elem2dof and ia/ja copied back)."""
    t0 = time.perf_counter()
    coords, elems = fa.mesh_kuhn_box(n, jitter=jitter, seed=seed)
    t1 = time.perf_counter()
    space = fa.Space.from_mesh(ctx, coords, elems, order)
    graph = fa.Graph.from_space(space)
    ctx.synchronize()
    t2 = time.perf_counter()
    k = fa.lognormal(len(elems), 0.0, 1.0, seed + 1)
    isb = fa.boundary_dofs(3, order, 1, coords, elems, space.elem2dof, space.ndof)
    log(f"setup: mesh {t1 - t0:.2f}s, device DofTable + pattern + SELL {t2 - t1:.2f}s: {len(elems)} tets, "
        f"{space.ndof} dofs, {graph.nnz} nnz")
    host = (coords, elems, space.elem2dof, space.ndof, graph.ia, graph.ja, k, isb)
    return space, graph, host


class Problem:
    def __init__(self, fa, ctx, built, order, qorder):
        self.space, self.graph, self.host = built
        coords, elems, e2d, ndof, ia, ja, k, isb = self.host
        t0 = time.perf_counter()
        self.fa, self.ctx = fa, ctx
        self.A = fa.Matrix(self.graph, fa.FP64_FP32)
        self.k = ctx.vector(len(elems), k)
        self.b = ctx.vector(ndof)
        self.x = ctx.vector(ndof)
        self.rows = fa.Rows(ctx, ndof, np.nonzero(isb)[0])
        self.order, self.qorder, self.ndof, self.ne, self.nnz = order, qorder, ndof, len(elems), len(ja)
        ctx.synchronize()
        log(f"device buffers {time.perf_counter() - t0:.1f}s, resident {ctx.device_bytes() / 2**30:.2f} GiB")

    def step(self, precision, rtol, inner_rtol=1e-2, inner_operator=0):
        fa = self.fa
        self.A.assemble(fa.DIFFUSION, self.qorder, kdiff=self.k, precision=fa.FP64_FP32)
        fa.assemble_linear(self.space, self.b, fa.SRC_CONST, self.order, (1.0,))
        self.A.zero_rows(self.rows, None, self.b, mode=fa.ELIMINATION_SYMMETRIC)
        self.x.zero()
        return self.A.solve(self.b, self.x, rtol=rtol, maxit=100000, precision=precision, inner_rtol=inner_rtol,
                            inner_operator=inner_operator)


# solver forms: (precision, operator of the FP32 inner iterations)
FORMS = {"fp64": ("FP64", "INNER_AUTO"), "mixed": ("FP32", "INNER_ASSEMBLED"),
         "mixed_element": ("FP32", "INNER_ELEMENT")}


def time_form(prob, form, rtol, steps, warmup, inner_rtol):
    ctx, fa = prob.ctx, prob.fa
    precision, op = (getattr(fa, name) for name in FORMS[form])
    for _ in range(warmup):
        prob.step(precision, rtol, inner_rtol, op)
    ctx.synchronize()
    ctx.reset_timing()
    # events around every 64th launch of each timed kernel: the kernel averages
    # come from the timed region (about 16 samples per solve) while the events
    # cost little of it (each recorded pair leaves the queue idle ~8 us on
    # either side; every 16th launch cost 1.2 % of a C2 step)
    ctx.set_timing(64)
    lat = []
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        st = prob.step(precision, rtol, inner_rtol, op)
        ctx.synchronize()
        lat.append(time.perf_counter() - s)
        its.append(st)
    total = time.perf_counter() - t0
    ctx.set_timing(False)
    timers = {name: ctx.timing(name)
              for name in ("spmv_fp64", "spmv_fp32", "elemop_fp32", "assemble_gather", "assemble_geo")}
    return total, lat, its, timers


def read_traffic(path):
    """'traffic <kernel> <bytes>' lines of a committed counter summary."""
    out = {}
    if os.path.exists(path):
        for ln in open(path):
            f = ln.split()
            if len(f) == 3 and f[0] == "traffic":
                out[f[1]] = int(f[2])
    return out


def cpu_threads():
    """Host threads for the all-cores baseline: OMP_NUM_THREADS when set (the
    GPU box sets it to the job's 16-core share), else the visible CPUs."""
    n = int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(orc, host, order, qorder, rtol, xref):
    """The FP64 CPU oracle on this host's cores, standing in for Feel++'s MPI
    CPU path: (ii) all cores, a full measured assemble (threads owning row
    blocks, ghost elements recomputed) + Dirichlet + threaded Jacobi-CG to rtol;
    (i) one core, assembly of a slice of the elements and 20 CG iterations on
    the same matrix, extrapolated to a full assemble+solve."""
    coords, elems, e2d, ndof, ia, ja, k, isb = host
    ne = len(elems)
    T = cpu_threads()
    t0 = time.perf_counter()
    val = orc.assemble_scalar_mt(3, order, qorder, coords, elems, e2d, ia, ja, diff=1.0, kdiff=k, threads=T)
    b = orc.assemble_rhs(3, order, 1, qorder, coords, elems, e2d, ndof, orc.FK_CONST, [1.0])
    orc.dirichlet(1, ia, ja, val, b, isb)
    t_asm = time.perf_counter() - t0
    t1 = time.perf_counter()
    x, its, _, reason = orc.pcg_mt(ia, ja, val, b, rtol=rtol, maxit=100000, threads=T)
    t_solve = time.perf_counter() - t1
    err = float(np.linalg.norm(x - xref) / np.linalg.norm(xref))
    log(f"cpu all cores ({T}): assemble {t_asm:.2f}s, solve {t_solve:.2f}s ({its} its, reason {reason}), "
        f"rel err vs GPU FP64 {err:.2e}")
    # one core
    frac_e = max(1, ne // 16)
    t2 = time.perf_counter()
    orc.assemble_scalar(3, order, qorder, coords, elems[:frac_e], e2d[:frac_e], ia, ja, diff=1.0, kdiff=k[:frac_e])
    t_asm1 = (time.perf_counter() - t2) * ne / frac_e
    n_it = 20
    t3 = time.perf_counter()
    orc.pcg(ia, ja, val, b, rtol=1e-300, maxit=n_it)
    t_it1 = (time.perf_counter() - t3) / n_it
    one = 1.0 / (t_asm1 + t_it1 * its)
    return {
        "value": 1.0 / (t_asm + t_solve),
        "unit": "solves/s",
        "cores": T,
        "kind": "port",
        "sample": (f"the full problem, measured: FP64 oracle assembly on {T} threads ({t_asm:.2f}s incl. rhs and "
                   f"Dirichlet) + threaded Jacobi-CG to rtol {rtol:g} ({its} iterations, {t_solve:.2f}s); "
                   f"solution vs the GPU FP64 solve {err:.1e}"),
        "one_core": {"value": one, "unit": "solves/s", "cores": 1, "extrapolated": True,
                     "sample": (f"assembly of {frac_e} of {ne} elements ({t_asm1 * frac_e / ne:.2f}s, scaled "
                                f"x{ne / frac_e:.0f}) + {n_it} CG iterations ({t_it1 * 1e3:.1f} ms each) "
                                f"extrapolated to {its}")},
    }


def affine_families(fa, ctx, orders, sizes):
    from feelpp_amd.families import Family
    fams = {}
    for k in orders:
        for n in sizes:
            fams[(k, n)] = Family(ctx, k, n, seed=100 * k + n)
            F = fams[(k, n)]
            log(f"family P{k} n={n}: {F.ndof} dofs, {F.nnz} nnz, setup {F.setup_s:.1f}s")
    return fams


def run_batch(fa, B, fams, ids, draws, chunk, keep=0):
    """Submit all draws at once, step until done; returns per-request latency (s), solve data,
    wall time and the solutions of the first `keep` draws (by draw index)."""
    t0 = time.perf_counter()
    sub = {}
    for i, (key, mu) in enumerate(draws):
        sub[B.submit(ids[key], mu)] = (key, t0, i)
    lat, stats, sols = {}, {}, {}
    last_log = t0
    while len(lat) < len(sub):
        if time.perf_counter() - last_log > 10:
            last_log = time.perf_counter()
            log(f"batch: {len(lat)}/{len(sub)} finished after {last_log - t0:.0f}s")
        B.step(chunk)
        while True:
            got = B.poll()
            if not got:
                break
            rid, sd = got
            xs = B.solution(rid, fams[sub[rid][0]].ndof)
            lat[rid] = time.perf_counter() - sub[rid][1]
            stats[rid] = sd
            if sub[rid][2] < keep:
                sols[sub[rid][2]] = xs
    return lat, stats, time.perf_counter() - t0, sols


def bench_batched(fa, args):
    """512 affine Poisson problems (orders P1-P4, Kuhn n in {8,12,16,24,32}) in one batch."""
    ctx = fa.Context(0)
    orders = [int(o) for o in args.orders.split(",")]
    sizes = [int(n) for n in args.sizes.split(",")]
    fams = affine_families(fa, ctx, orders, sizes)
    rng = np.random.default_rng(args.seed)
    keys = list(fams)
    draws = [(keys[rng.integers(len(keys))], np.exp(rng.uniform(np.log(0.1), np.log(10.0), 4)))
             for _ in range(args.problems)]
    B = fa.Batch(ctx, rtol=args.rtol, inner_rtol=args.inner_rtol or 1e-2, precision=fa.FP32, use_graph=True,
                 group_spmv=not args.no_group)
    ids = {k: B.add_family(F.terms, F.rhs, F.rows) for k, F in fams.items()}
    run_batch(fa, B, fams, ids, draws[:min(8, len(draws))], args.chunk)      # warm-up
    keep = min(len(draws), 64)
    lat, stats, wall, sols = run_batch(fa, B, fams, ids, draws, args.chunk, keep=keep)
    its = [s["its"] for s in stats.values()]
    ok = all(s["converged"] for s in stats.values())
    log(f"batched mixed: {len(draws) / wall:.2f} solves/s, mean {np.mean(its):.0f} iterations")
    del B
    # the same engine and draws, plain FP64 Jacobi-CG (like for like: the same batching and grouped term
    # stream, without the precision change)
    B64 = fa.Batch(ctx, rtol=args.rtol, precision=fa.FP64, use_graph=True, group_spmv=not args.no_group)
    ids64 = {k: B64.add_family(F.terms, F.rhs, F.rows) for k, F in fams.items()}
    run_batch(fa, B64, fams, ids64, draws[:min(8, len(draws))], args.chunk)
    lat64, stats64, wall64, sols64 = run_batch(fa, B64, fams, ids64, draws, args.chunk, keep=keep)
    ok64 = all(s["converged"] for s in stats64.values())
    err_mixed = max(float(np.linalg.norm(sols[i] - sols64[i]) / np.linalg.norm(sols64[i])) for i in range(keep))
    log(f"batched fp64: {len(draws) / wall64:.2f} solves/s; mixed vs fp64 max rel err {err_mixed:.2e} "
        f"over the first {keep} requests")
    del B64
    # plain FP64 GPU form on a sample: assemble A(mu) + Dirichlet + FP64 Jacobi-CG, one at a time
    sample = draws[:args.plain_sample]
    t1 = time.perf_counter()
    for key, mu in sample:
        F = fams[key]
        A = fa.Matrix(F.graph, fa.FP64)
        A.assemble(fa.DIFFUSION, F.qorder, kdiff=ctx.vector(F.ne, mu[F.sub]))
        b = ctx.vector(F.ndof)
        fa.assemble_linear(F.space, b, fa.SRC_CONST, F.order, (1.0,))
        A.zero_rows(F.rows, None, b, mode=fa.ELIMINATION_SYMMETRIC)
        x = ctx.vector(F.ndof)
        A.solve(b, x, rtol=args.rtol, maxit=100000)
    plain_s = (time.perf_counter() - t1) / max(len(sample), 1) * len(draws)
    L = np.array(list(lat.values()))
    line = {
        "metric": "batched affine Poisson assemble+solve/s (FP32 CG in FP64 refinement)",
        "value": round(len(draws) / wall, 3), "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "config": {"workload": f"{len(draws)} affine Poisson problems, P{orders} on Kuhn n in {sizes}, "
                               f"4-term subdomain coefficient mu log-uniform [0.1,10], rtol {args.rtol:g}"},
        "wall_s": round(wall, 3), "converged": ok, "iterations_mean": float(np.mean(its)),
        "latency_ms": {"p50": round(float(np.percentile(L, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(L, 99)) * 1e3, 2)},
        "batched_fp64_solves_per_s": round(len(draws) / wall64, 3),
        "batched_fp64_converged": ok64,
        "batched_fp64_grouped": not args.no_group,
        "speedup_vs_batched_fp64": round(wall64 / wall, 3),
        "rel_err_mixed_vs_batched_fp64": err_mixed,
        "rel_err_sample": keep,
        "plain_fp64_sequential_solves_per_s": round(len(draws) / plain_s, 3),
        "plain_fp64_sample": len(sample),
        "speedup_vs_plain_fp64": round(plain_s / wall, 3),
    }
    print(json.dumps(line), flush=True)


def bench_beam(fa, args):
    """Isotropic linear elasticity on the 1 x 0.25 x 0.25 Kuhn beam (40x10x10
    cells, 24 000 tets), vector Lagrange P3 and P4: Young's modulus log-uniform
    in [1, 100] per element, Poisson ratio uniform in [0.2, 0.35], left face
    clamped, body force (0, 0, -1), quadrature order 2k; FP64 assembly, FP64
    and mixed-precision Jacobi-CG."""
    ctx = fa.Context(0)
    res = {}
    for k in [int(o) for o in args.orders.split(",")]:
        t0 = time.perf_counter()
        coords, elems = fa.mesh_kuhn_box(40, 10, 10, lx=1.0, ly=0.25, lz=0.25, jitter=0.0, seed=args.seed)
        ne = len(elems)
        E = fa.uniform(ne, 1.0, 100.0, args.seed + 1, log=True)
        nu = fa.uniform(ne, 0.2, 0.35, args.seed + 2)
        space = fa.Space.from_mesh(ctx, coords, elems, k, 3)
        g = fa.Graph.from_space(space)
        e2d, ndof, ja = space.elem2dof, space.ndof, g.ja
        isb = fa.boundary_dofs(3, k, 3, coords, elems, e2d, ndof, sel=1)
        log(f"beam P{k}: setup {time.perf_counter() - t0:.1f}s, {ne} tets, {ndof} dofs, {len(ja)} nnz")
        A = fa.Matrix(g, fa.FP64_FP32)
        young, poisson = ctx.vector(ne, E), ctx.vector(ne, nu)
        b, x = ctx.vector(ndof), ctx.vector(ndof)
        rows = fa.Rows(ctx, ndof, np.nonzero(isb)[0])
        qo = 2 * k

        def assemble():
            A.assemble(fa.ELASTICITY, qo, young=young, poisson=poisson, precision=fa.FP64_FP32)
            fa.assemble_linear(space, b, fa.SRC_VECTOR, qo, (0.0, 0.0, -1.0))
            A.zero_rows(rows, None, b, mode=fa.ELIMINATION_SYMMETRIC)

        assemble()
        ctx.synchronize()
        reps = max(1, args.steps)
        t1 = time.perf_counter()
        for _ in range(reps):
            assemble()
        ctx.synchronize()
        t_asm = (time.perf_counter() - t1) / reps
        # the bilinear assembly alone on each element-matrix path (the default
        # above takes the faster of the two, timed on its first assembly)
        bilinear = {}
        for name, m in (("fused", fa.MATRICES_FUSED), ("mfma", fa.MATRICES_MFMA)):
            for rep in range(reps + 1):
                if rep == 1:
                    ctx.synchronize()
                    tb = time.perf_counter()
                A.assemble(fa.ELASTICITY, qo, young=young, poisson=poisson, precision=fa.FP64_FP32, matrices=m)
            ctx.synchronize()
            bilinear[name] = round((time.perf_counter() - tb) / reps * 1e3, 3)
        log(f"beam P{k}: bilinear assembly ms by element-matrix path {bilinear}")
        assemble()
        x.zero()
        st = A.solve(b, x, rtol=1e-13, maxit=200000)
        xref = x.get()
        forms = {}
        for form, prec in (("fp64", fa.FP64), ("mixed", fa.FP32)):
            x.zero()
            ctx.synchronize()
            t2 = time.perf_counter()
            st = A.solve(b, x, rtol=args.rtol, maxit=200000, precision=prec)
            ctx.synchronize()
            t_solve = time.perf_counter() - t2
            err = float(np.linalg.norm(x.get() - xref) / np.linalg.norm(xref))
            forms[form] = {"assemble_ms": round(t_asm * 1e3, 3), "solve_ms": round(t_solve * 1e3, 2),
                           "solves_per_s": round(1.0 / (t_asm + t_solve), 4), "its": st["its"],
                           "reliable_updates": st["outer_its"], "converged": st["converged"],
                           "rel_err_vs_fp64": err}
            log(f"beam P{k} {form}: {forms[form]}")
        res[f"P{k}"] = {"dofs": ndof, "nnz": len(ja), "elements": ne,
                        "elements_per_s_assembly": round(ne / t_asm, 1), "bilinear_ms": bilinear, "forms": forms}
    best = {kk: max(v["forms"].values(), key=lambda f: f["solves_per_s"] if f["rel_err_vs_fp64"] <= 1e-8 else 0)
            for kk, v in res.items()}
    line = {
        "metric": "P3/P4 elasticity beam assemble+solve/s at 1e-8 rel err vs FP64",
        "value": {kk: v["solves_per_s"] for kk, v in best.items()}, "unit": "solves/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f64 assembly, f32+f64 solve",
        "config": {"workload": "isotropic elasticity, 1x0.25x0.25 Kuhn beam 40x10x10, E log-uniform [1,100], "
                               f"nu uniform [0.2,0.35], clamped x=0, f=(0,0,-1), quadrature 2k, rtol {args.rtol:g}"},
        "results": res,
    }
    print(json.dumps(line), flush=True)


def memory_reference():
    """A plain 1 GiB device copy, 5 timed launches after one warm-up (1 GiB
    read + 1 GiB written each): it calibrates the HBM byte counters of the
    FETCH_SIZE / WRITE_SIZE passes and gives a copy rate beside the kernels."""
    import torch
    nbytes = 1 << 30
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(5):
        b.copy_(a)
    ev[1].record()
    torch.cuda.synchronize()
    out = {"copy_TBps": round(2 * nbytes / (ev[0].elapsed_time(ev[1]) / 5e3) / 1e12, 3)}
    log(f"copy kernel: {nbytes} bytes read + {nbytes} bytes written per launch, 6 launches; "
        f"{out['copy_TBps']} TB/s")
    del a, b
    torch.cuda.empty_cache()
    return out


def bench_kernels(fa, args):
    """Short fixed runs of the hot kernels for hardware-counter passes: a plain
    device copy of known size (calibrates the HBM byte counters), then on the
    C2 problem 3 assemblies (scatter kernel) and 20-iteration solves: FP64 (SELL
    SpMV), mixed with the assembled FP32 matrix (FP32 SELL SpMV) and mixed with
    the element-by-element FP32 operator."""
    stream = memory_reference()
    order, qorder = args.order, 2 * args.order
    ctx = fa.Context(0)
    prob = Problem(fa, ctx, build_problem(fa, ctx, args.n, order, 0.15, 2024), order, qorder)
    host = prob.host
    for _ in range(3):
        prob.A.assemble(fa.DIFFUSION, qorder, kdiff=prob.k, precision=fa.FP64_FP32)
    fa.assemble_linear(prob.space, prob.b, fa.SRC_CONST, order, (1.0,))
    prob.A.zero_rows(prob.rows, None, prob.b, mode=fa.ELIMINATION_SYMMETRIC)
    for prec, op in ((fa.FP64, fa.INNER_AUTO), (fa.FP32, fa.INNER_ASSEMBLED), (fa.FP32, fa.INNER_ELEMENT)):
        prob.x.zero()
        st = prob.A.solve(prob.b, prob.x, rtol=1e-30, maxit=20, precision=prec, inner_rtol=1e-30, check_every=20,
                          inner_operator=op)
        log(f"precision {prec} inner operator {op}: {st}")
    ctx.synchronize()
    print(json.dumps({"kernels": "copy 1 GiB x6; C2 assembly x3; solves of 20 "
                                 "iterations: FP64, mixed with the assembled FP32 matrix, mixed with the element "
                                 "operator", "memory": stream,
                      "nnz": prob.nnz, "dofs": prob.ndof, "elements": prob.ne}), flush=True)


def bench_beam_kernels(fa, args):
    """Short fixed run of the block SpMV for hardware-counter passes: a plain
    1 GiB device copy (counter calibration), then on the P3 or P4 elasticity
    beam (40x10x10 Kuhn cells, 24 000 tets) one assembly and 40-iteration FP64
    and FP32 solves (3x3 block SELL SpMV kernels)."""
    stream = memory_reference()
    ctx = fa.Context(0)
    k = args.order if args.order in (3, 4) else 4
    coords, elems = fa.mesh_kuhn_box(40, 10, 10, lx=1.0, ly=0.25, lz=0.25, jitter=0.0, seed=args.seed)
    ne = len(elems)
    space = fa.Space.from_mesh(ctx, coords, elems, k, 3)
    g = fa.Graph.from_space(space)
    ndof = space.ndof
    isb = fa.boundary_dofs(3, k, 3, coords, elems, space.elem2dof, ndof, sel=1)
    A = fa.Matrix(g, fa.FP64_FP32)
    if args.spmv_shape >= 0:   # a fixed launch shape instead of the first-solve tuning
        for prec in (fa.FP64, fa.FP32):
            g.set_spmv_shape(prec, args.spmv_shape)
    young = ctx.vector(ne, fa.uniform(ne, 1.0, 100.0, args.seed + 1, log=True))
    poisson = ctx.vector(ne, fa.uniform(ne, 0.2, 0.35, args.seed + 2))
    bvec, x = ctx.vector(ndof), ctx.vector(ndof)
    rows = fa.Rows(ctx, ndof, np.nonzero(isb)[0])
    A.assemble(fa.ELASTICITY, 2 * k, young=young, poisson=poisson, precision=fa.FP64_FP32)
    fa.assemble_linear(space, bvec, fa.SRC_VECTOR, 2 * k, (0.0, 0.0, -1.0))
    A.zero_rows(rows, None, bvec, mode=fa.ELIMINATION_SYMMETRIC)
    ctx.synchronize()
    its = 40
    out = {}
    for name, prec in (("fp64", fa.FP64), ("fp32", fa.FP32)):
        x.zero()
        ctx.synchronize()
        t = time.perf_counter()
        st = A.solve(bvec, x, rtol=1e-30, maxit=its, precision=prec, inner_rtol=1e-30, check_every=its)
        ctx.synchronize()
        out[name] = {"ms": round((time.perf_counter() - t) * 1e3, 2), "its": st["its"]}
        log(f"beam P{k} {name}: {st}")
    nnz = g.nnz
    print(json.dumps({"kernels": f"copy 1 GiB x6; P{k} beam elasticity: one assembly, "
                                 f"{its}-iteration FP64 and FP32 solves", "memory": stream,
                      "dofs": ndof, "nnz": nnz, "block_slots": nnz // 9,
                      "algorithmic_bytes": {"fp64": 8 * nnz + 4 * (nnz // 9) + 20 * ndof,
                                            "fp32": 4 * nnz + 4 * (nnz // 9) + 12 * ndof},
                      "solves": out}), flush=True)


def bench_elastic(fa, args):
    """Short fixed run of the element-matrix kernels for hardware-counter
    passes: a plain 1 GiB device copy (counter calibration), then four FP64+FP32
    bilinear assemblies on each element-matrix path (contracted inside the
    scatter; materialised on the FP64 matrix cores then scattered) for the P4
    elasticity beam (40x10x10 Kuhn cells, 24 000 tets, 811 923 dofs) and for P4
    diffusion on the Kuhn 32^3 cube (196 608 tets, quadrature order 6)."""
    stream = memory_reference()
    ctx = fa.Context(0)
    k = args.order if args.order in (3, 4) else 4
    coords, elems = fa.mesh_kuhn_box(40, 10, 10, lx=1.0, ly=0.25, lz=0.25, jitter=0.0, seed=args.seed)
    ne = len(elems)
    space = fa.Space.from_mesh(ctx, coords, elems, k, 3)
    A = fa.Matrix(fa.Graph.from_space(space), fa.FP64_FP32)
    young = ctx.vector(ne, fa.uniform(ne, 1.0, 100.0, args.seed + 1, log=True))
    poisson = ctx.vector(ne, fa.uniform(ne, 0.2, 0.35, args.seed + 2))
    out = {}
    for name, m in (("fused", fa.MATRICES_FUSED), ("mfma", fa.MATRICES_MFMA)):
        for rep in range(4):
            if rep == 1:
                ctx.synchronize()
                t = time.perf_counter()
            A.assemble(fa.ELASTICITY, 2 * k, young=young, poisson=poisson, precision=fa.FP64_FP32, matrices=m)
        ctx.synchronize()
        out[name] = round((time.perf_counter() - t) / 3 * 1e3, 3)
    # scalar P4 diffusion on the Kuhn 32^3 cube (196 608 tets), the largest problem of the batched mix
    sc, se = fa.mesh_kuhn_box(32, jitter=0.15, seed=args.seed)
    sspace = fa.Space.from_mesh(ctx, sc, se, 4, 1)
    S = fa.Matrix(fa.Graph.from_space(sspace), fa.FP64_FP32)
    kd = ctx.vector(len(se), fa.lognormal(len(se), 0.0, 1.0, args.seed))
    sout = {}
    for name, m in (("fused", fa.MATRICES_FUSED), ("mfma", fa.MATRICES_MFMA)):
        for rep in range(4):
            if rep == 1:
                ctx.synchronize()
                t = time.perf_counter()
            S.assemble(fa.DIFFUSION, 6, kdiff=kd, precision=fa.FP64_FP32, matrices=m)
        ctx.synchronize()
        sout[name] = round((time.perf_counter() - t) / 3 * 1e3, 3)
    print(json.dumps({"kernels": f"copy 1 GiB x6; P{k} beam elasticity and P4 Kuhn 32^3 "
                                 "diffusion assemblies, x4 per element-matrix path", "memory": stream,
                      "elements": ne, "dofs": space.ndof, "bilinear_ms": out,
                      "scalar_p4": {"elements": len(se), "dofs": sspace.ndof, "bilinear_ms": sout}}), flush=True)


def bench_serving(fa, args):
    """Poisson-arrival stream of affine Poisson requests (C4's size/order mix) at several rates."""
    ctx = fa.Context(0)
    orders = [int(o) for o in args.orders.split(",")]
    sizes = [int(n) for n in args.sizes.split(",")]
    fams = affine_families(fa, ctx, orders, sizes)
    keys = list(fams)
    # saturation throughput of the same mix: requests submitted at once (closed loop)
    rng = np.random.default_rng(args.seed)
    sat_draws = [(keys[rng.integers(len(keys))], np.exp(rng.uniform(np.log(0.1), np.log(10.0), 4)))
                 for _ in range(min(args.requests, 128))]
    B = fa.Batch(ctx, rtol=args.rtol, inner_rtol=args.inner_rtol or 1e-2, precision=fa.FP32, use_graph=True,
                 group_spmv=not args.no_group, hbm_budget=int(args.hbm_gib * 2**30), backfill=args.backfill)
    ids = {k: B.add_family(F.terms, F.rhs, F.rows) for k, F in fams.items()}
    run_batch(fa, B, fams, ids, sat_draws[:4], args.chunk)
    _, _, sat_wall, _ = run_batch(fa, B, fams, ids, sat_draws, args.chunk)
    saturation = len(sat_draws) / sat_wall
    log(f"saturation: {saturation:.2f} solves/s ({len(sat_draws)} requests submitted at once)")
    del B
    out = []
    for rate in [float(r) for r in args.rates.split(",")]:
        rng = np.random.default_rng(int(rate) + args.seed)
        nreq = int(min(args.requests, max(20, rate * args.seconds)))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
        draws = [(keys[rng.integers(len(keys))], np.exp(rng.uniform(np.log(0.1), np.log(10.0), 4)))
                 for _ in range(nreq)]
        B = fa.Batch(ctx, rtol=args.rtol, inner_rtol=args.inner_rtol or 1e-2, precision=fa.FP32, use_graph=True,
                     group_spmv=not args.no_group, hbm_budget=int(args.hbm_gib * 2**30), backfill=args.backfill)
        ids = {k: B.add_family(F.terms, F.rhs, F.rows) for k, F in fams.items()}
        run_batch(fa, B, fams, ids, draws[:4], args.chunk)   # warm-up (graph capture, allocations)
        t0 = time.perf_counter()
        nxt, sub, lat = 0, {}, []
        last = t0
        while len(lat) < nreq:
            now = time.perf_counter() - t0
            while nxt < nreq and arrivals[nxt] <= now:
                key, mu = draws[nxt]
                sub[B.submit(ids[key], mu)] = (key, arrivals[nxt])
                nxt += 1
            act, _ = B.step(args.chunk)
            while True:
                got = B.poll()
                if not got:
                    break
                rid, sd = got
                B.solution(rid, fams[sub[rid][0]].ndof)
                done_t = time.perf_counter() - t0
                lat.append(done_t - sub[rid][1])
                last = done_t
            if act == 0 and nxt < nreq:
                time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.01)))
        L = np.array(lat)
        rec = {"rate_per_s": rate, "requests": nreq, "p50_ms": round(float(np.percentile(L, 50)) * 1e3, 2),
               "p99_ms": round(float(np.percentile(L, 99)) * 1e3, 2),
               "sustained_solves_per_s": round(nreq / (last - arrivals[0]), 2),
               "load": round(rate / saturation, 3)}
        log(f"serving {rec}")
        out.append(rec)
    print(json.dumps({"metric": "serving p50/p99 latency and sustained solves/s (one MI355X)", "n_gpus": 1,
                      "config": {"workload": f"Poisson arrivals, affine Poisson P{orders} Kuhn n in {sizes}, "
                                             f"rtol {args.rtol:g}, chunk {args.chunk} CG iterations, "
                                             f"HBM budget {args.hbm_gib:g} GiB, backfill {args.backfill}"},
                      "saturation_solves_per_s": round(saturation, 2), "rates": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50,
                    help="timed assemble+solve steps per form (p50/p99 are over these)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="cubes per side of the Kuhn mesh")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--no-group", action="store_true",
                    help="batched/serving: one SpMV per instance instead of one per family group")
    ap.add_argument("--inner-rtol", type=float, default=None,
                    help="residual reduction between FP64 reliable updates of the FP32 solves (default 1e-2)")
    ap.add_argument("--forms", default="fp64,mixed,mixed_element")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="c2", choices=["c2", "batched", "serving", "beam", "kernels", "elastic", "beamkernels"])
    ap.add_argument("--orders", default="1,2,3,4")
    ap.add_argument("--sizes", default="8,12,16,24,32")
    ap.add_argument("--problems", type=int, default=512)
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--chunk", type=int, default=16)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--rates", default="5,10,20,30,50,200,1000,2000")
    ap.add_argument("--spmv-shape", type=int, default=-1,
                    help="pin the SpMV launch shape (index into the candidate list) instead of tuning it")
    ap.add_argument("--backfill", type=int, default=8,
                    help="serving: steps a request that does not fit the HBM budget lets smaller ones pass it")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--seconds", type=float, default=8.0)
    ap.add_argument("--hbm-gib", type=float, default=64.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import feelpp_amd as fa
    if args.config == "batched":
        if args.rtol == 1e-10:
            args.rtol = 1e-9
        return bench_batched(fa, args)
    if args.config == "kernels":
        return bench_kernels(fa, args)
    if args.config == "elastic":
        return bench_elastic(fa, args)
    if args.config == "beamkernels":
        return bench_beam_kernels(fa, args)
    if args.config == "beam":
        if args.orders == "1,2,3,4":
            args.orders = "3,4"
        return bench_beam(fa, args)
    if args.config == "serving":
        if args.rtol == 1e-10:
            args.rtol = 1e-9
        return bench_serving(fa, args)
    order, qorder = args.order, 2 * args.order
    ctx = fa.Context(0)
    prob = Problem(fa, ctx, build_problem(fa, ctx, args.n, order, 0.15, 2024), order, qorder)
    host = prob.host
    if args.spmv_shape >= 0:   # a fixed launch shape instead of the first-solve tuning
        for prec in (fa.FP64, fa.FP32):
            prob.graph.set_spmv_shape(prec, args.spmv_shape)

    # FP64 reference solution (tight) for the error figure
    st = prob.step(fa.FP64, 1e-13)
    xref = prob.x.get()
    log(f"FP64 reference solve: {st}")

    results = {}
    for form in args.forms.split(","):
        total, lat, sts, timers = time_form(prob, form, args.rtol, args.steps, args.warmup, args.inner_rtol or 1e-2)
        x = prob.x.get()
        err = float(np.linalg.norm(x - xref) / np.linalg.norm(xref))
        results[form] = dict(total=total, lat=lat, st=sts[-1], timers=timers, err=err)
        log(f"{form}: {args.steps / total:.3f} solves/s, p50 {np.median(lat) * 1e3:.1f} ms, err {err:.2e}, "
            f"{sts[-1]}, timers {timers}")

    ok = [f for f in results if results[f]["err"] <= 1e-8 and results[f]["st"]["converged"]]
    best = max(ok, key=lambda f: args.steps / results[f]["total"]) if ok else list(results)[0]
    R = results[best]
    sps = args.steps / R["total"]
    lat = np.array(R["lat"])
    n, nnz = prob.ndof, prob.nnz
    # HBM bytes per launch from the committed counter passes (calibrated FETCH_SIZE + WRITE_SIZE),
    # valid for the workload they were taken on
    traffic = read_traffic(os.path.join(ROOT, "profiles", "hbm_traffic_p2_kuhn64.txt"))
    measured_workload = args.n == 64 and order == 2

    def spmv_roofline(form, timer):
        nl, ms = results[form]["timers"][timer]
        if timer == "spmv_fp64":
            kernel, nbytes = "kSellSpmv<double,1024,2,true>", 12 * nnz + 20 * n + 4
        elif timer == "spmv_fp32":
            kernel, nbytes = "kSellSpmv<float,1024,2,true>", 8 * nnz + 12 * n + 4
        else:
            # element operator: pass 1 streams 6 factors + 10 dof ids per element, writes 10 element
            # results, gathers p once; pass 2 reads one index and one element result per incidence,
            # p, the Dirichlet flag and writes A p per row
            ne, nloc = prob.ne, 10
            kernel = "kElemApplyP2+kElemGather<256>"
            nbytes = ne * (6 * 4 + nloc * 4 + nloc * 4) + 4 * n + ne * nloc * 8 + n * (4 + 1 + 4) + 4 * (n // 64 + 1)
        avg = ms / max(nl, 1)
        ach = nbytes / (avg * 1e-3) / 1e9 if nl else 0.0
        counted = traffic.get(kernel) if measured_workload else None
        out = {"bound": "hbm", "kernel": kernel, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": counted,
               "avg_launch_us": round(avg * 1e3, 2), "launches": nl, "algorithmic_bytes": nbytes}
        if counted and nl:
            # the same launches measured against the counted HBM bytes (calibrated FETCH_SIZE + WRITE_SIZE)
            out["achieved_counted"] = round(counted / (avg * 1e-3) / 1e9, 1)
            out["frac_counted"] = round(counted / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        return out

    if best == "fp64":
        roof = spmv_roofline(best, "spmv_fp64")
    elif R["timers"]["elemop_fp32"][0] > 0:
        roof = spmv_roofline(best, "elemop_fp32")
    else:
        roof = spmv_roofline(best, "spmv_fp32")
    others = {}
    for f, timer in (("mixed", "spmv_fp32"), ("mixed_element", "elemop_fp32"), ("fp64", "spmv_fp64")):
        if f in results and f != best and results[f]["timers"][timer][0] > 0:
            others[f] = spmv_roofline(f, timer)
    # scatter (owner-computes gather of element contributions into the FP64 CSR values):
    # algorithmic bytes = nnz*8 values + E*nloc*4 elem2dof + E*8 coefficient + E*4*4 + V*3*8 geometry
    ng, gms = R["timers"]["assemble_gather"]
    nv = (args.n + 1) ** 3
    nloc = (order + 1) * (order + 2) * (order + 3) // 6
    bytes_scatter = 8 * nnz + prob.ne * nloc * 4 + prob.ne * 8 + prob.ne * 16 + nv * 24
    g_ms = gms / max(ng, 1)
    g_ach = bytes_scatter / (g_ms * 1e-3) / 1e9 if ng else 0.0
    scatter = {"kernel": "kGatherScalar<100,8,true>" if order == 2 else "kGatherScalar", "bound": "hbm",
               "achieved": round(g_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(g_ach / HBM_PEAK_GBS, 4), "avg_launch_us": round(g_ms * 1e3, 2),
               "algorithmic_bytes": bytes_scatter,
               "traffic": traffic.get("kGatherScalar<100,8,true>") if measured_workload else None}

    cpu = None
    if not args.no_cpu:
        import oracle as orc
        cpu = cpu_baseline(orc, host, order, qorder, args.rtol, xref)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC,
        "value": round(sps, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(R["total"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"P2 Poisson, Kuhn {args.n}^3 unit cube, lognormal(0,1) P0 conductivity, f=1, "
                        f"homogeneous Dirichlet, quadrature order {qorder}, Jacobi-CG rtol {args.rtol:g}",
            "elements": prob.ne, "dofs": n, "nnz": nnz, "form": best,
        },
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "elements_per_s": round(prob.ne * sps, 1),
        "nnz_per_s": round(nnz * sps, 1),
        "iterations": R["st"]["its"],
        "outer_iterations": R["st"]["outer_its"],
        "rel_err_vs_fp64": R["err"],
        "forms": {f: {"solves_per_s": round(args.steps / v["total"], 4), "its": v["st"]["its"],
                      "rel_err_vs_fp64": v["err"], "p50_ms": round(float(np.median(v["lat"])) * 1e3, 3)}
                  for f, v in results.items()},
        "roofline": dict(roof, scatter=scatter, other_forms=others),
        "cpu_baseline": cpu and {k: v for k, v in cpu.items()
                                 if k in ("value", "unit", "cores", "kind", "sample", "one_core")},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
