// Jacobi-preconditioned CG on the device: plain FP64 and FP32-inside-FP64
// iterative refinement.  Semantics follow SolverLinearPetsc::solve
// (feelalg/solverlinearpetsc.cpp:575-640) with KSPCG + PCJACOBI: stopping test
// on the preconditioned residual norm relative to the initial one (UIRNorm,
// :610), KSPSetTolerances(rtol, atol, dtol, maxit) (:592).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

using namespace feelpp_amd;

namespace {

#define TRY_CALL(expr)                       \
    do {                                     \
        int rc_ = (expr);                    \
        if (rc_ != FEEL_HIP_OK) return rc_;  \
    } while (0)

// y = A x in the solver numbering (SELL image), optional p.Ap partials
template <typename T>
void spmv(feel_hip_ctx* c, const feel_hip_graph* g, const T* val, const T* x, T* y, double* part, const CgState* st) {
    constexpr int U = sizeof(T) == 8 ? 4 : 8;
    if (g->compact)
        hipLaunchKernelGGL((kSellSpmv<T, kSpmvBlock, U, true>), dim3(kSpmvGrid), dim3(kSpmvBlock), 0, c->stream, g->n,
                           g->nslices, g->slicePtr, g->scol, g->colBase, g->ccptr, g->packed, val, x, y, part, st);
    else
        hipLaunchKernelGGL((kSellSpmv<T, kSpmvBlock, U, false>), dim3(kSpmvGrid), dim3(kSpmvBlock), 0, c->stream, g->n,
                           g->nslices, g->slicePtr, g->scol, g->colBase, g->ccptr, g->packed, val, x, y, part, st);
}

int gridN(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 8192)); }

template <typename T>
int allocWork(feel_hip_ctx* c, T** p, int64_t n) {
    if (*p) return FEEL_HIP_OK;
    return c->alloc((void**)p, sizeof(T) * n);
}

// Refresh the SELL values (and the Jacobi diagonal, solver numbering) from the CSR values.
int ensureSell(feel_hip_mat* A, bool need64, bool need32) {
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    bool changed = false;
    if (need64 && !A->sell64Valid) {
        if (!A->v64) return setError(FEEL_HIP_EINVAL, "matrix holds no FP64 values");
        TRY_CALL(allocWork(c, &A->s64, g->nslots));
        hipLaunchKernelGGL(kCsrToSell<double>, dim3(gridN(g->nslots)), dim3(kBlock), 0, c->stream, g->nslots, g->slot2csr,
                           A->v64, (const float*)nullptr, A->s64);
        A->sell64Valid = changed = true;
    }
    if (need32 && !A->sell32Valid) {
        TRY_CALL(allocWork(c, &A->s32, g->nslots));
        hipLaunchKernelGGL(kCsrToSell<float>, dim3(gridN(g->nslots)), dim3(kBlock), 0, c->stream, g->nslots, g->slot2csr,
                           A->v32 ? (const double*)nullptr : A->v64, A->v32, A->s32);
        A->sell32Valid = changed = true;
    }
    if (changed || !A->dinvValid) {
        hipLaunchKernelGGL(kDiagInvSell, dim3(gridN(g->n)), dim3(kBlock), 0, c->stream, g->n, g->sdiag,
                           A->sell64Valid ? A->s64 : nullptr, A->sell32Valid ? A->s32 : nullptr,
                           A->sell64Valid ? A->dinv64 : nullptr, A->dinv32);
        A->dinvValid = true;
    }
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

struct CgRun {
    int its = 0, reason = 0;
    double dp = 0, rnorm0 = 0;
};

// Jacobi-PCG on (val, dinv) from x (initial guess, zero when x0zero) for b.
// Kernel sequence per iteration it: direction(it) -> spmv+p.Ap -> update(it).
template <typename T>
int runCg(feel_hip_mat* A, const T* val, const T* dinv, const T* b, T* x, bool x0zero, double rtol, double atol,
          double dtol, int maxit, int checkEvery, T* r, T* z, T* p, T* w, const char* timerName, CgRun& out) {
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    const int64_t n = g->n;
    double* zrPart = c->part;
    double* zzPart = c->part + kReduceBlocks;
    double* dpiPart = c->part + 3 * kReduceBlocks;   // kSpmvGrid partials
    CgState* st = c->state;
    double* betaSlots = reinterpret_cast<double*>(st + 1);
    double* dpiSlots = betaSlots + 2;
    CgState h{};
    h.rtol = rtol; h.atol = atol; h.dtol = dtol; h.maxit = maxit;
    FEEL_HIP_CHECK(hipMemcpyAsync(st, &h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    if (x0zero)
        FEEL_HIP_CHECK(hipMemsetAsync(w, 0, sizeof(T) * n, c->stream));
    else
        spmv<T>(c, g, val, x, w, nullptr, nullptr);
    hipLaunchKernelGGL(kCgInit<T>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, b, w, r, z, dinv, zrPart, zzPart);
    FEEL_HIP_CHECK(hipGetLastError());
    int it = 0;
    while (true) {
        const int chunk = checkEvery > 0 ? checkEvery : std::min(64, std::max(8, it / 4));
        for (int k = 0; k < chunk; ++k, ++it) {
            hipLaunchKernelGGL(kCgDirection<T>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, it, z, p, zrPart, zzPart,
                               st, betaSlots);
            c->tic(timerName);
            spmv<T>(c, g, val, p, w, dpiPart, st);
            c->toc();
            hipLaunchKernelGGL(kCgUpdate<T>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, it, x, r, z, p, w, dinv,
                               dpiPart, zrPart, zzPart, st, betaSlots, dpiSlots);
            if (it > maxit) break;
        }
        FEEL_HIP_CHECK(hipGetLastError());
        FEEL_HIP_CHECK(hipMemcpyAsync(c->hstate, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
        FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
        if (c->hstate->done) break;
        if (it > maxit + 1) return setError(FEEL_HIP_EDEVICE, "CG state machine did not terminate");
    }
    out.its = c->hstate->its;
    out.reason = c->hstate->reason;
    out.dp = c->hstate->dp;
    out.rnorm0 = c->hstate->rnorm0;
    return FEEL_HIP_OK;
}

// FP64 residual r = b - A x and its preconditioned norm (host-side fixed-order sum)
int residual64(feel_hip_mat* A, const double* b, const double* x, double* r, double* w, double& dp) {
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    c->tic("spmv_fp64");
    spmv<double>(c, g, A->s64, x, w, nullptr, nullptr);
    c->toc();
    hipLaunchKernelGGL(kResidual64, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, g->n, b, w, A->dinv64, r,
                       c->part + 2 * kReduceBlocks);
    FEEL_HIP_CHECK(hipGetLastError());
    std::vector<double> h(kReduceBlocks);
    FEEL_HIP_CHECK(hipMemcpyAsync(h.data(), c->part + 2 * kReduceBlocks, sizeof(double) * kReduceBlocks,
                                  hipMemcpyDeviceToHost, c->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    double s = 0;
    for (double v : h) s += v;
    dp = std::sqrt(s);
    return FEEL_HIP_OK;
}

}  // namespace

extern "C" {

void feel_hip_solve_opts_default(feel_hip_solve_opts* o) {
    if (!o) return;
    // Feel++ backend defaults: ksp-rtol 1e-8, ksp-maxit 1000 (feelalg/backend.cpp:732-762,941);
    // PETSc defaults atol 1e-50, dtol 1e5.
    o->rtol = 1e-8;
    o->atol = 1e-50;
    o->dtol = 1e5;
    o->maxit = 1000;
    o->precision = FEEL_HIP_FP64;
    o->inner_rtol = 1e-4;
    o->inner_maxit = 100000;
    o->max_outer = 50;
    o->check_every = 0;
}

int feel_hip_solve_pcg(feel_hip_mat* A, const feel_hip_vec* b, feel_hip_vec* x, const feel_hip_solve_opts* opts,
                       feel_hip_solve_data* out) {
    if (!A || !b || !x || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    const int64_t n = g->n;
    if (b->n != n || x->n != n) return setError(FEEL_HIP_EINVAL, "vector size mismatch");
    feel_hip_solve_opts o;
    if (opts) o = *opts;
    else feel_hip_solve_opts_default(&o);
    if (o.maxit < 0 || !(o.rtol >= 0) || !(o.atol >= 0)) return setError(FEEL_HIP_EINVAL, "bad tolerances");
    *out = feel_hip_solve_data{};
    if (o.precision == FEEL_HIP_FP64) {
        if (!A->v64) return setError(FEEL_HIP_EINVAL, "FP64 solve needs FP64 matrix values");
        TRY_CALL(allocWork(c, &A->r64, n));
        TRY_CALL(allocWork(c, &A->z64, n));
        TRY_CALL(allocWork(c, &A->p64, n));
        TRY_CALL(allocWork(c, &A->w64, n));
        TRY_CALL(allocWork(c, &A->bp64, n));
        TRY_CALL(allocWork(c, &A->xp64, n));
        TRY_CALL(ensureSell(A, true, false));
        hipLaunchKernelGGL(kPermuteIn, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, b->d, A->bp64);
        hipLaunchKernelGGL(kPermuteIn, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, x->d, A->xp64);
        CgRun run;
        TRY_CALL(runCg<double>(A, A->s64, A->dinv64, A->bp64, A->xp64, false, o.rtol, o.atol, o.dtol, o.maxit,
                               o.check_every, A->r64, A->z64, A->p64, A->w64, "spmv_fp64", run));
        hipLaunchKernelGGL(kPermuteOut, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, A->xp64, x->d);
        FEEL_HIP_CHECK(hipGetLastError());
        c->collectTimers();
        out->its = run.its;
        out->reason = run.reason;
        out->residual = run.dp;
        out->rnorm0 = run.rnorm0;
        out->outer_its = 1;
        out->converged = run.reason > 0;
        return FEEL_HIP_OK;
    }
    if (o.precision != FEEL_HIP_FP32) return setError(FEEL_HIP_EINVAL, "solve precision must be FP64 or FP32");
    // FP32 Jacobi-CG inside FP64 iterative refinement: FP64 residual, FP64 stopping test.
    if (!A->v64) return setError(FEEL_HIP_EINVAL, "mixed-precision solve needs FP64 matrix values for the residual");
    TRY_CALL(allocWork(c, &A->r64, n));
    TRY_CALL(allocWork(c, &A->w64, n));
    TRY_CALL(allocWork(c, &A->r32, n));
    TRY_CALL(allocWork(c, &A->z32, n));
    TRY_CALL(allocWork(c, &A->p32, n));
    TRY_CALL(allocWork(c, &A->w32, n));
    TRY_CALL(allocWork(c, &A->x32, n));
    TRY_CALL(allocWork(c, &A->b32, n));
    TRY_CALL(allocWork(c, &A->bp64, n));
    TRY_CALL(allocWork(c, &A->xp64, n));
    TRY_CALL(ensureSell(A, true, true));
    hipLaunchKernelGGL(kPermuteIn, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, b->d, A->bp64);
    hipLaunchKernelGGL(kPermuteIn, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, x->d, A->xp64);
    double dp = 0;
    TRY_CALL(residual64(A, A->bp64, A->xp64, A->r64, A->w64, dp));
    const double rnorm0 = dp, ttol = std::max(o.rtol * dp, o.atol);
    int total = 0, outer = 0, reason = 0;
    while (true) {
        if (dp != dp) { reason = -9; break; }
        if (dp <= ttol) { reason = dp < o.atol ? 3 : 2; break; }
        if (outer > 0 && dp >= o.dtol * rnorm0) { reason = -4; break; }
        if (total >= o.maxit || outer >= o.max_outer) { reason = -3; break; }
        const double s = 1.0 / dp;
        hipLaunchKernelGGL(kScaleToF32, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, A->r64, s, A->b32);
        FEEL_HIP_CHECK(hipMemsetAsync(A->x32, 0, sizeof(float) * n, c->stream));
        CgRun run;
        const int budget = std::min(o.inner_maxit, o.maxit - total);
        TRY_CALL(runCg<float>(A, A->s32, A->dinv32, A->b32, A->x32, true, o.inner_rtol, 0.0, 1e5, budget, o.check_every,
                              A->r32, A->z32, A->p32, A->w32, "spmv_fp32", run));
        total += run.its;
        ++outer;
        if (run.reason == -8 || run.reason == -9) { reason = run.reason; break; }
        hipLaunchKernelGGL(kAxpyF32, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, dp, A->x32, A->xp64);
        FEEL_HIP_CHECK(hipGetLastError());
        TRY_CALL(residual64(A, A->bp64, A->xp64, A->r64, A->w64, dp));
    }
    hipLaunchKernelGGL(kPermuteOut, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, A->xp64, x->d);
    FEEL_HIP_CHECK(hipGetLastError());
    c->collectTimers();
    out->its = total;
    out->reason = reason;
    out->residual = dp;
    out->rnorm0 = rnorm0;
    out->outer_its = outer;
    out->converged = reason > 0;
    return FEEL_HIP_OK;
}

}  // extern "C"
