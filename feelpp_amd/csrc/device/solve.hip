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
#include "../host/krylov.hpp"

using namespace feelpp_amd;

namespace {

#define TRY_CALL(expr)                       \
    do {                                     \
        int rc_ = (expr);                    \
        if (rc_ != FEEL_HIP_OK) return rc_;  \
    } while (0)

SellCols sellCols(const feel_hip_graph* g) { return SellCols{g->slicePtr, g->scol, g->colBase, g->ccptr, g->packed}; }

// SpMV launch shapes the autotuner chooses from: (block size, grid).  The
// per-lane unroll is fixed per precision, so every shape computes each row
// with the same operations in the same order; shapes differ only in how rows
// map to waves and in the number of p.Ap block partials.
struct SpmvShape { int block, grid; };
constexpr SpmvShape kSpmvShapes[] = {{1024, 2048}, {1024, 1024}, {1024, 4096}, {512, 2048}, {512, 1024},
                                     {256, 1024}, {256, 2048}, {256, 4096}};
constexpr int kNumSpmvShapes = sizeof(kSpmvShapes) / sizeof(kSpmvShapes[0]);
// a shape replaces the default only when it is this much faster, so that the
// pick is stable from run to run where shapes are close
constexpr double kShapeMargin = 0.95;

template <typename T, int BS, bool COMPACT>
void launchSpmv(feel_hip_ctx* c, const feel_hip_graph* g, int grid, const T* val, const T* x, T* y, double* part,
                const CgState* st, const T* dinv, double* part2) {
    constexpr int U = 2;   // quads per step
    hipLaunchKernelGGL((kSellSpmv<T, BS, U, COMPACT>), dim3(grid), dim3(BS), 0, c->stream, g->n, g->nslices,
                       sellCols(g), val, x, y, part, st, dinv, part2);
}

// (dinv, part2: the y.D^{-1}y partials, scalar image only)
template <typename T>
void launchShape(feel_hip_ctx* c, const feel_hip_graph* g, const SpmvShape& sh, const T* val, const T* x, T* y,
                 double* part, const CgState* st, const T* dinv = nullptr, double* part2 = nullptr) {
    if (g->block) {   // the block image: per-slice lanes per node row
        if (sh.block == 1024)
            hipLaunchKernelGGL((kBlockSellSpmv<T, 1024>), dim3(sh.grid), dim3(1024), 0, c->stream, g->nslices,
                               g->slicePtr, g->sliceRow, g->sliceShift, g->scol, val, x, y, part, st);
        else if (sh.block == 512)
            hipLaunchKernelGGL((kBlockSellSpmv<T, 512>), dim3(sh.grid), dim3(512), 0, c->stream, g->nslices,
                               g->slicePtr, g->sliceRow, g->sliceShift, g->scol, val, x, y, part, st);
        else
            hipLaunchKernelGGL((kBlockSellSpmv<T, 256>), dim3(sh.grid), dim3(256), 0, c->stream, g->nslices,
                               g->slicePtr, g->sliceRow, g->sliceShift, g->scol, val, x, y, part, st);
        return;
    }
    if (g->compact) {
        if (sh.block == 1024) launchSpmv<T, 1024, true>(c, g, sh.grid, val, x, y, part, st, dinv, part2);
        else if (sh.block == 512) launchSpmv<T, 512, true>(c, g, sh.grid, val, x, y, part, st, dinv, part2);
        else launchSpmv<T, 256, true>(c, g, sh.grid, val, x, y, part, st, dinv, part2);
    } else {
        if (sh.block == 1024) launchSpmv<T, 1024, false>(c, g, sh.grid, val, x, y, part, st, dinv, part2);
        else if (sh.block == 512) launchSpmv<T, 512, false>(c, g, sh.grid, val, x, y, part, st, dinv, part2);
        else launchSpmv<T, 256, false>(c, g, sh.grid, val, x, y, part, st, dinv, part2);
    }
}

// Autotuner: times every candidate shape on this graph's matrix (2 warm-up
// and 5 timed launches each, HIP events on the context stream) and keeps the
// fastest, or the default unless another beats it by the margin.  Runs once
// per graph and precision; part receives partials of at most kSpmvPartCap blocks.
template <typename T>
int tuneSpmv(feel_hip_ctx* c, feel_hip_graph* g, const T* val, const T* x, T* y, double* part) {
    int& pick = g->spmvShape[sizeof(T) == 8 ? 0 : 1];
    hipEvent_t a, b;
    FEEL_HIP_CHECK(hipEventCreate(&a));
    FEEL_HIP_CHECK(hipEventCreate(&b));
    double best = 1e300, def = 1e300;
    int bestI = 0;
    for (int i = 0; i < kNumSpmvShapes; ++i) {
        for (int w = 0; w < 2; ++w) launchShape<T>(c, g, kSpmvShapes[i], val, x, y, part, nullptr);
        (void)hipEventRecord(a, c->stream);
        for (int r = 0; r < 5; ++r) launchShape<T>(c, g, kSpmvShapes[i], val, x, y, part, nullptr);
        (void)hipEventRecord(b, c->stream);
        FEEL_HIP_CHECK(hipEventSynchronize(b));
        float ms = 0;
        FEEL_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        if (i == 0) def = ms;
        if (ms < best) { best = ms; bestI = i; }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    FEEL_HIP_CHECK(hipGetLastError());
    pick = best < kShapeMargin * def ? bestI : 0;
    return FEEL_HIP_OK;
}

// y = A x in the solver numbering (SELL image), optional p.Ap partials (and,
// on the scalar image, y.D^{-1}y partials in part2), with the graph's tuned
// launch shape (tuned on the first call).  Returns the number of partials
// written (the grid).
template <typename T>
int spmv(feel_hip_ctx* c, feel_hip_graph* g, const T* val, const T* x, T* y, double* part, const CgState* st,
         int* npart, const T* dinv = nullptr, double* part2 = nullptr) {
    const int idx = sizeof(T) == 8 ? 0 : 1;
    if (part2 && g->block) return setError(FEEL_HIP_EUNSUPPORTED, "y.D^-1y partials are computed on the scalar image only");
    if (g->spmvShape[idx] < 0) TRY_CALL(tuneSpmv<T>(c, g, val, x, y, c->part + 3 * kReduceBlocks));
    const SpmvShape& sh = kSpmvShapes[g->spmvShape[idx]];
    launchShape<T>(c, g, sh, val, x, y, part, st, dinv, part2);
    if (npart) *npart = sh.grid;
    return FEEL_HIP_OK;
}

int gridN(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 8192)); }

template <typename T>
int allocWork(feel_hip_ctx* c, T** p, int64_t n) {
    if (*p) return FEEL_HIP_OK;
    return c->alloc((void**)p, sizeof(T) * n);
}

// Refresh the SELL values (and the Jacobi diagonal, solver numbering) from the
// CSR values.  On the scalar image the FP32 values are those of
// D^{-1/2} A D^{-1/2}, rounded from the FP64 image, when every diagonal entry
// is positive (s32Scaled: the FP32 iteration then runs on the scaled system);
// otherwise, and on the block image, those of A.
int ensureSell(feel_hip_mat* A, bool need64, bool need32) {
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    bool changed = false;
    const bool scalable = need32 && !g->block && A->v64;
    if (need64 && need32 && !scalable && !A->sell64Valid && !A->sell32Valid && A->v64) {
        TRY_CALL(allocWork(c, &A->s64, g->nslots));
        TRY_CALL(allocWork(c, &A->s32, g->nslots));
        hipLaunchKernelGGL(kCsrToSellBoth, dim3(gridN(g->nslots)), dim3(kBlock), 0, c->stream, g->nslots, g->slot2csr,
                           A->v64, A->v32, A->s64, A->s32);
        A->sell64Valid = A->sell32Valid = changed = true;
        A->s32Scaled = false;
    }
    if ((need64 || scalable) && !A->sell64Valid) {
        if (!A->v64) return setError(FEEL_HIP_EINVAL, "matrix holds no FP64 values");
        TRY_CALL(allocWork(c, &A->s64, g->nslots));
        hipLaunchKernelGGL(kCsrToSell<double>, dim3(gridN(g->nslots)), dim3(kBlock), 0, c->stream, g->nslots, g->slot2csr,
                           A->v64, (const float*)nullptr, A->s64);
        A->sell64Valid = changed = true;
    }
    if (need32 && !A->sell32Valid && !scalable) {
        TRY_CALL(allocWork(c, &A->s32, g->nslots));
        hipLaunchKernelGGL(kCsrToSell<float>, dim3(gridN(g->nslots)), dim3(kBlock), 0, c->stream, g->nslots, g->slot2csr,
                           A->v32 ? (const double*)nullptr : A->v64, A->v32, A->s32);
        A->sell32Valid = changed = true;
        A->s32Scaled = false;
    }
    if (changed || !A->dinvValid) {
        hipLaunchKernelGGL(kDiagInvSell, dim3(gridN(g->n)), dim3(kBlock), 0, c->stream, g->n, g->sdiag,
                           A->sell64Valid ? A->s64 : nullptr,
                           A->sell32Valid && !A->s32Scaled ? A->s32 : nullptr,
                           A->sell64Valid ? A->dinv64 : nullptr, A->dinv32);
        A->dinvValid = true;
    }
    if (scalable && !A->sell32Valid) {
        TRY_CALL(allocWork(c, &A->s32, g->nslots));
        TRY_CALL(allocWork(c, &A->dsq64, g->n));
        TRY_CALL(allocWork(c, &A->diagFlag, 1));
        FEEL_HIP_CHECK(hipMemsetAsync(A->diagFlag, 0, sizeof(int), c->stream));
        hipLaunchKernelGGL(kDiagSqrt, dim3(gridN(g->n)), dim3(kBlock), 0, c->stream, g->n, A->dinv64, A->dsq64,
                           A->diagFlag);
        int bad = 0;
        FEEL_HIP_CHECK(hipMemcpyAsync(&bad, A->diagFlag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
        if (!bad) {
            const int grid = (int)std::min<int64_t>((g->nslices + kBlock / 64 - 1) / (kBlock / 64), 16384);
            hipLaunchKernelGGL(kSellScale32, dim3(std::max(grid, 1)), dim3(kBlock), 0, c->stream, g->n, g->nslices,
                               g->slicePtr, g->scol, A->s64, A->dsq64, A->s32);
        } else {
            hipLaunchKernelGGL(kCsrToSell<float>, dim3(gridN(g->nslots)), dim3(kBlock), 0, c->stream, g->nslots,
                               g->slot2csr, A->v64, (const float*)nullptr, A->s32);
        }
        A->s32Scaled = !bad;
        A->sell32Valid = true;
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
    double* dpiPart = c->part + 3 * kReduceBlocks;   // up to kSpmvPartCap partials
    CgState* st = c->state;
    double* betaSlots = reinterpret_cast<double*>(st + 1);
    double* dpiSlots = betaSlots + 2;
    CgState h{};
    h.rtol = rtol; h.atol = atol; h.dtol = dtol; h.maxit = maxit;
    FEEL_HIP_CHECK(hipMemcpyAsync(st, &h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    if (x0zero)
        FEEL_HIP_CHECK(hipMemsetAsync(w, 0, sizeof(T) * n, c->stream));
    else
        TRY_CALL(spmv<T>(c, g, val, x, w, nullptr, nullptr, nullptr));
    hipLaunchKernelGGL(kCgInit<T>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, b, w, r, z, dinv, zrPart, zzPart);
    FEEL_HIP_CHECK(hipGetLastError());
    int it = 0;
    while (true) {
        const int chunk = checkEvery > 0 ? checkEvery : std::min(64, std::max(8, it / 4));
        for (int k = 0; k < chunk; ++k, ++it) {
            hipLaunchKernelGGL(kCgDirection<T>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, it, z, (const T*)p, p,
                               zrPart, zzPart, st, betaSlots, (double*)nullptr);
            c->tic(timerName);
            int npart = 0;
            TRY_CALL(spmv<T>(c, g, val, p, w, dpiPart, st, &npart));
            c->toc();
            hipLaunchKernelGGL(kCgUpdate<T>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, it, x, r, z, p, w, dinv,
                               dpiPart, npart, zrPart, zzPart, st, betaSlots, dpiSlots, (double*)nullptr);
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
    TRY_CALL(spmv<double>(c, g, A->s64, x, w, nullptr, nullptr, nullptr));
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

struct MixedRun {
    int its = 0, reason = 0, updates = 0, fp64Its = 0;
    double dp = 0, rnorm0 = 0;
};

// Mixed precision pays while eps32 * kappa stays small: an FP32 operator and
// FP32 Krylov vectors perturb CG by about that much per iteration, and beyond
// kappa ~ 5e4 the FP32 iterations outnumber the FP64 ones by more than the
// FP32 iteration is cheaper (about 0.6x the FP64 one, HBM-bound).
constexpr double kMixedKappaLimit = 5e4;
constexpr int kKappaMinIts = 30;
// the estimate is also taken at host polls during the first iterations of a
// chain (every kKappaEvery up to kKappaPollIts), so an ill-conditioned solve
// hands over early instead of at its first reliable update
constexpr int kKappaEvery = 128;
constexpr int kKappaPollIts = 1024;

// Jacobi-PCG in FP32 with reliable updates in FP64 (solver numbering, on
// A->bp64 / A->xp64).  One FP32 CG runs on the scaled residual r / sigma and
// accumulates its correction in xs: on the assembled scalar image two kernels
// per iteration (SpMV with the p.Ap and w.D^{-1}w partials, then kCgStep: the
// update fused with the next direction), otherwise the direction / SpMV /
// update kernels of the FP64 solver.  Whenever its preconditioned residual norm has dropped by
// inner_rtol since the last reliable update, or reaches the tolerance, the
// device stops and the host folds sigma * xs into the FP64 solution,
// recomputes the true FP64 residual and tests it (KSPConvergedDefault on the
// FP64 preconditioned residual), then restarts the FP32 residual from it and
// continues with the same search direction, so the Krylov recurrence is not
// restarted: beta = (r_new, z_new) / (r_old, z_old).
// A reliable update whose true residual exceeds the FP32 estimate by more than
// this factor restarts the search direction (the FP32 recurrence drifted).
constexpr double kDriftRestart = 2.0;

int runMixedCg(feel_hip_mat* A, const feel_hip_solve_opts& o, bool elemOp, MixedRun& out) {
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    const int64_t n = g->n;
    double* zrPart = c->part;
    double* zzPart = c->part + kReduceBlocks;
    double* dpiPart = c->part + 3 * kReduceBlocks;   // up to kSpmvPartCap partials
    // the pipelined iteration (assembled FP32 values on the scalar image):
    // w.D^{-1}w partials of the SpMV, and a second z.r / z.z set and search
    // direction buffer, taken by iteration parity
    const bool pipelined = !elemOp && !g->block;
    // on the scaled FP32 image (every diagonal entry positive) the iteration
    // runs on D^{-1/2} A D^{-1/2}: no D^{-1} stream in the SpMV or the update
    const bool scaled = pipelined && A->s32Scaled;
    double* wdwPart = dpiPart + kSpmvPartCap;
    double* zrPartB = wdwPart + kSpmvPartCap;
    double* zzPartB = zrPartB + kReduceBlocks;
    auto zrSet = [&](int k) { return pipelined && (k & 1) ? zrPartB : zrPart; };
    auto zzSet = [&](int k) { return pipelined && (k & 1) ? zzPartB : zzPart; };
    if (pipelined) TRY_CALL(allocWork(c, &A->q32, n));
    float* const dirs[2] = {A->p32, pipelined ? A->q32 : A->p32};
    CgState* st = c->state;
    double* betaSlots = reinterpret_cast<double*>(st + 1);
    double* dpiSlots = betaSlots + 2;
    double dp = 0;
    TRY_CALL(residual64(A, A->bp64, A->xp64, A->r64, A->w64, dp));
    const double rnorm0 = dp, ttol = std::max(o.rtol * dp, o.atol);
    const double delta = o.inner_rtol > 0 && o.inner_rtol < 1 ? o.inner_rtol : 1e-2;
    out.rnorm0 = rnorm0;
    out.dp = dp;
    if (dp != dp) { out.reason = -9; return FEEL_HIP_OK; }
    if (dp <= ttol) { out.reason = dp < o.atol ? 3 : 2; return FEEL_HIP_OK; }
    if (o.maxit == 0) { out.reason = -3; return FEEL_HIP_OK; }
    // the device works in units of sigma: its residual norm starts at 1 and
    // its stopping threshold is stop / sigma
    const double sigma = dp;
    CgState h{};
    h.rtol = 0.0; h.dtol = o.dtol; h.maxit = o.maxit;
    int it = 0;
    // scaled image: the device norm of a chain is ||r~|| = ||D^{-1/2} r|| / sigma,
    // not the preconditioned one, so each chain's thresholds are taken relative
    // to its starting norm chainStart, and chainScale converts its norms back
    double chainStart = 1.0, chainScale = sigma;
    std::vector<double> hpart(kReduceBlocks);
    auto reset = [&](double dpNow, bool restart) -> int {
        double stop = std::max(ttol, delta * dpNow) / sigma;
        if (scaled) {
            hipLaunchKernelGGL(kCgScaledReset, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, A->r64, A->dsq64,
                               1.0 / sigma, A->r32, A->x32, zrSet(it), zzSet(it));
            FEEL_HIP_CHECK(hipGetLastError());
            FEEL_HIP_CHECK(hipMemcpyAsync(hpart.data(), zzSet(it), sizeof(double) * kReduceBlocks,
                                          hipMemcpyDeviceToHost, c->stream));
            FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
            double rr = 0;
            for (double v : hpart) rr += v;
            chainStart = std::sqrt(rr);
            chainScale = chainStart > 0 ? dpNow / chainStart : sigma;
            stop = chainStart * (std::max(ttol, delta * dpNow) / dpNow);
        }
        h.atol = stop;   // threshold at iteration 0
        h.ttol = stop;   // threshold afterwards
        h.done = 0; h.reason = 0; h.its = it;
        if (restart) h.fresh = it;
        FEEL_HIP_CHECK(hipMemcpyAsync(st, &h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
        if (!scaled)
            hipLaunchKernelGGL(kCgStepReset, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, A->r64, 1.0 / sigma,
                               A->dinv32, A->r32, A->z32, A->x32, zrSet(it), zzSet(it));
        FEEL_HIP_CHECK(hipGetLastError());
        // h must stay valid until the copy has run
        FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
        return FEEL_HIP_OK;
    };
    TRY_CALL(reset(dp, true));
    // the pipelined iteration starts each chain (and resumes after a reliable
    // update) with the classic direction kernel: stopping test, p = z + b p_prev
    bool needDirection = true;
    int reason = 0;
    // Launches queued past the iteration that stops the device are early
    // exits; the chunk is sized from the observed convergence rate so that
    // few are wasted before each reliable update.
    double lastDp = scaled ? chainStart : 1.0;
    int lastIt = it, chunk = 16, lastKappa = 0;
    // condition estimate of the current Lanczos chain (since the last direction restart)
    auto tooIllConditioned = [&](bool& ill) -> int {
        const int k = std::min(it, kHistMax) - h.fresh;
        ill = false;
        if (k < kKappaMinIts) return FEEL_HIP_OK;
        std::vector<double> hh(2 * (size_t)k);
        FEEL_HIP_CHECK(hipMemcpy(hh.data(), c->hist + 2 * (size_t)h.fresh, sizeof(double) * hh.size(),
                                 hipMemcpyDeviceToHost));
        // an order-of-magnitude test: the largest eigenvalue to 1 % and one
        // Sturm count at lmax / limit, instead of resolving the smallest
        ill = cgConditionExceeds(hh.data(), k, kMixedKappaLimit);
        return FEEL_HIP_OK;
    };
    while (true) {
        if (o.check_every > 0) chunk = o.check_every;
        for (int k = 0; k < chunk && it <= o.maxit; ++k, ++it) {
            if (pipelined) {
                // direction it in dirs[it & 1], the previous one in the other buffer
                float* p = dirs[it & 1];
                const int tested = needDirection ? 1 : 0;
                if (needDirection) {
                    hipLaunchKernelGGL(kCgDirection<float>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, it,
                                       scaled ? A->r32 : A->z32, (const float*)dirs[(it + 1) & 1], p, zrSet(it),
                                       zzSet(it), st, betaSlots, c->hist);
                    needDirection = false;
                }
                int npart = 0;
                c->tic("spmv_fp32");
                TRY_CALL(spmv<float>(c, g, A->s32, p, A->w32, dpiPart, st, &npart,
                                     scaled ? (const float*)nullptr : A->dinv32, wdwPart));
                c->toc();
                if (scaled)
                    hipLaunchKernelGGL((kCgStep<float, true>), dim3(kStepBlocks), dim3(kStepBlock), 0, c->stream, n, it,
                                       A->x32, A->r32, (const float*)p, dirs[(it + 1) & 1], A->w32,
                                       (const float*)nullptr, dpiPart, wdwPart, npart, zrSet(it), zzSet(it),
                                       zrSet(it + 1), zzSet(it + 1), st, betaSlots, dpiSlots, c->hist, tested);
                else
                    hipLaunchKernelGGL((kCgStep<float, false>), dim3(kStepBlocks), dim3(kStepBlock), 0, c->stream, n, it,
                                       A->x32, A->r32, (const float*)p, dirs[(it + 1) & 1], A->w32, A->dinv32,
                                       dpiPart, wdwPart, npart, zrSet(it), zzSet(it), zrSet(it + 1), zzSet(it + 1), st,
                                       betaSlots, dpiSlots, c->hist, tested);
                continue;
            }
            hipLaunchKernelGGL(kCgDirection<float>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, it, A->z32,
                               (const float*)A->p32, A->p32, zrPart, zzPart, st, betaSlots, c->hist);
            int npart = 0;
            if (elemOp) {
                c->tic("elemop_fp32");
                TRY_CALL(elemOpApply(A, A->p32, A->w32, dpiPart, st, &npart));
            } else {
                c->tic("spmv_fp32");
                TRY_CALL(spmv<float>(c, g, A->s32, A->p32, A->w32, dpiPart, st, &npart));
            }
            c->toc();
            hipLaunchKernelGGL(kCgUpdate<float>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, n, it, A->x32,
                               A->r32, A->z32, A->p32, A->w32, A->dinv32, dpiPart, npart, zrPart, zzPart, st,
                               betaSlots, dpiSlots, c->hist);
        }
        FEEL_HIP_CHECK(hipGetLastError());
        FEEL_HIP_CHECK(hipMemcpyAsync(c->hstate, st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
        FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
        bool handover = false;
        if (!c->hstate->done && it - h.fresh >= kKappaMinIts && it - h.fresh <= kKappaPollIts &&
            it - lastKappa >= kKappaEvery && it <= o.maxit) {
            lastKappa = it;
            TRY_CALL(tooIllConditioned(handover));
        }
        if (!c->hstate->done && !handover) {
            if (it > o.maxit + 1) return setError(FEEL_HIP_EDEVICE, "CG state machine did not terminate");
            const double est = c->hstate->dp;
            const double rate = it > lastIt && est > 0 && est < lastDp ? std::log(est / lastDp) / (it - lastIt) : 0.0;
            const double left = rate < 0 ? std::log(c->hstate->ttol / est) / rate : 64.0;
            chunk = (int)std::min(64.0, std::max(4.0, std::ceil(left)));
            lastDp = est;
            lastIt = it;
            continue;
        }
        if (!handover) {
            it = c->hstate->its;
            reason = c->hstate->reason;
        }
        h.rnorm0 = c->hstate->rnorm0;   // set by the device at iteration 0
        const double estimate = c->hstate->dp * (scaled ? chainScale : sigma);
        // fold the FP32 correction into the FP64 solution and take the true residual
        hipLaunchKernelGGL(kAxpyF32, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, sigma, A->x32,
                           scaled ? (const double*)A->dsq64 : nullptr, A->xp64);
        FEEL_HIP_CHECK(hipGetLastError());
        TRY_CALL(residual64(A, A->bp64, A->xp64, A->r64, A->w64, dp));
        ++out.updates;
        if (reason < 0) break;   // -3, -4, -8, -9, -10 from the device tests
        if (dp != dp) { reason = -9; break; }
        if (dp <= ttol) { reason = dp < o.atol ? 3 : 2; break; }
        if (dp >= o.dtol * rnorm0) { reason = -4; break; }
        if (it >= o.maxit || out.updates >= o.max_outer + 1) { reason = -3; break; }
        if (!handover) TRY_CALL(tooIllConditioned(handover));
        if (handover) {
            // finish in FP64 Jacobi-CG from the current solution, same absolute tolerance
            TRY_CALL(allocWork(c, &A->r64, n));
            TRY_CALL(allocWork(c, &A->z64, n));
            TRY_CALL(allocWork(c, &A->p64, n));
            TRY_CALL(allocWork(c, &A->w64, n));
            CgRun run;
            TRY_CALL(runCg<double>(A, A->s64, A->dinv64, A->bp64, A->xp64, false, 0.0, ttol, o.dtol * rnorm0 / dp,
                                   o.maxit - it, o.check_every, A->r64, A->z64, A->p64, A->w64, "spmv_fp64", run));
            out.fp64Its = run.its;
            it += run.its;
            dp = run.dp;
            reason = run.reason > 0 ? (dp < o.atol ? 3 : 2) : run.reason;
            break;
        }
        TRY_CALL(reset(dp, dp > kDriftRestart * estimate));
        needDirection = true;
        lastDp = scaled ? chainStart : dp / sigma;
        lastIt = it;
        chunk = 16;
    }
    out.its = it;
    out.reason = reason;
    out.dp = dp;
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
    o->inner_rtol = 1e-2;
    o->inner_maxit = 100000;
    o->max_outer = 50;
    o->check_every = 0;
    o->inner_operator = FEEL_HIP_INNER_AUTO;
}

int feel_hip_graph_spmv_shapes(void) { return kNumSpmvShapes; }

int feel_hip_graph_set_spmv_shape(feel_hip_graph* g, int precision, int shape) {
    if (!g) return setError(FEEL_HIP_EINVAL, "null argument");
    if (precision != FEEL_HIP_FP64 && precision != FEEL_HIP_FP32) return setError(FEEL_HIP_EINVAL, "bad precision");
    if (shape < -1 || shape >= kNumSpmvShapes) return setError(FEEL_HIP_EINVAL, "bad SpMV launch shape");
    g->spmvShape[precision == FEEL_HIP_FP64 ? 0 : 1] = shape;
    return FEEL_HIP_OK;
}

int feel_hip_graph_spmv_shape(const feel_hip_graph* g, int precision, int* shape, int* block, int* grid) {
    if (!g || !shape) return setError(FEEL_HIP_EINVAL, "null argument");
    if (precision != FEEL_HIP_FP64 && precision != FEEL_HIP_FP32) return setError(FEEL_HIP_EINVAL, "bad precision");
    *shape = g->spmvShape[precision == FEEL_HIP_FP64 ? 0 : 1];
    if (block) *block = *shape >= 0 ? kSpmvShapes[*shape].block : 0;
    if (grid) *grid = *shape >= 0 ? kSpmvShapes[*shape].grid : 0;
    return FEEL_HIP_OK;
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
        out->fp64_its = run.its;
        out->converged = run.reason > 0;
        return FEEL_HIP_OK;
    }
    if (o.precision != FEEL_HIP_FP32) return setError(FEEL_HIP_EINVAL, "solve precision must be FP64 or FP32");
    // FP32 Jacobi-CG with FP64 reliable updates: FP64 solution, residual and stopping test.
    if (!A->v64) return setError(FEEL_HIP_EINVAL, "mixed-precision solve needs FP64 matrix values for the residual");
    TRY_CALL(allocWork(c, &A->r64, n));
    TRY_CALL(allocWork(c, &A->w64, n));
    TRY_CALL(allocWork(c, &A->r32, n));
    TRY_CALL(allocWork(c, &A->z32, n));
    TRY_CALL(allocWork(c, &A->p32, n));
    TRY_CALL(allocWork(c, &A->w32, n));
    TRY_CALL(allocWork(c, &A->x32, n));
    TRY_CALL(allocWork(c, &A->bp64, n));
    TRY_CALL(allocWork(c, &A->xp64, n));
    if (o.inner_operator < FEEL_HIP_INNER_AUTO || o.inner_operator > FEEL_HIP_INNER_ELEMENT)
        return setError(FEEL_HIP_EINVAL, "unknown inner operator");
    if (o.inner_operator == FEEL_HIP_INNER_ELEMENT) TRY_CALL(elemOpPrepare(A));
    // AUTO takes the assembled FP32 values: on a P2 Poisson matrix the element
    // operator moves about as many bytes per product (its element results are
    // written, then gathered by row) and takes longer
    const bool elemOp = o.inner_operator == FEEL_HIP_INNER_ELEMENT;
    // the element operator replaces the FP32 SELL values of the inner iterations
    TRY_CALL(ensureSell(A, true, !elemOp));
    hipLaunchKernelGGL(kPermuteIn, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, b->d, A->bp64);
    hipLaunchKernelGGL(kPermuteIn, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, x->d, A->xp64);
    MixedRun run;
    TRY_CALL(runMixedCg(A, o, elemOp, run));
    hipLaunchKernelGGL(kPermuteOut, dim3(gridN(n)), dim3(kBlock), 0, c->stream, n, g->perm, A->xp64, x->d);
    FEEL_HIP_CHECK(hipGetLastError());
    c->collectTimers();
    out->its = run.its;
    out->reason = run.reason;
    out->residual = run.dp;
    out->rnorm0 = run.rnorm0;
    out->outer_its = run.updates;
    out->fp64_its = run.fp64Its;
    out->converged = run.reason > 0;
    return FEEL_HIP_OK;
}

}  // extern "C"
