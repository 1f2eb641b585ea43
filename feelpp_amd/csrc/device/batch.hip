// Batched and continuously batched assemble-and-solve of affine-parameter
// Poisson families: A(mu) = sum_t mu_t A_t over resident term matrices
// (the affine decomposition Feel++'s reduced-basis layer assembles once,
// mor/mor/feel/feelmor/modelcrbbase.hpp:1244), homogeneous Dirichlet rows,
// Jacobi-CG in FP32 with FP64 reliable updates (or FP64 throughout).  Every in-flight instance advances in the same kernel launches:
// launches run over a work list of (instance, slice range) items, each
// instance keeps its own CG state and its own iteration origin, so instances
// admitted at different times share launches (continuous batching).
#include <algorithm>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
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

constexpr int kItemSlices = 64;   // slices (64 rows each) per work item
constexpr int kMaxTerms = 8;
// a reliable update whose true residual exceeds the FP32 estimate by more
// than this factor restarts the search direction
constexpr double kBatchDriftRestart = 2.0;

struct BInst {
    int64_t n, nslices;
    const uint32_t* slicePtr;
    const int32_t* col;
    SellCols cols;             // column stream (packed 16/32-bit when compact)
    int compact;
    const int32_t* sdiag;
    const uint8_t* mask;       // per slot: 0 keep, 1 zero, 2 unit diagonal (Dirichlet)
    const double* b64;         // family rhs, solver numbering, Dirichlet rows zero
    const double* terms[kMaxTerms];
    int nterms;
    double mu[kMaxTerms];
    double *val64, *dinv64, *x64, *r64, *w64;
    void *valT, *dinvT, *xT, *rT, *zT, *pT, *wT, *bT;
    CgState* st;
    double* slots;             // beta[2], dpi[2]
    double* part;              // per-instance partials: z.r | z.z | p.Ap (nitems each) | grouped p.Ap
    int nitems;
    int apOff, nap;            // p.Ap partials of the current SpMV: part[apOff .. apOff + nap)
    int start;                 // global iteration at which the current inner solve began
    int pstride;               // p of the FP32 iterations: element r at pT[r * pstride] (8 in a family group)
    double scale;              // FP32 residual = scale * r64 (scale = 1 / sigma, fixed per request)
    double xscale;             // x64 += xscale * xT
};
struct BItem {
    int inst, s0, s1, item;    // item: index of the item within its instance (its partial slot)
};

template <typename T>
__device__ inline T* tp(void* p) { return reinterpret_cast<T*>(p); }

// per-instance sum of item partials, identical in every block of the instance
__device__ inline double instReduce(const double* __restrict__ part, int n) {
    double v = 0;
    for (int i = (int)threadIdx.x; i < n; i += kBlock) v += part[i];
    return blockReduceSum(v);
}

__device__ inline bool instDone(const CgState* st) { return blockDone(st); }

// rows of an item
__device__ inline void itemRows(const BInst& I, const BItem& it, int64_t& r0, int64_t& r1) {
    r0 = (int64_t)it.s0 * 64;
    r1 = std::min<int64_t>((int64_t)it.s1 * 64, I.n);
}

// A(mu) on the SELL slots of the item, FP64 and T copies
template <typename T>
__global__ __launch_bounds__(kBlock) void kBCombine(const BItem* __restrict__ items, const BInst* __restrict__ insts) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    const uint32_t a = I.slicePtr[it.s0], b = I.slicePtr[it.s1];
    for (uint32_t s = a + threadIdx.x; s < b; s += kBlock) {
        const uint8_t m = I.mask[s];
        double v = 0.0;
        if (m == 0)
            for (int t = 0; t < I.nterms; ++t) v += I.mu[t] * I.terms[t][s];
        else if (m == 2)
            v = 1.0;
        I.val64[s] = v;
        if (I.valT) tp<T>(I.valT)[s] = (T)v;   // the grouped engine reads the terms instead
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kBDiag(const BItem* __restrict__ items, const BInst* __restrict__ insts) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    int64_t r0, r1;
    itemRows(I, it, r0, r1);
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
        const double d = I.val64[I.sdiag[r]];
        const double inv = d != 0.0 ? 1.0 / d : 1.0;
        I.dinv64[r] = inv;
        const double dT = (double)(T)d;   // the diagonal of the T copy of A(mu)
        tp<T>(I.dinvT)[r] = (T)(dT != 0.0 ? 1.0 / dT : 1.0);
    }
}

// x64 += xscale * xT (skipped when xscale == 0)
template <typename T>
__global__ __launch_bounds__(kBlock) void kBAxpy(const BItem* __restrict__ items, const BInst* __restrict__ insts) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    if (I.xscale == 0.0) return;
    int64_t r0, r1;
    itemRows(I, it, r0, r1);
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) I.x64[r] += I.xscale * (double)tp<T>(I.xT)[r];
}

// w64 = A64 x64 on the item's slices (one wavefront per slice, the SELL row
// kernel of the single solver)
__global__ __launch_bounds__(kBlock) void kBSpmv64(const BItem* __restrict__ items, const BInst* __restrict__ insts) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    const int lane = threadIdx.x & 63;
    for (int s = it.s0 + (int)(threadIdx.x >> 6); s < it.s1; s += kBlock / 64) {
        const double acc = I.compact ? sellRow<double, 2, true>(I.cols, s, lane, I.val64, I.x64)
                                     : sellRow<double, 2, false>(I.cols, s, lane, I.val64, I.x64);
        const int64_t r = (int64_t)s * 64 + lane;
        if (r < I.n) I.w64[r] = acc;
    }
}

// r64 = b - w64; partial ||D^{-1} r64||^2 per item
__global__ __launch_bounds__(kBlock) void kBResidual(const BItem* __restrict__ items, const BInst* __restrict__ insts) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    double* zzPart = I.part + I.nitems;
    int64_t r0, r1;
    itemRows(I, it, r0, r1);
    double zz = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
        const double ri = I.b64[r] - I.w64[r];
        I.r64[r] = ri;
        const double zi = I.dinv64[r] * ri;
        zz += zi * zi;
    }
    zz = blockReduceSum(zz);
    if (threadIdx.x == 0) zzPart[it.item] = zz;
}

// one block per instance of the work list: out[i] = sum of its z.z partials
__global__ __launch_bounds__(kBlock) void kBInstSum(const BInst* __restrict__ insts, double* __restrict__ out) {
    const BInst& I = insts[blockIdx.x];
    const double v = instReduce(I.part + I.nitems, I.nitems);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// Dirichlet mask on the SELL slots (solver numbering): boundary row -> unit diagonal,
// boundary column -> zero (MatZeroRowsColumns), padding -> zero
__global__ __launch_bounds__(kBlock) void kBMask(int64_t n, int64_t nslices, const uint32_t* __restrict__ slicePtr,
                                                const int32_t* __restrict__ col, const int32_t* __restrict__ slot2csr,
                                                const int32_t* __restrict__ perm, const uint8_t* __restrict__ flag,
                                                uint8_t* __restrict__ mask) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t s = wave; s < nslices; s += nwaves) {
        const uint32_t base = slicePtr[s];
        const int width = (int)((slicePtr[s + 1] - base) >> 6);
        const int64_t r = s * 64 + lane;
        for (int k = 0; k < width; ++k) {
            const size_t slot = sellSlot(base, k, lane);
            uint8_t m = 1;
            if (r < n && slot2csr[slot] >= 0) {
                const int32_t c = col[slot];
                if (flag[perm[r]]) m = (c == r) ? 2 : 1;
                else m = flag[perm[c]] ? 1 : 0;
            }
            mask[slot] = m;
        }
    }
}

// Per slice, the terms with a nonzero FP32 value on a kept slot (mask 0): bit t
// of tmask[s].  A term that is zero over a slice adds only zeros to every
// a = sum_t mu_t a_t there, so the grouped SpMV does not read its plane.
template <typename T>
struct TermPtrs {
    const T* t[kMaxTerms];
};
template <typename V>
__global__ __launch_bounds__(kBlock) void kBTermMask(int64_t nslices, int nterms, const uint32_t* __restrict__ slicePtr,
                                                    TermPtrs<V> T, const uint8_t* __restrict__ mask,
                                                    uint8_t* __restrict__ tmask) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t s = wave; s < nslices; s += nwaves) {
        const uint32_t b0 = slicePtr[s], b1 = slicePtr[s + 1];
        unsigned m = 0;
        for (int t = 0; t < nterms; ++t) {
            bool nz = false;
            for (uint32_t q = b0 + lane; q < b1; q += 64) nz |= mask[q] == 0 && T.t[t][q] != (V)0;
            if (__ballot(nz)) m |= 1u << t;
        }
        if (lane == 0) tmask[s] = (uint8_t)m;
    }
}

// permuted rhs with zero Dirichlet rows
__global__ __launch_bounds__(kBlock) void kBRhs(int64_t n, const int32_t* __restrict__ perm, const uint8_t* __restrict__ flag,
                                               const double* __restrict__ b, double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = flag[perm[i]] ? 0.0 : b[perm[i]];
}

// inner solve set-up: b = scale * r64, x = 0, r = b, z = D^{-1} r; partials z.r, z.z
template <typename T>
__global__ __launch_bounds__(kBlock) void kBInnerInit(const BItem* __restrict__ items, const BInst* __restrict__ insts) {
    const BItem it = items[blockIdx.x];
    const BInst& I = insts[it.inst];
    int64_t r0, r1;
    itemRows(I, it, r0, r1);
    double zr = 0.0, zz = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
        const T bi = (T)(I.scale * I.r64[r]);
        tp<T>(I.bT)[r] = bi;
        tp<T>(I.xT)[r] = (T)0;
        tp<T>(I.rT)[r] = bi;
        const T zi = tp<T>(I.dinvT)[r] * bi;
        tp<T>(I.zT)[r] = zi;
        zr += (double)zi * (double)bi;
        zz += (double)zi * (double)zi;
    }
    zr = blockReduceSum(zr);
    zz = blockReduceSum(zz);
    if (threadIdx.x == 0) { I.part[it.item] = zr; I.part[I.nitems + it.item] = zz; }
}

// CG direction with the convergence test (as kCgDirection, per instance); the
// direction (re)starts as p = z at it == st->fresh
template <typename T>
__global__ __launch_bounds__(kBlock) void kBDirection(const BItem* __restrict__ items, const BInst* __restrict__ insts,
                                                     const int* __restrict__ gcount, int k) {
    const int bid = blockIdx.x;
    const BItem item = items[bid];
    const BInst& I = insts[item.inst];
    CgState* st = I.st;
    if (instDone(st)) return;
    const int it = *gcount + k - I.start;
    const double beta = instReduce(I.part, I.nitems);
    const double dp = sqrt(instReduce(I.part + I.nitems, I.nitems));
    const double ttol = it == 0 ? fmax(st->rtol * dp, st->atol) : st->ttol;
    const double rnorm0 = it == 0 ? dp : st->rnorm0;
    int reason = 0;
    if (dp != dp) reason = -9;
    else if (dp <= ttol) reason = dp < st->atol ? 3 : 2;
    else if (it > 0 && dp >= st->dtol * rnorm0) reason = -4;
    else if (it >= st->maxit) reason = -3;
    else if (beta == 0.0) reason = 3;
    else if (it > 0 && beta * I.slots[(it - 1) & 1] < 0.0) reason = -8;   // indefinite preconditioner
    if (item.item == 0 && threadIdx.x == 0) {
        st->dp = dp;
        if (it == 0) { st->ttol = ttol; st->rnorm0 = rnorm0; }
        if (reason) { st->done = 1; st->reason = reason; st->its = reason == -8 ? it + 1 : it; }
        I.slots[it & 1] = beta;
    }
    if (reason) return;
    const bool first = it == st->fresh;
    const double bb = first ? 0.0 : beta / I.slots[(it - 1) & 1];
    int64_t r0, r1;
    itemRows(I, item, r0, r1);
    T* p = tp<T>(I.pT);
    const T* z = tp<T>(I.zT);
    const int ps = I.pstride;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock)
        p[r * ps] = first ? z[r] : (T)((double)z[r] + bb * (double)p[r * ps]);
}

// w = A p on the item's slices (one wavefront per slice, the SELL row kernel
// of the single solver), partial p.Ap
template <typename T>
__global__ __launch_bounds__(kBlock) void kBSpmv(const BItem* __restrict__ items, const BInst* __restrict__ insts) {
    constexpr int U = 2;   // quads per step
    const BItem item = items[blockIdx.x];
    const BInst& I = insts[item.inst];
    if (instDone(I.st)) return;
    const int lane = threadIdx.x & 63;
    const T* val = tp<T>(I.valT);
    const T* x = tp<T>(I.pT);
    T* y = tp<T>(I.wT);
    const SellCols S = I.cols;
    double dot = 0.0;
    for (int s = item.s0 + (int)(threadIdx.x >> 6); s < item.s1; s += kBlock / 64) {
        const T acc = I.compact ? sellRow<T, U, true>(S, s, lane, val, x) : sellRow<T, U, false>(S, s, lane, val, x);
        const int64_t r = (int64_t)s * 64 + lane;
        if (r < I.n) {
            y[r] = acc;
            dot += (double)x[r] * (double)acc;
        }
    }
    dot = blockReduceSum(dot);
    if (threadIdx.x == 0) I.part[2 * I.nitems + item.item] = dot;
}

constexpr int kGroupMax = 8;
constexpr int kGroupBlock = 256;   // 4 wavefronts
constexpr int kXcds = 8;           // accelerator complex dies, each with its own L2
constexpr int kGroupSlices = 16;   // slices per grouped-SpMV block (4 per wavefront)
// A family group: up to kGroupMax in-flight instances of one family (all FP32
// or all FP64) whose search directions are stored interleaved, P[r * 8 +
// slot], so that one 32-byte (FP32) or 64-byte (FP64) row brings every
// instance's p at a column.
struct BGroup {
    int nterms, compact, nitems;
    int64_t n;
    const void* P;                       // interleaved search directions of the group (T)
    int inst[kGroupMax];                 // per slot: position in the launch's instance array, -1 if none
    float muF[kGroupMax][kMaxTerms];     // mu rounded to FP32 (FP32 groups)
    double muD[kGroupMax][kMaxTerms];    // mu (FP64 groups)
    const void* terms[kMaxTerms];        // family terms, SELL values in T
    const uint8_t* mask;
    const uint8_t* tmask;                // per slice: terms nonzero there (bit t)
    SellCols cols;
};
struct BGItem {
    int group, s0, s1, item;   // item: index of the block's p.Ap partial (s0 / kGroupSlices)
};

typedef Quad<float>::type f32x4;
typedef Quad<double>::type f64x4;

template <typename T>
__device__ inline T groupMu(const BGroup& G, int b, int t);
template <>
__device__ inline float groupMu<float>(const BGroup& G, int b, int t) { return G.muF[b][t]; }
template <>
__device__ inline double groupMu<double>(const BGroup& G, int b, int t) { return G.muD[b][t]; }

// four consecutive T values: FP32 non-temporally (one 16-byte load; the term
// stream is read once), FP64 through L2 (two 16-byte loads of one line)
__device__ inline void load4(const float* p, float* v) {
    const f32x4 q = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}
__device__ inline void load4(const double* p, double* v) {
    const f64x4 q = *reinterpret_cast<const f64x4*>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
}

// One quad (4 consecutive entries) of one slice row: columns, unit diagonal
// of a Dirichlet row, and the term values (zero on Dirichlet slots)
template <typename T, int NT>
struct GroupQuad {
    int32_t col[4];
    T e[4];
    T a[4][NT];
};

template <typename T, int NT, bool NARROW>
__device__ inline void loadQuad(const BGroup& G, unsigned tm, size_t q0, const void* cp, int32_t cb, int kq,
                                GroupQuad<T, NT>& d) {
    // q0: slot of the lane's first quad; quad kq starts 256 kq slots further
    const size_t q = q0 + ((size_t)kq << 8);
    if (NARROW) {
        const u16x4 c = static_cast<const u16x4*>(cp)[(size_t)kq * 64];
        d.col[0] = cb + c.x; d.col[1] = cb + c.y; d.col[2] = cb + c.z; d.col[3] = cb + c.w;
    } else {
        const i32x4 c = static_cast<const i32x4*>(cp)[(size_t)kq * 64];
        d.col[0] = c.x; d.col[1] = c.y; d.col[2] = c.z; d.col[3] = c.w;
    }
    const u8x4 m = *reinterpret_cast<const u8x4*>(G.mask + q);
    const uint8_t mm[4] = {m.x, m.y, m.z, m.w};
    T tv[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        // tm is wave-uniform: a term that is zero over the slice is not read
        if ((tm >> t) & 1u) load4(static_cast<const T*>(G.terms[t]) + q, tv[t]);
        else tv[t][0] = tv[t][1] = tv[t][2] = tv[t][3] = (T)0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        d.e[u] = mm[u] == 2 ? (T)1 : (T)0;
#pragma unroll
        for (int t = 0; t < NT; ++t) d.a[u][t] = mm[u] == 0 ? tv[t][u] : (T)0;
    }
}

// the 8 interleaved p values of the group at row (or column) r
__device__ inline void groupP(const float* __restrict__ P, int64_t r, float* pv) {
    const f32x4* q = reinterpret_cast<const f32x4*>(P) + 2 * r;
    const f32x4 lo = q[0], hi = q[1];
    pv[0] = lo.x; pv[1] = lo.y; pv[2] = lo.z; pv[3] = lo.w;
    pv[4] = hi.x; pv[5] = hi.y; pv[6] = hi.z; pv[7] = hi.w;
}
__device__ inline void groupP(const double* __restrict__ P, int64_t r, double* pv) {
    const f64x4* q = reinterpret_cast<const f64x4*>(P) + 2 * r;
    const f64x4 lo = q[0], hi = q[1];
    pv[0] = lo.x; pv[1] = lo.y; pv[2] = lo.z; pv[3] = lo.w;
    pv[4] = hi.x; pv[5] = hi.y; pv[6] = hi.z; pv[7] = hi.w;
}
__device__ inline void groupPStore(float* __restrict__ P, int64_t r, const float* pv) {
    f32x4* q = reinterpret_cast<f32x4*>(P) + 2 * r;
    q[0] = f32x4{pv[0], pv[1], pv[2], pv[3]};
    q[1] = f32x4{pv[4], pv[5], pv[6], pv[7]};
}
__device__ inline void groupPStore(double* __restrict__ P, int64_t r, const double* pv) {
    f64x4* q = reinterpret_cast<f64x4*>(P) + 2 * r;
    q[0] = f64x4{pv[0], pv[1], pv[2], pv[3]};
    q[1] = f64x4{pv[4], pv[5], pv[6], pv[7]};
}

// acc_b += A(mu_b) p_b on the quad for every active instance b of the group
template <typename T, int NT>
__device__ inline void applyQuad(const BGroup& G, unsigned amask, unsigned tm, const GroupQuad<T, NT>& d, T* acc) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        T pv[kGroupMax];
        groupP(static_cast<const T*>(G.P), d.col[u], pv);
        T v[kGroupMax];
#pragma unroll
        for (int b = 0; b < kGroupMax; ++b) v[b] = d.e[u];
        // terms in order; one that is zero over the slice (tm, wave-uniform)
        // would only add zeros and is skipped
#pragma unroll
        for (int t = 0; t < NT; ++t)
            if ((tm >> t) & 1u)
#pragma unroll
                for (int b = 0; b < kGroupMax; ++b)
                    if ((amask >> b) & 1u) v[b] += groupMu<T>(G, b, t) * d.a[u][t];
#pragma unroll
        for (int b = 0; b < kGroupMax; ++b)
            if ((amask >> b) & 1u) acc[b] += v[b] * pv[b];
    }
}

// one slice row, the next quad's matrix data loaded while the current quad's
// gathers and products run
template <typename T, int NT, bool NARROW>
__device__ inline void groupRow(const BGroup& G, unsigned amask, unsigned tm, size_t q0, const void* cp, int32_t cb,
                                int nq, T* acc) {
    if (nq <= 0) return;
    GroupQuad<T, NT> cur;
    loadQuad<T, NT, NARROW>(G, tm, q0, cp, cb, 0, cur);
    for (int k = 1; k < nq; ++k) {
        GroupQuad<T, NT> nxt;
        loadQuad<T, NT, NARROW>(G, tm, q0, cp, cb, k, nxt);
        applyQuad<T, NT>(G, amask, tm, cur, acc);
        cur = nxt;
    }
    applyQuad<T, NT>(G, amask, tm, cur, acc);
}

// Grouped SpMV of an affine family (the CG iterations in T): the in-flight
// instances of one family share the sparsity pattern and the term matrices, so
// a group streams each slot's term values, Dirichlet mask and column once and
// every instance b combines a = sum_t mu_bt a_t on the fly (mask 1: zero,
// mask 2: unit diagonal) before w_b += a * p_b[col], the p of all instances at
// a column coming from one row of the interleaved directions.  Each instance's
// arithmetic is independent of the others in the group, so a result is the
// same alone and in any mix.
template <typename T, int NT>
__global__ __launch_bounds__(kGroupBlock) void kBSpmvGroup(const BGItem* __restrict__ gitems,
                                                          const BGroup* __restrict__ groups,
                                                          const BInst* __restrict__ insts) {
    const BGItem it = gitems[blockIdx.x];
    if (it.group < 0) return;
    const BGroup& G = groups[it.group];
    __shared__ unsigned actMask;
    __shared__ double red[kGroupBlock / 64][kGroupMax];
    if (threadIdx.x < 64) {
        const int b = threadIdx.x;
        const int ii = b < kGroupMax ? G.inst[b] : -1;
        const bool on = ii >= 0 && !insts[ii].st->done;
        const unsigned long long bal = __ballot(on);
        if (threadIdx.x == 0) actMask = (unsigned)bal;
    }
    __syncthreads();
    const unsigned amask = __builtin_amdgcn_readfirstlane(actMask);
    if (!amask) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double dot[kGroupMax];
#pragma unroll
    for (int b = 0; b < kGroupMax; ++b) dot[b] = 0.0;
    for (int sl = it.s0 + wv; sl < it.s1; sl += kGroupBlock / 64) {
        const uint32_t base = G.cols.slicePtr[sl];
        const int nq = (int)((G.cols.slicePtr[sl + 1] - base) >> 8);
        const size_t q0 = (size_t)base + ((size_t)lane << 2);
        T acc[kGroupMax];
#pragma unroll
        for (int b = 0; b < kGroupMax; ++b) acc[b] = (T)0;
        const int32_t cb = G.compact ? G.cols.colBase[sl] : -1;
        const unsigned tm = __builtin_amdgcn_readfirstlane((unsigned)G.tmask[sl]);
        if (cb >= 0)
            groupRow<T, NT, true>(G, amask, tm, q0, reinterpret_cast<const u16x4*>(G.cols.packed + G.cols.cptr[sl]) + lane, cb,
                               nq, acc);
        else if (G.compact)
            groupRow<T, NT, false>(G, amask, tm, q0, reinterpret_cast<const i32x4*>(G.cols.packed + G.cols.cptr[sl]) + lane,
                                cb, nq, acc);
        else
            groupRow<T, NT, false>(G, amask, tm, q0, reinterpret_cast<const i32x4*>(G.cols.col + base) + lane, cb, nq, acc);
        const int64_t r = (int64_t)sl * 64 + lane;
        if (r < G.n) {
            T pr[kGroupMax];
            groupP(static_cast<const T*>(G.P), r, pr);
#pragma unroll
            for (int b = 0; b < kGroupMax; ++b)
                if ((amask >> b) & 1u) {
                    tp<T>(insts[G.inst[b]].wT)[r] = acc[b];
                    dot[b] += (double)pr[b] * (double)acc[b];
                }
        }
    }
    // per-instance p.Ap partial of this item: wave sums, then the waves in order
#pragma unroll
    for (int b = 0; b < kGroupMax; ++b) {
        if ((amask >> b) & 1u) {
            const double v = waveReduceSum(dot[b]);
            if (lane == 0) red[wv][b] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < kGroupMax && ((amask >> threadIdx.x) & 1u)) {
        const int b = threadIdx.x;
        double v = 0.0;
        for (int w = 0; w < kGroupBlock / 64; ++w) v += red[w][b];
        const BInst& I = insts[G.inst[b]];
        I.part[3 * I.nitems + it.item] = v;
    }
}

// Direction and update kernels of a family group: one block covers one item
// (64 slices) of every instance of the group, so the interleaved P is read
// and written in whole rows.  Per instance the operations, their order and the
// rows each thread handles are those of kBDirection and kBUpdate, so the
// results are the same as theirs, bit for bit.
template <typename T>
__global__ __launch_bounds__(kBlock) void kBDirectionGroup(const BGItem* __restrict__ gitems,
                                                          const BGroup* __restrict__ groups,
                                                          const BInst* __restrict__ insts, const int* __restrict__ gcount, int k) {
    const BGItem item = gitems[blockIdx.x];
    if (item.group < 0) return;
    const BGroup& G = groups[item.group];
    __shared__ int sdone[kGroupMax];
    __shared__ int smode[kGroupMax];      // 0 skip, 1 p = z + bb p, 2 p = z
    __shared__ double sbb[kGroupMax];
    if (threadIdx.x < kGroupMax) {
        const int ii = G.inst[threadIdx.x];
        sdone[threadIdx.x] = ii < 0 ? 1 : insts[ii].st->done;
        smode[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int b = 0; b < kGroupMax; ++b) {
        if (sdone[b]) continue;
        const BInst& I = insts[G.inst[b]];
        CgState* st = I.st;
        const int it = *gcount + k - I.start;
        const double beta = instReduce(I.part, I.nitems);
        const double dp = sqrt(instReduce(I.part + I.nitems, I.nitems));
        const double ttol = it == 0 ? fmax(st->rtol * dp, st->atol) : st->ttol;
        const double rnorm0 = it == 0 ? dp : st->rnorm0;
        int reason = 0;
        if (dp != dp) reason = -9;
        else if (dp <= ttol) reason = dp < st->atol ? 3 : 2;
        else if (it > 0 && dp >= st->dtol * rnorm0) reason = -4;
        else if (it >= st->maxit) reason = -3;
        else if (beta == 0.0) reason = 3;
        else if (it > 0 && beta * I.slots[(it - 1) & 1] < 0.0) reason = -8;   // indefinite preconditioner
        if (item.item == 0 && threadIdx.x == 0) {
            st->dp = dp;
            if (it == 0) { st->ttol = ttol; st->rnorm0 = rnorm0; }
            if (reason) { st->done = 1; st->reason = reason; st->its = reason == -8 ? it + 1 : it; }
            I.slots[it & 1] = beta;
        }
        if (!reason && threadIdx.x == 0) {
            const bool first = it == st->fresh;
            smode[b] = first ? 2 : 1;
            sbb[b] = first ? 0.0 : beta / I.slots[(it - 1) & 1];
        }
    }
    __syncthreads();
    const T* z[kGroupMax];
    int mode[kGroupMax];
    double bb[kGroupMax];
#pragma unroll
    for (int b = 0; b < kGroupMax; ++b) {
        mode[b] = smode[b];
        bb[b] = sbb[b];
        z[b] = mode[b] ? tp<T>(insts[G.inst[b]].zT) : nullptr;
    }
    const int64_t r0 = (int64_t)item.s0 * 64, r1 = std::min<int64_t>((int64_t)item.s1 * 64, G.n);
    T* P = static_cast<T*>(const_cast<void*>(G.P));
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kBlock) {
        // every load of the row before any store: the instances' loads overlap
        T zv[kGroupMax];
#pragma unroll
        for (int b = 0; b < kGroupMax; ++b) zv[b] = mode[b] ? z[b][r] : (T)0;
        T pv[kGroupMax];
        groupP(P, r, pv);
#pragma unroll
        for (int b = 0; b < kGroupMax; ++b)
            if (mode[b]) pv[b] = mode[b] == 2 ? zv[b] : (T)((double)zv[b] + bb[b] * (double)pv[b]);
        groupPStore(P, r, pv);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kBUpdateGroup(const BGItem* __restrict__ gitems,
                                                       const BGroup* __restrict__ groups,
                                                       const BInst* __restrict__ insts, const int* __restrict__ gcount, int k) {
    const BGItem item = gitems[blockIdx.x];
    if (item.group < 0) return;
    const BGroup& G = groups[item.group];
    __shared__ int sdone[kGroupMax];
    __shared__ int sact[kGroupMax];
    __shared__ double sa[kGroupMax];
    if (threadIdx.x < kGroupMax) {
        const int ii = G.inst[threadIdx.x];
        sdone[threadIdx.x] = ii < 0 ? 1 : insts[ii].st->done;
        sact[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int b = 0; b < kGroupMax; ++b) {
        if (sdone[b]) continue;
        const BInst& I = insts[G.inst[b]];
        CgState* st = I.st;
        const int it = *gcount + k - I.start;
        const double dpi = instReduce(I.part + I.apOff, I.nap);
        const bool indefinite = dpi == 0.0 || (it > 0 && dpi * I.slots[2 + ((it - 1) & 1)] <= 0.0);
        if (item.item == 0 && threadIdx.x == 0) {
            I.slots[2 + (it & 1)] = dpi;
            if (indefinite) { st->done = 1; st->reason = -10; st->its = it + 1; }
        }
        if (!indefinite && threadIdx.x == 0) {
            sact[b] = 1;
            sa[b] = I.slots[it & 1] / dpi;
        }
    }
    __syncthreads();
    // active set as a wave-uniform mask, so the per-instance pointers stay scalar
    unsigned m = 0;
    for (int b = 0; b < kGroupMax; ++b) m |= (unsigned)(sact[b] != 0) << b;
    const unsigned amask = __builtin_amdgcn_readfirstlane(m);
    int act[kGroupMax];
    double a[kGroupMax], zr[kGroupMax], zz[kGroupMax];
#pragma unroll
    for (int b = 0; b < kGroupMax; ++b) {
        act[b] = (amask >> b) & 1u;
        a[b] = sa[b];
        zr[b] = 0.0;
        zz[b] = 0.0;
    }
    const int64_t r0 = (int64_t)item.s0 * 64, r1 = std::min<int64_t>((int64_t)item.s1 * 64, G.n);
    const T* P = static_cast<const T*>(G.P);
    T *X[kGroupMax], *Rv[kGroupMax], *Z[kGroupMax];
    const T *W[kGroupMax], *D[kGroupMax];
#pragma unroll
    for (int b = 0; b < kGroupMax; ++b) {
        const BInst* I = act[b] ? &insts[G.inst[b]] : nullptr;
        X[b] = I ? tp<T>(I->xT) : nullptr;
        Rv[b] = I ? tp<T>(I->rT) : nullptr;
        Z[b] = I ? tp<T>(I->zT) : nullptr;
        W[b] = I ? tp<T>(I->wT) : nullptr;
        D[b] = I ? tp<T>(I->dinvT) : nullptr;
    }
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
        T pv[kGroupMax];
        groupP(P, i, pv);
        // every load of the row before any store: the instances' loads overlap
        T xv[kGroupMax], rv[kGroupMax], wv[kGroupMax], dv[kGroupMax];
#pragma unroll
        for (int b = 0; b < kGroupMax; ++b) {
            if (!act[b]) continue;
            xv[b] = X[b][i];
            rv[b] = Rv[b][i];
            wv[b] = W[b][i];
            dv[b] = D[b][i];
        }
#pragma unroll
        for (int b = 0; b < kGroupMax; ++b) {
            if (!act[b]) continue;
            X[b][i] = (T)((double)xv[b] + a[b] * (double)pv[b]);
            const T ri = (T)((double)rv[b] - a[b] * (double)wv[b]);
            Rv[b][i] = ri;
            const T zi = dv[b] * ri;
            Z[b][i] = zi;
            zr[b] += (double)zi * (double)ri;
            zz[b] += (double)zi * (double)zi;
        }
    }
    for (int b = 0; b < kGroupMax; ++b) {
        if (!act[b]) continue;
        const double vzr = blockReduceSum(zr[b]);
        const double vzz = blockReduceSum(zz[b]);
        const BInst& I = insts[G.inst[b]];
        if (threadIdx.x == 0) { I.part[item.item] = vzr; I.part[I.nitems + item.item] = vzz; }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kBUpdate(const BItem* __restrict__ items, const BInst* __restrict__ insts,
                                                  const int* __restrict__ gcount, int k) {
    const int bid = blockIdx.x;
    const BItem item = items[bid];
    const BInst& I = insts[item.inst];
    CgState* st = I.st;
    if (instDone(st)) return;
    const int it = *gcount + k - I.start;
    const double dpi = instReduce(I.part + I.apOff, I.nap);
    const bool indefinite = dpi == 0.0 || (it > 0 && dpi * I.slots[2 + ((it - 1) & 1)] <= 0.0);
    if (item.item == 0 && threadIdx.x == 0) {
        I.slots[2 + (it & 1)] = dpi;
        if (indefinite) { st->done = 1; st->reason = -10; st->its = it + 1; }
    }
    if (indefinite) return;
    const double a = I.slots[it & 1] / dpi;
    int64_t r0, r1;
    itemRows(I, item, r0, r1);
    T *x = tp<T>(I.xT), *r = tp<T>(I.rT), *z = tp<T>(I.zT);
    const T *p = tp<T>(I.pT), *w = tp<T>(I.wT), *dinv = tp<T>(I.dinvT);
    const int ps = I.pstride;
    double zr = 0.0, zz = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
        x[i] = (T)((double)x[i] + a * (double)p[i * ps]);
        const T ri = (T)((double)r[i] - a * (double)w[i]);
        r[i] = ri;
        const T zi = dinv[i] * ri;
        z[i] = zi;
        zr += (double)zi * (double)ri;
        zz += (double)zi * (double)zi;
    }
    zr = blockReduceSum(zr);
    zz = blockReduceSum(zz);
    if (threadIdx.x == 0) { I.part[item.item] = zr; I.part[I.nitems + item.item] = zz; }
}

__global__ void kBBump(int* gcount, int iterations) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *gcount += iterations;
}

__global__ __launch_bounds__(kBlock) void kBPermuteOut(int64_t n, const int32_t* __restrict__ perm,
                                                      const double* __restrict__ in, double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[perm[i]] = in[i];
}

}  // namespace

// ============================================================== host engine
struct BatchFamily {
    feel_hip_graph* g = nullptr;
    int nterms = 0;
    std::vector<double*> terms;   // SELL FP64 values per term
    std::vector<float*> termsF;   // the same rounded to FP32 (grouped SpMV of FP32 batches)
    uint8_t* mask = nullptr;
    uint8_t* tmask = nullptr;     // per slice: FP32 terms nonzero on a kept slot (grouped SpMV)
    double* b64 = nullptr;        // permuted rhs with zero Dirichlet rows
    int64_t n = 0, nslices = 0, nslots = 0;
    int nitems = 0;
    size_t instBytes = 0;
    std::vector<int> freeBufs;    // reusable instance buffer sets
    // family groups of the grouped SpMV: interleaved search directions
    // P[n x 8] and the request holding each slot (-1 = free); an instance keeps
    // its slot from admission until it finishes
    struct PGroup {
        void* P = nullptr;   // T rows of kGroupMax
        int64_t req[kGroupMax];
    };
    std::vector<PGroup> pgroups;
};

struct InstBuffers {
    double *val64 = nullptr, *dinv64 = nullptr, *x64 = nullptr, *r64 = nullptr, *w64 = nullptr, *out = nullptr;
    double* part = nullptr;
    void *valT = nullptr, *dinvT = nullptr, *xT = nullptr, *rT = nullptr, *zT = nullptr, *pT = nullptr, *wT = nullptr,
         *bT = nullptr;
};

enum class Phase { Pending, Resid, Inner, Done, Released };

struct Request {
    int64_t id = 0;
    int family = 0;
    std::vector<double> mu;
    Phase phase = Phase::Pending;
    int buf = -1, state = -1;
    int start = 0;
    double dp = 0, rnorm0 = 0, ttol = 0, xscale = 0;
    double sigma = 0;      // FP32 residual scale (the initial FP64 residual norm)
    double estimate = 0;   // FP32 estimate of the FP64 residual norm when the device stopped
    int its = 0, outer = 0, reason = 0;
    int fresh = 0;         // iteration at which the search direction last (re)started
    double devRnorm0 = 0;  // the device's initial residual norm (FP32 units), for the dtol test
    bool unpermuted = false;   // solution copied to Feel++ numbering
    int pgroup = -1, pslot = -1;   // family group slot of the grouped SpMV
    int waited = 0;        // steps spent at the head of the queue without fitting (backfill)
};

struct feel_hip_batch {
    feel_hip_ctx* ctx = nullptr;
    feel_hip_batch_opts opts{};
    int64_t used = 0;
    std::vector<BatchFamily> fams;
    std::vector<InstBuffers> bufs;
    std::map<int64_t, Request> reqs;
    std::deque<int64_t> pending;
    std::deque<int64_t> finished;   // done, not yet polled
    int64_t nextId = 0;
    CgState* states = nullptr;
    double* slots = nullptr;
    int nStates = 0;
    std::vector<int> freeStates;
    BItem* dItems = nullptr;
    BInst* dInsts = nullptr;
    BGroup* dGroups = nullptr;
    BGItem* dGItems = nullptr;
    int capGroups = 0, capGItems = 0, nGItems = 0;
    BGItem* dGDItems = nullptr;   // group items of the direction / update kernels (64 slices each)
    int capGDItems = 0, nGDItems = 0;
    double* dSums = nullptr;
    int capItems = 0, capInsts = 0;
    int* gcount = nullptr;
    int hgcount = 0;
    hipGraphExec_t exec = nullptr;
    int execItems = -1, execChunk = -1, execGItems = -1, execGDItems = -1, execTerms = -1;
    uint64_t execVersion = ~0ULL, version = 0;
};

namespace {

template <typename T>
int alloc(feel_hip_ctx* c, T** p, size_t count) {
    return c->alloc((void**)p, sizeof(T) * count);
}

size_t tsize(const feel_hip_batch* B) { return B->opts.precision == FEEL_HIP_FP32 ? 4 : 8; }

// z.r, z.z and p.Ap partials per item, then one p.Ap partial per grouped-SpMV block
size_t partCount(const BatchFamily& f) {
    return 3 * (size_t)f.nitems + (size_t)((f.nslices + kGroupSlices - 1) / kGroupSlices);
}

// the grouped engine reads the family terms and keeps p in the family
// group: its instances hold no T copy of A(mu) and no p of their own
bool groupedEngine(const feel_hip_batch* B) { return B->opts.group_spmv != 0; }

size_t instanceBytes(const feel_hip_batch* B, const BatchFamily& f) {
    const size_t t = tsize(B);
    if (groupedEngine(B)) return f.nslots * 8 + f.n * (8 * 5 + t * 6) + sizeof(double) * partCount(f);
    return f.nslots * (8 + t) + f.n * (8 * 5 + t * 7) + sizeof(double) * partCount(f);
}

int allocInstance(feel_hip_batch* B, int fam, int* out) {
    BatchFamily& f = B->fams[fam];
    if (!f.freeBufs.empty()) {
        *out = f.freeBufs.back();
        f.freeBufs.pop_back();
        return FEEL_HIP_OK;
    }
    feel_hip_ctx* c = B->ctx;
    const size_t t = tsize(B);
    InstBuffers ib;
    TRY_CALL(alloc(c, &ib.val64, f.nslots));
    if (!groupedEngine(B)) {
        TRY_CALL(c->alloc(&ib.valT, t * f.nslots));
        TRY_CALL(c->alloc(&ib.pT, t * f.n));
    }
    for (double** p : {&ib.dinv64, &ib.x64, &ib.r64, &ib.w64, &ib.out}) TRY_CALL(alloc(c, p, f.n));
    TRY_CALL(alloc(c, &ib.part, partCount(f)));
    for (void** p : {&ib.dinvT, &ib.xT, &ib.rT, &ib.zT, &ib.wT, &ib.bT}) TRY_CALL(c->alloc(p, t * f.n));
    B->bufs.push_back(ib);
    B->used += (int64_t)instanceBytes(B, f);
    *out = (int)B->bufs.size() - 1;
    return FEEL_HIP_OK;
}

int ensureStates(feel_hip_batch* B, int need) {
    if (need <= B->nStates) return FEEL_HIP_OK;
    const int cap = std::max(need, 2 * B->nStates + 16);
    CgState* ns = nullptr;
    double* nsl = nullptr;
    TRY_CALL(alloc(B->ctx, &ns, cap));
    TRY_CALL(alloc(B->ctx, &nsl, 4 * (size_t)cap));
    if (B->nStates) {
        FEEL_HIP_CHECK(hipMemcpyAsync(ns, B->states, sizeof(CgState) * B->nStates, hipMemcpyDeviceToDevice, B->ctx->stream));
        FEEL_HIP_CHECK(hipMemcpyAsync(nsl, B->slots, sizeof(double) * 4 * B->nStates, hipMemcpyDeviceToDevice, B->ctx->stream));
        FEEL_HIP_CHECK(hipStreamSynchronize(B->ctx->stream));
        B->ctx->release(B->states, sizeof(CgState) * B->nStates);
        B->ctx->release(B->slots, sizeof(double) * 4 * B->nStates);
    }
    for (int s = cap - 1; s >= B->nStates; --s) B->freeStates.push_back(s);
    B->states = ns;
    B->slots = nsl;
    B->nStates = cap;
    ++B->version;
    return FEEL_HIP_OK;
}

// Work list (and per-instance descriptors) for a set of requests.
// Work list of the requests rs: one BInst per request, its 64-slice items, and
// with groups set the family groups of the grouped SpMV (instances of one
// family in request order, kGroupMax per group; group items ordered by item
// within a family so the groups of a family stream the same slices together).
int buildWork(feel_hip_batch* B, const std::vector<Request*>& rs, int* nItems, bool groups = false) {
    std::vector<BItem> items;
    std::vector<BInst> insts;
    std::vector<BGroup> grp;
    std::vector<BGItem> gitems, gditems;
    for (Request* r : rs) {
        const BatchFamily& f = B->fams[r->family];
        const InstBuffers& ib = B->bufs[r->buf];
        BInst I{};
        I.n = f.n;
        I.nslices = f.nslices;
        I.slicePtr = f.g->slicePtr;
        I.col = f.g->scol;
        I.cols = SellCols{f.g->slicePtr, f.g->scol, f.g->colBase, f.g->ccptr, f.g->packed};
        I.compact = f.g->compact ? 1 : 0;
        I.sdiag = f.g->sdiag;
        I.mask = f.mask;
        I.b64 = f.b64;
        I.nterms = f.nterms;
        for (int t = 0; t < f.nterms; ++t) { I.terms[t] = f.terms[t]; I.mu[t] = r->mu[t]; }
        I.val64 = ib.val64; I.dinv64 = ib.dinv64; I.x64 = ib.x64; I.r64 = ib.r64; I.w64 = ib.w64;
        I.valT = ib.valT; I.dinvT = ib.dinvT; I.xT = ib.xT; I.rT = ib.rT; I.zT = ib.zT; I.pT = ib.pT;
        I.wT = ib.wT; I.bT = ib.bT;
        I.pstride = 1;
        if (r->pgroup >= 0) {
            I.pT = static_cast<char*>(f.pgroups[r->pgroup].P) + tsize(B) * r->pslot;
            I.pstride = kGroupMax;
        }
        I.st = B->states + r->state;
        I.slots = B->slots + 4 * (size_t)r->state;
        I.part = ib.part;
        I.nitems = f.nitems;
        I.apOff = groups ? 3 * f.nitems : 2 * f.nitems;
        I.nap = groups ? (int)((f.nslices + kGroupSlices - 1) / kGroupSlices) : f.nitems;
        I.start = r->start;
        I.scale = r->sigma > 0 ? 1.0 / r->sigma : 0.0;
        I.xscale = r->xscale;
        // grouped instances get their items below, interleaved by group
        if (r->pgroup < 0)
            for (int64_t s = 0, j = 0; s < f.nslices; s += kItemSlices, ++j)
                items.push_back({(int)insts.size(), (int)s, (int)std::min<int64_t>(f.nslices, s + kItemSlices), (int)j});
        insts.push_back(I);
    }
    // Items of the instances of one family group: for each slice range, the
    // group's instances side by side, so the strided accesses of the direction
    // and update kernels to the interleaved P meet in L2.
    {
        std::map<std::pair<int, int>, std::vector<int>> byGroup;   // (family, group) -> positions in rs
        for (size_t i = 0; i < rs.size(); ++i)
            if (rs[i]->pgroup >= 0) byGroup[{rs[i]->family, rs[i]->pgroup}].push_back((int)i);
        for (auto& kv : byGroup) {
            const BatchFamily& f = B->fams[kv.first.first];
            for (int64_t s = 0, j = 0; s < f.nslices; s += kItemSlices, ++j)
                for (int i : kv.second)
                    items.push_back({i, (int)s, (int)std::min<int64_t>(f.nslices, s + kItemSlices), (int)j});
        }
    }
    if (groups) {
        // one BGroup per family group with members in this launch
        std::map<int64_t, int> pos;   // request id -> position in rs
        for (size_t i = 0; i < rs.size(); ++i) pos[rs[i]->id] = (int)i;
        std::map<std::pair<int, int>, bool> seen;
        for (Request* r : rs) {
            if (r->pgroup < 0 || seen.count({r->family, r->pgroup})) continue;
            seen[{r->family, r->pgroup}] = true;
            const BatchFamily& f = B->fams[r->family];
            const BatchFamily::PGroup& pg = f.pgroups[r->pgroup];
            BGroup G{};
            G.nterms = f.nterms;
            G.compact = f.g->compact ? 1 : 0;
            G.nitems = f.nitems;
            G.n = f.n;
            G.P = pg.P;
            for (int b = 0; b < kGroupMax; ++b) {
                auto it = pg.req[b] >= 0 ? pos.find(pg.req[b]) : pos.end();
                G.inst[b] = it == pos.end() ? -1 : it->second;
                if (G.inst[b] >= 0)
                    for (int t = 0; t < f.nterms; ++t) {
                        G.muD[b][t] = rs[G.inst[b]]->mu[t];
                        G.muF[b][t] = (float)rs[G.inst[b]]->mu[t];
                    }
            }
            for (int t = 0; t < kMaxTerms; ++t) {
                const int u = t < f.nterms ? t : 0;
                G.terms[t] = B->opts.precision == FEEL_HIP_FP32 ? (const void*)f.termsF[u] : (const void*)f.terms[u];
            }
            G.mask = f.mask;
            G.tmask = f.tmask;
            G.cols = SellCols{f.g->slicePtr, f.g->scol, f.g->colBase, f.g->ccptr, f.g->packed};
            const int g = (int)grp.size();
            grp.push_back(G);
            for (int64_t s = 0; s < f.nslices; s += kGroupSlices)
                gitems.push_back({g, (int)s, (int)std::min<int64_t>(f.nslices, s + kGroupSlices), (int)(s / kGroupSlices)});
            for (int64_t s = 0, j = 0; s < f.nslices; s += kItemSlices, ++j)
                gditems.push_back({g, (int)s, (int)std::min<int64_t>(f.nslices, s + kItemSlices), (int)j});
        }
        // Workgroups go to the XCDs round robin (block i to XCD i % 8): give each
        // XCD one contiguous eighth of the (group, slice range) sequence, so the
        // blocks sharing an L2 gather from neighbouring rows of the same
        // instances.  Short eighths are padded with empty blocks (group -1).
        const size_t N = gitems.size(), per = (N + kXcds - 1) / kXcds;
        std::vector<BGItem> placed(per * kXcds, BGItem{-1, 0, 0, 0});
        for (int x = 0; x < kXcds; ++x) {
            const size_t b0 = N * x / kXcds, b1 = N * (x + 1) / kXcds;
            for (size_t i = b0; i < b1; ++i) placed[(i - b0) * kXcds + x] = gitems[i];
        }
        gitems.swap(placed);
    }
    feel_hip_ctx* c = B->ctx;
    if ((int)items.size() > B->capItems) {
        c->release(B->dItems, sizeof(BItem) * B->capItems);
        B->capItems = (int)items.size() * 2;
        TRY_CALL(alloc(c, &B->dItems, B->capItems));
        ++B->version;
    }
    if ((int)insts.size() > B->capInsts) {
        c->release(B->dInsts, sizeof(BInst) * B->capInsts);
        c->release(B->dSums, sizeof(double) * B->capInsts);
        B->capInsts = (int)insts.size() * 2;
        TRY_CALL(alloc(c, &B->dInsts, B->capInsts));
        TRY_CALL(alloc(c, &B->dSums, B->capInsts));
        ++B->version;
    }
    if ((int)grp.size() > B->capGroups) {
        c->release(B->dGroups, sizeof(BGroup) * B->capGroups);
        B->capGroups = (int)grp.size() * 2;
        TRY_CALL(alloc(c, &B->dGroups, B->capGroups));
        ++B->version;
    }
    if ((int)gitems.size() > B->capGItems) {
        c->release(B->dGItems, sizeof(BGItem) * B->capGItems);
        B->capGItems = (int)gitems.size() * 2;
        TRY_CALL(alloc(c, &B->dGItems, B->capGItems));
        ++B->version;
    }
    if ((int)gditems.size() > B->capGDItems) {
        c->release(B->dGDItems, sizeof(BGItem) * B->capGDItems);
        B->capGDItems = (int)gditems.size() * 2;
        TRY_CALL(alloc(c, &B->dGDItems, B->capGDItems));
        ++B->version;
    }
    if (!gditems.empty())
        FEEL_HIP_CHECK(hipMemcpyAsync(B->dGDItems, gditems.data(), sizeof(BGItem) * gditems.size(), hipMemcpyHostToDevice,
                                      c->stream));
    if (!grp.empty()) {
        FEEL_HIP_CHECK(hipMemcpyAsync(B->dGroups, grp.data(), sizeof(BGroup) * grp.size(), hipMemcpyHostToDevice, c->stream));
        FEEL_HIP_CHECK(
            hipMemcpyAsync(B->dGItems, gitems.data(), sizeof(BGItem) * gitems.size(), hipMemcpyHostToDevice, c->stream));
    }
    if (groups) {
        B->nGItems = (int)gitems.size();
        B->nGDItems = (int)gditems.size();
    }
    if (!items.empty()) {
        FEEL_HIP_CHECK(hipMemcpyAsync(B->dItems, items.data(), sizeof(BItem) * items.size(), hipMemcpyHostToDevice, c->stream));
        FEEL_HIP_CHECK(hipMemcpyAsync(B->dInsts, insts.data(), sizeof(BInst) * insts.size(), hipMemcpyHostToDevice, c->stream));
        FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));   // host vectors go out of scope
    }
    *nItems = (int)items.size();
    return FEEL_HIP_OK;
}

// per-instance ||D^{-1} r64|| after kBResidual on the same work list
int residualNorms(feel_hip_batch* B, const std::vector<Request*>& rs) {
    feel_hip_ctx* c = B->ctx;
    hipLaunchKernelGGL(kBInstSum, dim3((int)rs.size()), dim3(kBlock), 0, c->stream, B->dInsts, B->dSums);
    FEEL_HIP_CHECK(hipGetLastError());
    std::vector<double> h(rs.size());
    FEEL_HIP_CHECK(hipMemcpyAsync(h.data(), B->dSums, sizeof(double) * rs.size(), hipMemcpyDeviceToHost, c->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < rs.size(); ++i) rs[i]->dp = std::sqrt(h[i]);
    return FEEL_HIP_OK;
}

// grouped SpMV over every group of the current work list, instantiated for
// the largest term count of the list: a family with fewer terms pads them with
// its first term and mu = 0 (the repeated loads hit the cache lines just read)
template <typename T>
void launchGroupSpmv(feel_hip_batch* B, int nterms) {
    feel_hip_ctx* c = B->ctx;
    const dim3 grid(B->nGItems), block(kGroupBlock);
    switch (nterms) {
#define FEEL_GROUP_CASE(NT) \
    case NT: hipLaunchKernelGGL((kBSpmvGroup<T, NT>), grid, block, 0, c->stream, B->dGItems, B->dGroups, B->dInsts); break;
        FEEL_GROUP_CASE(1) FEEL_GROUP_CASE(2) FEEL_GROUP_CASE(3) FEEL_GROUP_CASE(4)
        FEEL_GROUP_CASE(5) FEEL_GROUP_CASE(6) FEEL_GROUP_CASE(7) FEEL_GROUP_CASE(8)
#undef FEEL_GROUP_CASE
    }
}

template <typename T>
int launchIterations(feel_hip_batch* B, int nItems, int chunk, int nterms) {
    feel_hip_ctx* c = B->ctx;
    const bool grouped = nterms > 0;
    auto body = [&]() {
        for (int k = 0; k < chunk; ++k) {
            // iteration k of the chunk: the counter moves once per chunk
            if (grouped) {
                hipLaunchKernelGGL(kBDirectionGroup<T>, dim3(B->nGDItems), dim3(kBlock), 0, c->stream, B->dGDItems,
                                   B->dGroups, B->dInsts, B->gcount, k);
                launchGroupSpmv<T>(B, nterms);
                hipLaunchKernelGGL(kBUpdateGroup<T>, dim3(B->nGDItems), dim3(kBlock), 0, c->stream, B->dGDItems, B->dGroups,
                                   B->dInsts, B->gcount, k);
            } else {
                hipLaunchKernelGGL(kBDirection<T>, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts,
                                   B->gcount, k);
                hipLaunchKernelGGL(kBSpmv<T>, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts);
                hipLaunchKernelGGL(kBUpdate<T>, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts,
                                   B->gcount, k);
            }
        }
        hipLaunchKernelGGL(kBBump, dim3(1), dim3(64), 0, c->stream, B->gcount, chunk);
    };
    if (!B->opts.use_graph) {
        body();
        FEEL_HIP_CHECK(hipGetLastError());
        return FEEL_HIP_OK;
    }
    // the captured chunk reads the work list, descriptors and counter from device
    // memory, so it stays valid while those buffers keep their addresses
    const int gItems = grouped ? B->nGItems : 0, gdItems = grouped ? B->nGDItems : 0;
    if (!B->exec || B->execItems != nItems || B->execGItems != gItems || B->execGDItems != gdItems ||
        B->execTerms != nterms ||
        B->execChunk != chunk || B->execVersion != B->version) {
        if (B->exec) { (void)hipGraphExecDestroy(B->exec); B->exec = nullptr; }
        hipGraph_t graph;
        FEEL_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        body();
        FEEL_HIP_CHECK(hipStreamEndCapture(c->stream, &graph));
        FEEL_HIP_CHECK(hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        B->execItems = nItems;
        B->execGItems = gItems;
        B->execGDItems = gdItems;
        B->execTerms = nterms;
        B->execChunk = chunk;
        B->execVersion = B->version;
    }
    FEEL_HIP_CHECK(hipGraphLaunch(B->exec, c->stream));
    return FEEL_HIP_OK;
}

// a family group slot for a request of the grouped SpMV (a new group
// when every slot of the family's groups is taken)
bool hasFreeSlot(const BatchFamily& f) {
    for (const auto& pg : f.pgroups)
        if (pg.P)
            for (int b = 0; b < kGroupMax; ++b)
                if (pg.req[b] < 0) return true;
    return false;
}

size_t groupBytes(const feel_hip_batch* B, const BatchFamily& f) { return tsize(B) * kGroupMax * (size_t)f.n; }

int assignSlot(feel_hip_batch* B, Request& r) {
    BatchFamily& f = B->fams[r.family];
    for (size_t g = 0; g < f.pgroups.size(); ++g)
        for (int b = 0; f.pgroups[g].P && b < kGroupMax; ++b)
            if (f.pgroups[g].req[b] < 0) {
                f.pgroups[g].req[b] = r.id;
                r.pgroup = (int)g;
                r.pslot = b;
                return FEEL_HIP_OK;
            }
    BatchFamily::PGroup pg;
    for (int b = 0; b < kGroupMax; ++b) pg.req[b] = -1;
    TRY_CALL(B->ctx->alloc((void**)&pg.P, groupBytes(B, f)));
    // slots without an instance are read (one row load serves all eight) but never used
    FEEL_HIP_CHECK(hipMemsetAsync(pg.P, 0, groupBytes(B, f), B->ctx->stream));
    B->used += (int64_t)groupBytes(B, f);
    pg.req[0] = r.id;
    f.pgroups.push_back(pg);
    r.pgroup = (int)f.pgroups.size() - 1;
    r.pslot = 0;
    return FEEL_HIP_OK;
}

void finish(feel_hip_batch* B, Request& r, int reason) {
    r.reason = reason;
    r.phase = Phase::Done;
    B->finished.push_back(r.id);
    if (r.pgroup >= 0) {   // the search direction is no longer needed
        B->fams[r.family].pgroups[r.pgroup].req[r.pslot] = -1;
        r.pgroup = r.pslot = -1;
    }
}

// copies the request's FP64 solution to Feel++ dof numbering (instance buffer `out`)
void permuteOut(feel_hip_batch* B, Request& r) {
    const BatchFamily& f = B->fams[r.family];
    const InstBuffers& ib = B->bufs[r.buf];
    hipLaunchKernelGGL(kBPermuteOut, dim3((int)std::min<int64_t>(4096, (f.n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       B->ctx->stream, f.n, f.g->perm, ib.x64, ib.out);
    r.unpermuted = true;
}

// Frees cached memory no request holds -- instance buffer sets kept from
// finished requests, family groups whose eight slots are all free -- of the
// families other than `keep` (whose cached memory the request being admitted
// counts on) until `need` more bytes fit the HBM budget.  Released groups stay
// in their family's list (request slot indices refer to it) with no storage.
// Returns whether the bytes fit.
bool releaseIdle(feel_hip_batch* B, int64_t need, int keep) {
    feel_hip_ctx* c = B->ctx;
    auto fits = [&] { return B->used + need <= B->opts.hbm_budget; };
    bool freed = false;
    for (size_t fi = 0; fi < B->fams.size(); ++fi) {
        if ((int)fi == keep) continue;
        BatchFamily& f = B->fams[fi];
        while (!fits() && !f.freeBufs.empty()) {
            InstBuffers& ib = B->bufs[f.freeBufs.back()];
            f.freeBufs.pop_back();
            for (double* q : {ib.val64, ib.dinv64, ib.x64, ib.r64, ib.w64, ib.out, ib.part}) if (q) (void)hipFree(q);
            for (void* q : {ib.valT, ib.dinvT, ib.xT, ib.rT, ib.zT, ib.pT, ib.wT, ib.bT}) if (q) (void)hipFree(q);
            ib = InstBuffers{};
            const int64_t bytes = (int64_t)instanceBytes(B, f);
            B->used -= bytes;
            c->bytes -= bytes;
            freed = true;
        }
        for (auto& pg : f.pgroups) {
            if (fits()) break;
            bool idle = pg.P != nullptr;
            for (int b = 0; idle && b < kGroupMax; ++b) idle = pg.req[b] < 0;
            if (!idle) continue;
            (void)hipFree(pg.P);
            pg.P = nullptr;
            B->used -= (int64_t)groupBytes(B, f);
            c->bytes -= (int64_t)groupBytes(B, f);
            freed = true;
        }
    }
    if (freed) ++B->version;
    return fits();
}

template <typename T>
int stepImpl(feel_hip_batch* B, int iterations, int* nActive, int* nDone) {
    feel_hip_ctx* c = B->ctx;
    int nItems = 0;
    // 1. admission by HBM footprint in submission order.  With backfill > 0 a
    //    request that does not fit lets later ones that fit pass it, until it
    //    has waited `backfill` steps; then nothing passes it until it fits.
    std::vector<Request*> admitted;
    bool blocked = false;
    for (auto it = B->pending.begin(); it != B->pending.end();) {
        Request& r = B->reqs[*it];
        BatchFamily& f = B->fams[r.family];
        const bool reuse = !f.freeBufs.empty();
        const bool grouped = B->opts.group_spmv != 0;
        const int64_t need = (reuse ? 0 : (int64_t)instanceBytes(B, f)) +
                             (grouped && !hasFreeSlot(f) ? (int64_t)groupBytes(B, f) : 0);
        if (need > 0 && B->opts.hbm_budget > 0 && B->used + need > B->opts.hbm_budget &&
            !releaseIdle(B, need, r.family)) {
            if (B->used == 0) return setError(FEEL_HIP_ENOMEM, "one instance exceeds the batch HBM budget");
            if (B->opts.backfill <= 0 || blocked) {
                if (B->opts.backfill <= 0) break;
                ++it;
                continue;
            }
            blocked = true;
            if (++r.waited > B->opts.backfill) break;
            ++it;
            continue;
        }
        TRY_CALL(ensureStates(B, B->nStates - (int)B->freeStates.size() + 1));
        TRY_CALL(allocInstance(B, r.family, &r.buf));
        if (grouped) TRY_CALL(assignSlot(B, r));
        r.state = B->freeStates.back();
        B->freeStates.pop_back();
        admitted.push_back(&r);
        it = B->pending.erase(it);
    }
    // 2. assembly of admitted instances: A(mu), D^{-1}; x = 0 so r = b
    if (!admitted.empty()) {
        for (Request* r : admitted) { r->dp = 0; r->xscale = 0; }
        TRY_CALL(buildWork(B, admitted, &nItems));
        hipLaunchKernelGGL(kBCombine<T>, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts);
        hipLaunchKernelGGL(kBDiag<T>, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts);
        for (Request* r : admitted) {
            const InstBuffers& ib = B->bufs[r->buf];
            const int64_t n = B->fams[r->family].n;
            FEEL_HIP_CHECK(hipMemsetAsync(ib.x64, 0, sizeof(double) * n, c->stream));
            FEEL_HIP_CHECK(hipMemsetAsync(ib.w64, 0, sizeof(double) * n, c->stream));
        }
        hipLaunchKernelGGL(kBResidual, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts);
        FEEL_HIP_CHECK(hipGetLastError());
        TRY_CALL(residualNorms(B, admitted));
        for (Request* r : admitted) {
            r->rnorm0 = r->dp;
            r->sigma = r->dp;
            r->ttol = std::max(B->opts.rtol * r->rnorm0, B->opts.atol);
            r->phase = Phase::Resid;
        }
    }
    // 3. FP64 residual holders: converged, or start an inner solve
    std::vector<Request*> toInner;
    for (auto& kv : B->reqs) {
        Request& r = kv.second;
        if (r.phase != Phase::Resid) continue;
        if (r.dp != r.dp) finish(B, r, -9);
        else if (r.dp <= r.ttol) finish(B, r, r.dp < B->opts.atol ? 3 : 2);
        else if (r.outer >= B->opts.max_outer || r.its >= B->opts.maxit) finish(B, r, -3);
        else toInner.push_back(&r);
    }
    // FP32 CG with reliable updates: the device stops once its residual
    // estimate has dropped by inner_rtol since the last FP64 residual (or
    // reaches the tolerance); the FP32 residual then restarts from the true one
    // and CG continues with the same direction and iteration count.
    std::vector<CgState> fresh(toInner.size());   // kept alive until buildWork synchronises
    if (!toInner.empty()) {
        for (size_t i = 0; i < toInner.size(); ++i) {
            Request* r = toInner[i];
            const bool restart = r->outer == 0 || r->dp > kBatchDriftRestart * r->estimate;
            if (restart) r->fresh = r->its;
            CgState& h = fresh[i];
            h = CgState{};
            const double stop = std::max(r->ttol, B->opts.inner_rtol * r->dp) / r->sigma;
            h.rtol = 0.0;
            h.atol = stop;
            h.ttol = stop;
            h.rnorm0 = r->devRnorm0;
            h.dtol = 1e5;
            h.maxit = B->opts.maxit;
            h.fresh = r->fresh;
            FEEL_HIP_CHECK(hipMemcpyAsync(B->states + r->state, &h, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
            // the chunk's first iteration runs as iteration r->its of this request
            r->start = B->hgcount - r->its;
            r->phase = Phase::Inner;
        }
        TRY_CALL(buildWork(B, toInner, &nItems));
        hipLaunchKernelGGL(kBInnerInit<T>, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts);
        FEEL_HIP_CHECK(hipGetLastError());
    }
    // 4. one chunk of iterations for every instance in an inner solve
    std::vector<Request*> inner;
    for (auto& kv : B->reqs)
        if (kv.second.phase == Phase::Inner) inner.push_back(&kv.second);
    if (!inner.empty() && iterations > 0) {
        const bool grouped = B->opts.group_spmv != 0;
        int maxTerms = 0;
        for (Request* r : inner) maxTerms = std::max(maxTerms, B->fams[r->family].nterms);
        TRY_CALL(buildWork(B, inner, &nItems, grouped));
        TRY_CALL(launchIterations<T>(B, nItems, iterations, grouped ? maxTerms : 0));
        B->hgcount += iterations;
        std::vector<CgState> hs(B->nStates);
        FEEL_HIP_CHECK(hipMemcpyAsync(hs.data(), B->states, sizeof(CgState) * B->nStates, hipMemcpyDeviceToHost, c->stream));
        FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
        std::vector<Request*> ended;
        std::vector<int> stopReason;   // per ended request: 0, or the reason the inner solve broke down
        for (Request* r : inner) {
            const CgState& st = hs[r->state];
            if (!st.done) continue;
            r->its = st.its;
            r->outer += 1;
            if (r->outer == 1) r->devRnorm0 = st.rnorm0;
            r->estimate = st.dp * r->sigma;
            if (st.reason == -9) { finish(B, *r, -9); continue; }   // x64 keeps the last finite update
            r->xscale = r->sigma;   // x64 += sigma * x_fp32 (the FP32 residual is r64 / sigma)
            ended.push_back(r);
            stopReason.push_back(st.reason == -8 || st.reason == -10 || st.reason == -4 ? st.reason : 0);
        }
        // 5. FP64 residual of the instances whose inner solve ended; a breakdown
        // (indefinite preconditioner or matrix, divergence) returns the iterate it reached with its true
        // FP64 residual, as KSPSolve leaves x at the last iterate
        if (!ended.empty()) {
            TRY_CALL(buildWork(B, ended, &nItems));
            hipLaunchKernelGGL(kBAxpy<T>, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts);
            hipLaunchKernelGGL(kBSpmv64, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts);
            hipLaunchKernelGGL(kBResidual, dim3(nItems), dim3(kBlock), 0, c->stream, B->dItems, B->dInsts);
            FEEL_HIP_CHECK(hipGetLastError());
            TRY_CALL(residualNorms(B, ended));
            for (size_t i = 0; i < ended.size(); ++i) {
                Request* r = ended[i];
                r->xscale = 0;
                r->phase = Phase::Resid;
                if (stopReason[i]) finish(B, *r, stopReason[i]);
            }
        }
    }
    // 6. solutions of the requests finished in this step, back to Feel++ numbering
    for (auto& kv : B->reqs)
        if (kv.second.phase == Phase::Done && kv.second.buf >= 0 && !kv.second.unpermuted) permuteOut(B, kv.second);
    int act = 0, done = 0;
    for (auto& kv : B->reqs) {
        if (kv.second.phase == Phase::Done) ++done;
        else if (kv.second.phase == Phase::Resid || kv.second.phase == Phase::Inner) ++act;
    }
    if (nActive) *nActive = act + (int)B->pending.size();
    if (nDone) *nDone = done;
    return FEEL_HIP_OK;
}

}  // namespace

extern "C" {

void feel_hip_batch_opts_default(feel_hip_batch_opts* o) {
    if (!o) return;
    o->rtol = 1e-8;
    o->atol = 1e-50;
    o->inner_rtol = 1e-2;
    o->maxit = 100000;
    o->inner_maxit = 100000;
    o->max_outer = 50;
    o->precision = FEEL_HIP_FP32;
    o->hbm_budget = 0;
    o->use_graph = 1;
    o->group_spmv = 1;
    o->backfill = 0;
}

int feel_hip_batch_create(feel_hip_ctx* ctx, const feel_hip_batch_opts* opts, feel_hip_batch** out) {
    if (!ctx || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_batch_opts o;
    if (opts) o = *opts;
    else feel_hip_batch_opts_default(&o);
    if (o.precision != FEEL_HIP_FP32 && o.precision != FEEL_HIP_FP64)
        return setError(FEEL_HIP_EINVAL, "batch precision must be FP32 (refined) or FP64");
    auto* B = new feel_hip_batch();
    B->ctx = ctx;
    B->opts = o;
    retain(ctx);
    int rc = alloc(ctx, &B->gcount, 1);
    if (rc == FEEL_HIP_OK && hipMemsetAsync(B->gcount, 0, sizeof(int), ctx->stream) != hipSuccess)
        rc = setError(FEEL_HIP_EDEVICE, "hipMemset failed");
    if (rc != FEEL_HIP_OK) { feel_hip_batch_destroy(B); return rc; }
    *out = B;
    return FEEL_HIP_OK;
}

int feel_hip_batch_destroy(feel_hip_batch* B) {
    if (!B) return FEEL_HIP_OK;
    feel_hip_ctx* c = B->ctx;
    (void)hipStreamSynchronize(c->stream);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    for (InstBuffers& ib : B->bufs) {
        for (double* p : {ib.val64, ib.dinv64, ib.x64, ib.r64, ib.w64, ib.out, ib.part}) if (p) (void)hipFree(p);
        for (void* p : {ib.valT, ib.dinvT, ib.xT, ib.rT, ib.zT, ib.pT, ib.wT, ib.bT}) if (p) (void)hipFree(p);
    }
    for (auto& f : B->fams)
        for (auto& pg : f.pgroups) (void)hipFree(pg.P);
    c->bytes -= B->used;   // instance and group buffers were counted through ctx->alloc
    for (auto& f : B->fams) {
        for (double* p : f.terms) c->release(p, sizeof(double) * f.nslots);
        for (float* p : f.termsF) c->release(p, sizeof(float) * f.nslots);
        c->release(f.mask, f.nslots);
        if (f.tmask) c->release(f.tmask, f.nslices);
        c->release(f.b64, sizeof(double) * f.n);
        unref(f.g);
    }
    c->release(B->states, sizeof(CgState) * B->nStates);
    c->release(B->slots, sizeof(double) * 4 * B->nStates);
    c->release(B->dItems, sizeof(BItem) * B->capItems);
    c->release(B->dInsts, sizeof(BInst) * B->capInsts);
    c->release(B->dGroups, sizeof(BGroup) * B->capGroups);
    c->release(B->dGItems, sizeof(BGItem) * B->capGItems);
    c->release(B->dGDItems, sizeof(BGItem) * B->capGDItems);
    c->release(B->dSums, sizeof(double) * B->capInsts);
    c->release(B->gcount, sizeof(int));
    delete B;
    unref(c);
    return FEEL_HIP_OK;
}

int feel_hip_batch_add_family(feel_hip_batch* B, int nterms, feel_hip_mat* const* terms, const feel_hip_vec* rhs,
                              const feel_hip_rows* dirichlet, int* family) {
    if (!B || !terms || !rhs || !family || nterms < 1) return setError(FEEL_HIP_EINVAL, "bad argument");
    if (nterms > kMaxTerms) return setError(FEEL_HIP_EUNSUPPORTED, "at most 8 affine terms");
    feel_hip_graph* g = terms[0]->g;
    for (int t = 0; t < nterms; ++t) {
        if (!terms[t] || terms[t]->g != g) return setError(FEEL_HIP_EINVAL, "all terms must share one graph");
        if (!terms[t]->v64) return setError(FEEL_HIP_EINVAL, "terms need FP64 values");
    }
    if (rhs->n != g->n) return setError(FEEL_HIP_EINVAL, "rhs size mismatch");
    if (g->block) return setError(FEEL_HIP_EUNSUPPORTED, "batched families need a scalar space");
    if (dirichlet && dirichlet->n != g->n) return setError(FEEL_HIP_EINVAL, "Dirichlet rows built for another size");
    feel_hip_ctx* c = B->ctx;
    BatchFamily f;
    f.g = g;
    f.nterms = nterms;
    f.n = g->n;
    f.nslices = g->nslices;
    f.nslots = g->nslots;
    f.nitems = (int)((f.nslices + kItemSlices - 1) / kItemSlices);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((f.nslots + kBlock - 1) / kBlock, 8192));
    for (int t = 0; t < nterms; ++t) {
        double* v = nullptr;
        TRY_CALL(alloc(c, &v, f.nslots));
        hipLaunchKernelGGL(kCsrToSell<double>, dim3(grid), dim3(kBlock), 0, c->stream, f.nslots, g->slot2csr, terms[t]->v64,
                           (const float*)nullptr, v);
        f.terms.push_back(v);
        if (B->opts.precision == FEEL_HIP_FP32 && B->opts.group_spmv) {   // FP32 groups stream FP32 terms
            float* vf = nullptr;
            TRY_CALL(alloc(c, &vf, f.nslots));
            hipLaunchKernelGGL(kCastF32, dim3(grid), dim3(kBlock), 0, c->stream, f.nslots, v, vf);
            f.termsF.push_back(vf);
        }
    }
    uint8_t* zeroFlags = nullptr;
    const uint8_t* flags = dirichlet ? dirichlet->flag : nullptr;
    if (!flags) {
        TRY_CALL(alloc(c, &zeroFlags, f.n));
        FEEL_HIP_CHECK(hipMemsetAsync(zeroFlags, 0, f.n, c->stream));
        flags = zeroFlags;
    }
    TRY_CALL(alloc(c, &f.mask, f.nslots));
    hipLaunchKernelGGL(kBMask, dim3(grid), dim3(kBlock), 0, c->stream, f.n, f.nslices, g->slicePtr, g->scol, g->slot2csr,
                       g->perm, flags, f.mask);
    if (B->opts.group_spmv) {
        TRY_CALL(alloc(c, &f.tmask, f.nslices));
        const int sgrid = (int)std::max<int64_t>(1, std::min<int64_t>((f.nslices * 64 + kBlock - 1) / kBlock, 8192));
        if (B->opts.precision == FEEL_HIP_FP32) {
            TermPtrs<float> T{};
            for (int t = 0; t < nterms; ++t) T.t[t] = f.termsF[t];
            hipLaunchKernelGGL(kBTermMask<float>, dim3(sgrid), dim3(kBlock), 0, c->stream, f.nslices, nterms,
                               g->slicePtr, T, f.mask, f.tmask);
        } else {
            TermPtrs<double> T{};
            for (int t = 0; t < nterms; ++t) T.t[t] = f.terms[t];
            hipLaunchKernelGGL(kBTermMask<double>, dim3(sgrid), dim3(kBlock), 0, c->stream, f.nslices, nterms,
                               g->slicePtr, T, f.mask, f.tmask);
        }
    }
    TRY_CALL(alloc(c, &f.b64, f.n));
    hipLaunchKernelGGL(kBRhs, dim3(grid), dim3(kBlock), 0, c->stream, f.n, g->perm, flags, rhs->d, f.b64);
    FEEL_HIP_CHECK(hipGetLastError());
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (zeroFlags) c->release(zeroFlags, f.n);
    retain(g);   // the family reads the graph's SELL image while the batch lives
    B->fams.push_back(std::move(f));
    *family = (int)B->fams.size() - 1;
    return FEEL_HIP_OK;
}

int feel_hip_batch_submit(feel_hip_batch* B, int family, const double* mu, int64_t* request) {
    if (!B || !mu || !request) return setError(FEEL_HIP_EINVAL, "null argument");
    if (family < 0 || family >= (int)B->fams.size()) return setError(FEEL_HIP_EINVAL, "unknown family");
    Request r;
    r.id = B->nextId++;
    r.family = family;
    r.mu.assign(mu, mu + B->fams[family].nterms);
    B->reqs[r.id] = r;
    B->pending.push_back(r.id);
    *request = r.id;
    return FEEL_HIP_OK;
}

int feel_hip_batch_step(feel_hip_batch* B, int iterations, int* n_active, int* n_done) {
    if (!B || iterations < 0) return setError(FEEL_HIP_EINVAL, "bad argument");
    try {
        return B->opts.precision == FEEL_HIP_FP32 ? stepImpl<float>(B, iterations, n_active, n_done)
                                                  : stepImpl<double>(B, iterations, n_active, n_done);
    } catch (const std::exception& e) {
        return setError(FEEL_HIP_EINVAL, e.what());
    }
}

int feel_hip_batch_poll(feel_hip_batch* B, int64_t* request, feel_hip_solve_data* sd) {
    if (!B || !request) return setError(FEEL_HIP_EINVAL, "null argument");
    // ids whose solution was already taken (and the request released) are skipped
    auto it = B->reqs.end();
    int64_t id = -1;
    while (!B->finished.empty() && it == B->reqs.end()) {
        id = B->finished.front();
        B->finished.pop_front();
        it = B->reqs.find(id);
    }
    if (it == B->reqs.end()) return 0;
    const Request& r = it->second;
    *request = id;
    if (sd) {
        sd->converged = r.reason > 0;
        sd->its = r.its;
        sd->residual = r.dp;
        sd->reason = r.reason;
        sd->outer_its = r.outer;
        sd->rnorm0 = r.rnorm0;
        sd->fp64_its = B->opts.precision == FEEL_HIP_FP64 ? r.its : 0;
    }
    return 1;
}

int feel_hip_batch_solution(feel_hip_batch* B, int64_t request, double* host_x) {
    if (!B) return setError(FEEL_HIP_EINVAL, "null batch");
    auto it = B->reqs.find(request);
    if (it == B->reqs.end() || it->second.phase != Phase::Done) return setError(FEEL_HIP_EINVAL, "request not finished");
    Request& r = it->second;
    const BatchFamily& f = B->fams[r.family];
    if (!r.unpermuted) permuteOut(B, r);
    if (host_x) {
        FEEL_HIP_CHECK(hipMemcpyAsync(host_x, B->bufs[r.buf].out, sizeof(double) * f.n, hipMemcpyDeviceToHost, B->ctx->stream));
    }
    FEEL_HIP_CHECK(hipStreamSynchronize(B->ctx->stream));
    // release the instance's buffers and CG state for later admissions
    B->fams[r.family].freeBufs.push_back(r.buf);
    B->freeStates.push_back(r.state);
    r.buf = -1;
    r.state = -1;
    r.phase = Phase::Released;
    B->reqs.erase(it);
    return FEEL_HIP_OK;
}

int64_t feel_hip_batch_device_bytes(const feel_hip_batch* B) { return B ? B->used : -1; }

}  // extern "C"
