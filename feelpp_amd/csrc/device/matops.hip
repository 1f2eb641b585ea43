// The rest of the MatrixSparse / Vector backend interface on the device:
// scale, axpy of matrices, copy, diagonal, transpose, matrix norms, energy,
// entry-wise insertion, and the vector reductions and indexed updates a
// MatrixSparse<double> / Vector<double> implementation needs beside assembly
// and solve (feelalg/matrixsparse.hpp, feelalg/vector.hpp).  Reductions write
// one partial per block of a fixed grid and the host adds them in block order,
// so every result is run-to-run deterministic; no floating-point atomics.
#include <algorithm>
#include <cmath>
#include <limits>
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

int gridOf(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16)); }

// lanes per CSR row from the mean row length (rows of 7 to ~220 entries here)
int lanesOf(const feel_hip_graph* g) {
    const double mean = g->n ? (double)g->nnz / (double)g->n : 0.0;
    return mean > 48 ? 16 : (mean > 12 ? 8 : 4);
}

__device__ inline double valAt(const double* v64, const float* v32, int64_t p) {
    return v64 ? v64[p] : (double)v32[p];
}

// ---------------------------------------------------------------- matrix kernels
// v64 *= a and the FP32 copy rounded from it (or v32 *= a when the matrix holds FP32 only)
__global__ __launch_bounds__(kBlock) void kMatScale(int64_t nnz, double a, double* __restrict__ v64,
                                                    float* __restrict__ v32) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        if (v64) {
            const double v = a * v64[p];
            v64[p] = v;
            if (v32) v32[p] = (float)v;
        } else {
            v32[p] = (float)(a * (double)v32[p]);
        }
    }
}

// Y += a X on one pattern
__global__ __launch_bounds__(kBlock) void kMatAxpySame(int64_t nnz, double a, const double* __restrict__ x64,
                                                       const float* __restrict__ x32, double* __restrict__ y64,
                                                       float* __restrict__ y32) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const double x = valAt(x64, x32, p);
        if (y64) {
            const double v = y64[p] + a * x;
            y64[p] = v;
            if (y32) y32[p] = (float)v;
        } else {
            y32[p] = (float)((double)y32[p] + a * x);
        }
    }
}

// position of column c in the sorted row [b, e) of ja, -1 when absent
__device__ inline int64_t findCol(const int32_t* __restrict__ ja, int64_t b, int64_t e, int32_t c) {
    while (b < e) {
        const int64_t m = (b + e) >> 1;
        const int32_t v = ja[m];
        if (v == c) return m;
        if (v < c) b = m + 1;
        else e = m;
    }
    return -1;
}

// Y += a X where X's pattern is contained in Y's (same row count): one thread
// per row of X, each entry located in Y's row by binary search.  apply = 0 only
// counts the entries of X missing from Y (per block partials in miss).
__global__ __launch_bounds__(kBlock) void kMatAxpySubset(int64_t n, double a, const int32_t* __restrict__ xia,
                                                         const int32_t* __restrict__ xja, const double* __restrict__ x64,
                                                         const float* __restrict__ x32, const int32_t* __restrict__ yia,
                                                         const int32_t* __restrict__ yja, double* __restrict__ y64,
                                                         float* __restrict__ y32, int apply, double* __restrict__ miss) {
    double missing = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t yb = yia[r], ye = yia[r + 1];
        for (int64_t p = xia[r]; p < xia[r + 1]; ++p) {
            const int64_t q = findCol(yja, yb, ye, xja[p]);
            if (q < 0) {
                missing += 1.0;
                continue;
            }
            if (!apply) continue;
            const double x = valAt(x64, x32, p);
            if (y64) {
                const double v = y64[q] + a * x;
                y64[q] = v;
                if (y32) y32[q] = (float)v;
            } else {
                y32[q] = (float)((double)y32[q] + a * x);
            }
        }
    }
    if (!apply) {
        missing = blockReduceSum(missing);
        if (threadIdx.x == 0) miss[blockIdx.x] = missing;
    }
}

// copy of the values of one matrix into another on the same pattern
__global__ __launch_bounds__(kBlock) void kMatCopy(int64_t nnz, const double* __restrict__ x64,
                                                   const float* __restrict__ x32, double* __restrict__ y64,
                                                   float* __restrict__ y32) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const double v = valAt(x64, x32, p);
        if (y64) y64[p] = v;
        if (y32) y32[p] = x32 ? x32[p] : (float)v;
    }
}

__global__ __launch_bounds__(kBlock) void kMatDiagonal(int64_t n, const int32_t* __restrict__ diagpos,
                                                       const double* __restrict__ v64, const float* __restrict__ v32,
                                                       double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = valAt(v64, v32, diagpos[i]);
}

// transpose map: tpos[p] = position of (j, i) for p = (i, j), -1 when the
// pattern lacks it; per block the count of missing entries
__global__ __launch_bounds__(kBlock) void kTransposeMap(int64_t n, const int32_t* __restrict__ ia,
                                                        const int32_t* __restrict__ ja, int32_t* __restrict__ tpos,
                                                        double* __restrict__ miss) {
    double missing = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        for (int64_t p = ia[r]; p < ia[r + 1]; ++p) {
            const int32_t c = ja[p];
            const int64_t q = findCol(ja, ia[c], ia[c + 1], (int32_t)r);
            tpos[p] = (int32_t)q;
            if (q < 0) missing += 1.0;
        }
    }
    missing = blockReduceSum(missing);
    if (threadIdx.x == 0) miss[blockIdx.x] = missing;
}

// At(i, j) = A(j, i) on a structurally symmetric pattern
__global__ __launch_bounds__(kBlock) void kMatTransposeVals(int64_t nnz, const int32_t* __restrict__ tpos,
                                                            const double* __restrict__ x64, const float* __restrict__ x32,
                                                            double* __restrict__ y64, float* __restrict__ y32) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = tpos[p];
        const double v = valAt(x64, x32, q);
        if (y64) y64[p] = v;
        if (y32) y32[p] = x32 ? x32[q] : (float)v;
    }
}

// Per block: the largest row sum of |A(i, j)| (kind 0), the largest column
// sum through the transpose map (kind 1: row c of the map lists column c),
// or the sum of A(i, j)^2 (kind 2).  L lanes per row.
template <int L>
__global__ __launch_bounds__(kBlock) void kMatNorm(int64_t n, const int32_t* __restrict__ ia,
                                                   const int32_t* __restrict__ tpos, const double* __restrict__ v64,
                                                   const float* __restrict__ v32, int kind, double* __restrict__ part) {
    const int lane = threadIdx.x & (L - 1);
    constexpr int RPB = kBlock / L;
    double acc = 0.0;   // max for kinds 0 and 1 (sums of absolute values are >= 0), sum for kind 2
    for (int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / L; r < n; r += (int64_t)gridDim.x * RPB) {
        double s = 0.0;
        for (int64_t p = ia[r] + lane; p < ia[r + 1]; p += L) {
            const double v = valAt(v64, v32, kind == 1 ? (int64_t)tpos[p] : p);
            s += kind == 2 ? v * v : fabs(v);
        }
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        if (kind == 2) {
            if (lane == 0) acc += s;
        } else {
            acc = fmax(acc, s);
        }
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kBlock; ++i) t = kind == 2 ? t + sh[i] : fmax(t, sh[i]);
        part[blockIdx.x] = t;
    }
}

// Per block partial of sum_i v_i (A u)_i with L lanes per row
template <int L>
__global__ __launch_bounds__(kBlock) void kMatEnergy(int64_t n, const int32_t* __restrict__ ia,
                                                     const int32_t* __restrict__ ja, const double* __restrict__ v64,
                                                     const float* __restrict__ v32, const double* __restrict__ u,
                                                     const double* __restrict__ v, double* __restrict__ part) {
    const int lane = threadIdx.x & (L - 1);
    constexpr int RPB = kBlock / L;
    double acc = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / L; r < n; r += (int64_t)gridDim.x * RPB) {
        double s = 0.0;
        for (int64_t p = ia[r] + lane; p < ia[r + 1]; p += L) s += valAt(v64, v32, p) * u[ja[p]];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        if (lane == 0) acc += v[r] * s;
    }
    acc = blockReduceSum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// y += A x (CSR, FP64), L lanes per row
template <int L>
__global__ __launch_bounds__(kBlock) void kMatMultAdd(int64_t n, const int32_t* __restrict__ ia,
                                                      const int32_t* __restrict__ ja, const double* __restrict__ v64,
                                                      const float* __restrict__ v32, const double* __restrict__ x,
                                                      double* __restrict__ y) {
    const int lane = threadIdx.x & (L - 1);
    constexpr int RPB = kBlock / L;
    for (int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / L; r < n; r += (int64_t)gridDim.x * RPB) {
        double s = 0.0;
        for (int64_t p = ia[r] + lane; p < ia[r + 1]; p += L) s += valAt(v64, v32, p) * x[ja[p]];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        if (lane == 0) y[r] += s;
    }
}

// INSERT (add = 0) or ADD (add = 1) of m entries at known CSR positions; the
// positions are distinct (the host merged repeated entries in order)
__global__ __launch_bounds__(kBlock) void kMatPut(int64_t m, const int64_t* __restrict__ pos,
                                                  const double* __restrict__ vals, int add, double* __restrict__ v64,
                                                  float* __restrict__ v32) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = pos[k];
        if (v64) {
            const double v = add ? v64[p] + vals[k] : vals[k];
            v64[p] = v;
            if (v32) v32[p] = (float)v;
        } else {
            v32[p] = (float)(add ? (double)v32[p] + vals[k] : vals[k]);
        }
    }
}

// ---------------------------------------------------------------- vector kernels
__global__ __launch_bounds__(kBlock) void kVecFill(int64_t n, double s, double* __restrict__ v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = s;
}

__global__ __launch_bounds__(kBlock) void kVecShift(int64_t n, double s, double* __restrict__ v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] += s;
}

// per block: sum (0), sum of |v| (1), sum of v^2 (2), min (3), max (4), max |v| (5)
__global__ __launch_bounds__(kBlock) void kVecReduce(int64_t n, const double* __restrict__ v, int kind,
                                                     double* __restrict__ part) {
    const bool extremum = kind >= 3;
    double acc = kind == 3 ? std::numeric_limits<double>::infinity()
                           : (kind == 4 ? -std::numeric_limits<double>::infinity() : 0.0);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = v[i];
        switch (kind) {
            case 0: acc += x; break;
            case 1: acc += fabs(x); break;
            case 2: acc += x * x; break;
            case 3: acc = fmin(acc, x); break;
            case 4: acc = fmax(acc, x); break;
            default: acc = fmax(acc, fabs(x)); break;
        }
    }
    if (!extremum) {
        acc = blockReduceSum(acc);
        if (threadIdx.x == 0) part[blockIdx.x] = acc;
        return;
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int i = 1; i < kBlock; ++i) t = kind == 3 ? fmin(t, sh[i]) : fmax(t, sh[i]);
        part[blockIdx.x] = t;
    }
}

// v[idx[k]] = vals[k] (add = 0) or v[idx[k]] += vals[k] (add = 1); distinct indices
__global__ __launch_bounds__(kBlock) void kVecPut(int64_t m, const int32_t* __restrict__ idx,
                                                  const double* __restrict__ vals, int add, double* __restrict__ v) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
        v[idx[k]] = add ? v[idx[k]] + vals[k] : vals[k];
}

// ---------------------------------------------------------------- host helpers
feel_hip_ctx* ctxOf(const feel_hip_mat* A) { return A->g->space->ctx; }

void invalidate(feel_hip_mat* A) {
    A->dinvValid = false;
    A->sell64Valid = A->sell32Valid = false;
    elemOpInvalidate(A);
}

// sum (or max) of the first count block partials in c->part, in block order
int hostPartials(feel_hip_ctx* c, int count, bool max, double* out) {
    std::vector<double> h(count);
    FEEL_HIP_CHECK(hipMemcpyAsync(h.data(), c->part, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    double s = max ? h[0] : 0.0;
    for (int i = 0; i < count; ++i) s = max ? std::max(s, h[i]) : s + h[i];
    *out = s;
    return FEEL_HIP_OK;
}

bool samePattern(const feel_hip_graph* a, const feel_hip_graph* b) {
    if (a == b) return true;
    return a->n == b->n && a->nnz == b->nnz && a->hia == b->hia && a->hja == b->hja;
}

// the transpose map of g, built on first use; EUNSUPPORTED when the pattern is
// not structurally symmetric
int transposeMap(feel_hip_graph* g) {
    if (g->symmetric == 1) return FEEL_HIP_OK;
    if (g->symmetric == -1) return setError(FEEL_HIP_EUNSUPPORTED, "the sparsity graph is not structurally symmetric");
    feel_hip_ctx* c = g->space->ctx;
    TRY_CALL(c->alloc((void**)&g->tpos, sizeof(int32_t) * (size_t)g->nnz));
    hipLaunchKernelGGL(kTransposeMap, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, g->tpos,
                       c->part);
    FEEL_HIP_CHECK(hipGetLastError());
    double missing = 0;
    TRY_CALL(hostPartials(c, kReduceBlocks, false, &missing));
    if (missing > 0) {
        c->release(g->tpos, sizeof(int32_t) * (size_t)g->nnz);
        g->tpos = nullptr;
        g->symmetric = -1;
        return setError(FEEL_HIP_EUNSUPPORTED, "the sparsity graph is not structurally symmetric");
    }
    g->symmetric = 1;
    return FEEL_HIP_OK;
}

// entries of (rows x cols) as CSR positions, -1 outside the pattern
int positionsOf(const feel_hip_graph* g, int nrows, const int32_t* rows, int ncols, const int32_t* cols,
                std::vector<int64_t>& pos) {
    pos.resize((size_t)nrows * ncols);
    for (int i = 0; i < nrows; ++i) {
        if (rows[i] < 0 || rows[i] >= g->n) return setError(FEEL_HIP_EINVAL, "row out of range");
        const int32_t* b = g->hja.data() + g->hia[rows[i]];
        const int32_t* e = g->hja.data() + g->hia[rows[i] + 1];
        for (int j = 0; j < ncols; ++j) {
            const int32_t* p = std::lower_bound(b, e, cols[j]);
            pos[(size_t)i * ncols + j] = (p == e || *p != cols[j]) ? -1 : (int64_t)(p - g->hja.data());
        }
    }
    return FEEL_HIP_OK;
}

// Repeated keys merged in input order: INSERT keeps the last value, ADD sums
// the values in order (the order MatSetValues / VecSetValues apply them).
template <typename K>
void mergeRepeated(std::vector<K>& key, std::vector<double>& val, bool add) {
    std::vector<size_t> ord(key.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
    std::vector<K> k2;
    std::vector<double> v2;
    for (size_t t = 0; t < ord.size(); ++t) {
        const size_t i = ord[t];
        if (!k2.empty() && k2.back() == key[i]) v2.back() = add ? v2.back() + val[i] : val[i];
        else { k2.push_back(key[i]); v2.push_back(val[i]); }
    }
    key.swap(k2);
    val.swap(v2);
}

// uploads (pos, val) into fresh device buffers and launches kMatPut
int putEntries(feel_hip_mat* A, std::vector<int64_t>& pos, std::vector<double>& val, int add) {
    feel_hip_ctx* c = ctxOf(A);
    mergeRepeated(pos, val, add != 0);
    const int64_t m = (int64_t)pos.size();
    if (m == 0) return FEEL_HIP_OK;
    int64_t* dpos = nullptr;
    double* dval = nullptr;
    TRY_CALL(c->alloc((void**)&dpos, sizeof(int64_t) * m));
    int rc = c->alloc((void**)&dval, sizeof(double) * m);
    if (rc == FEEL_HIP_OK) {
        hipError_t e = hipMemcpyAsync(dpos, pos.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dval, val.data(), sizeof(double) * m, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(kMatPut, dim3(gridOf(m)), dim3(kBlock), 0, c->stream, m, dpos, dval, add, A->v64, A->v32);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = setError(FEEL_HIP_EDEVICE, hipGetErrorString(e));
    }
    c->release(dpos, sizeof(int64_t) * m);
    c->release(dval, sizeof(double) * m);
    invalidate(A);
    return rc;
}

int putVector(feel_hip_vec* v, int64_t m, const int32_t* idx, const double* vals, int add) {
    if (!v || m < 0 || (m > 0 && (!idx || !vals))) return setError(FEEL_HIP_EINVAL, "bad argument");
    if (m == 0) return FEEL_HIP_OK;
    std::vector<int32_t> key(idx, idx + m);
    std::vector<double> val(vals, vals + m);
    for (int32_t i : key)
        if (i < 0 || i >= v->n) return setError(FEEL_HIP_EINVAL, "index out of range");
    mergeRepeated(key, val, add != 0);
    feel_hip_ctx* c = v->ctx;
    const int64_t u = (int64_t)key.size();
    int32_t* didx = nullptr;
    double* dval = nullptr;
    TRY_CALL(c->alloc((void**)&didx, sizeof(int32_t) * u));
    int rc = c->alloc((void**)&dval, sizeof(double) * u);
    if (rc == FEEL_HIP_OK) {
        hipError_t e = hipMemcpyAsync(didx, key.data(), sizeof(int32_t) * u, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dval, val.data(), sizeof(double) * u, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(kVecPut, dim3(gridOf(u)), dim3(kBlock), 0, c->stream, u, didx, dval, add, v->d);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = setError(FEEL_HIP_EDEVICE, hipGetErrorString(e));
    }
    c->release(didx, sizeof(int32_t) * u);
    c->release(dval, sizeof(double) * u);
    return rc;
}

}  // namespace

extern "C" {

int feel_hip_mat_scale(feel_hip_mat* A, double a) {
    if (!A) return setError(FEEL_HIP_EINVAL, "null matrix");
    feel_hip_ctx* c = ctxOf(A);
    const int64_t nnz = A->g->nnz;
    if (nnz) hipLaunchKernelGGL(kMatScale, dim3(gridOf(nnz)), dim3(kBlock), 0, c->stream, nnz, a, A->v64, A->v32);
    FEEL_HIP_CHECK(hipGetLastError());
    invalidate(A);
    return FEEL_HIP_OK;
}

int feel_hip_mat_axpy(feel_hip_mat* Y, double a, const feel_hip_mat* X, int structure) {
    if (!Y || !X) return setError(FEEL_HIP_EINVAL, "null matrix");
    if (structure != FEEL_HIP_SAME_NONZERO_PATTERN && structure != FEEL_HIP_DIFFERENT_NONZERO_PATTERN &&
        structure != FEEL_HIP_SUBSET_NONZERO_PATTERN)
        return setError(FEEL_HIP_EINVAL, "unknown matrix structure");
    feel_hip_graph *gy = Y->g, *gx = X->g;
    if (gx->space->ctx != gy->space->ctx) return setError(FEEL_HIP_EINVAL, "matrices on different contexts");
    if (gx->n != gy->n) return setError(FEEL_HIP_EINVAL, "size mismatch");
    feel_hip_ctx* c = ctxOf(Y);
    if (samePattern(gx, gy)) {
        if (gy->nnz)
            hipLaunchKernelGGL(kMatAxpySame, dim3(gridOf(gy->nnz)), dim3(kBlock), 0, c->stream, gy->nnz, a, X->v64, X->v32,
                               Y->v64, Y->v32);
    } else {
        if (structure == FEEL_HIP_SAME_NONZERO_PATTERN)
            return setError(FEEL_HIP_EPATTERN, "SAME_NONZERO_PATTERN given for matrices of different patterns");
        // the result must fit Y's graph: every entry of X must be in it
        hipLaunchKernelGGL(kMatAxpySubset, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, gy->n, a, gx->ia, gx->ja,
                           X->v64, X->v32, gy->ia, gy->ja, Y->v64, Y->v32, 0, c->part);
        FEEL_HIP_CHECK(hipGetLastError());
        double missing = 0;
        TRY_CALL(hostPartials(c, kReduceBlocks, false, &missing));
        if (missing > 0) return setError(FEEL_HIP_EPATTERN, "X has entries outside the sparsity graph of Y");
        hipLaunchKernelGGL(kMatAxpySubset, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, gy->n, a, gx->ia, gx->ja,
                           X->v64, X->v32, gy->ia, gy->ja, Y->v64, Y->v32, 1, c->part);
    }
    FEEL_HIP_CHECK(hipGetLastError());
    invalidate(Y);
    return FEEL_HIP_OK;
}

int feel_hip_mat_copy(feel_hip_mat* dst, const feel_hip_mat* src) {
    if (!dst || !src) return setError(FEEL_HIP_EINVAL, "null matrix");
    if (!samePattern(dst->g, src->g)) return setError(FEEL_HIP_EPATTERN, "matrices of different patterns");
    if (dst->g->space->ctx != src->g->space->ctx) return setError(FEEL_HIP_EINVAL, "matrices on different contexts");
    feel_hip_ctx* c = ctxOf(dst);
    const int64_t nnz = dst->g->nnz;
    if (nnz)
        hipLaunchKernelGGL(kMatCopy, dim3(gridOf(nnz)), dim3(kBlock), 0, c->stream, nnz, src->v64, src->v32, dst->v64,
                           dst->v32);
    FEEL_HIP_CHECK(hipGetLastError());
    invalidate(dst);
    return FEEL_HIP_OK;
}

int feel_hip_mat_diagonal(const feel_hip_mat* A, feel_hip_vec* d) {
    if (!A || !d) return setError(FEEL_HIP_EINVAL, "null argument");
    if (d->n != A->g->n) return setError(FEEL_HIP_EINVAL, "size mismatch");
    feel_hip_ctx* c = ctxOf(A);
    if (d->n)
        hipLaunchKernelGGL(kMatDiagonal, dim3(gridOf(d->n)), dim3(kBlock), 0, c->stream, d->n, A->g->diagpos, A->v64,
                           A->v32, d->d);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

int feel_hip_mat_transpose(const feel_hip_mat* A, feel_hip_mat* At) {
    if (!A || !At) return setError(FEEL_HIP_EINVAL, "null matrix");
    if (A == At) return setError(FEEL_HIP_EINVAL, "in-place transpose is not supported");
    if (!samePattern(A->g, At->g)) return setError(FEEL_HIP_EPATTERN, "the transpose needs a matrix on the same pattern");
    TRY_CALL(transposeMap(A->g));
    feel_hip_ctx* c = ctxOf(A);
    const int64_t nnz = A->g->nnz;
    if (nnz)
        hipLaunchKernelGGL(kMatTransposeVals, dim3(gridOf(nnz)), dim3(kBlock), 0, c->stream, nnz, A->g->tpos, A->v64,
                           A->v32, At->v64, At->v32);
    FEEL_HIP_CHECK(hipGetLastError());
    invalidate(At);
    return FEEL_HIP_OK;
}

int feel_hip_mat_norm(const feel_hip_mat* A, int type, double* out) {
    if (!A || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = ctxOf(A);
    int kind;
    if (type == FEEL_HIP_NORM_INF) kind = 0;
    else if (type == FEEL_HIP_NORM_1) kind = 1;
    else if (type == FEEL_HIP_NORM_FROBENIUS) kind = 2;
    else return setError(FEEL_HIP_EINVAL, "unknown norm type");
    if (kind == 1) TRY_CALL(transposeMap(g));
    if (g->n == 0) { *out = 0.0; return FEEL_HIP_OK; }
    const int L = lanesOf(g);
    if (L == 16)
        hipLaunchKernelGGL(kMatNorm<16>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, g->n, g->ia, g->tpos, A->v64,
                           A->v32, kind, c->part);
    else if (L == 8)
        hipLaunchKernelGGL(kMatNorm<8>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, g->n, g->ia, g->tpos, A->v64,
                           A->v32, kind, c->part);
    else
        hipLaunchKernelGGL(kMatNorm<4>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, g->n, g->ia, g->tpos, A->v64,
                           A->v32, kind, c->part);
    FEEL_HIP_CHECK(hipGetLastError());
    double v = 0;
    TRY_CALL(hostPartials(c, kReduceBlocks, kind != 2, &v));
    *out = kind == 2 ? std::sqrt(v) : v;
    return FEEL_HIP_OK;
}

int feel_hip_mat_energy(const feel_hip_mat* A, const feel_hip_vec* v, const feel_hip_vec* u, int transpose, double* out) {
    if (!A || !v || !u || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_graph* g = A->g;
    if (v->n != g->n || u->n != g->n) return setError(FEEL_HIP_EINVAL, "size mismatch");
    feel_hip_ctx* c = ctxOf(A);
    if (g->n == 0) { *out = 0.0; return FEEL_HIP_OK; }
    // v^T A^T u = u^T A v
    const double* a = transpose ? v->d : u->d;
    const double* b = transpose ? u->d : v->d;
    const int L = lanesOf(g);
    if (L == 16)
        hipLaunchKernelGGL(kMatEnergy<16>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, A->v64,
                           A->v32, a, b, c->part);
    else if (L == 8)
        hipLaunchKernelGGL(kMatEnergy<8>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, A->v64,
                           A->v32, a, b, c->part);
    else
        hipLaunchKernelGGL(kMatEnergy<4>, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, A->v64,
                           A->v32, a, b, c->part);
    FEEL_HIP_CHECK(hipGetLastError());
    return hostPartials(c, kReduceBlocks, false, out);
}

int feel_hip_mat_set_matrix(feel_hip_mat* A, int nrows, const int32_t* rows, int ncols, const int32_t* cols,
                            const double* data) {
    if (!A || nrows < 0 || ncols < 0 || (nrows && ncols && (!rows || !cols || !data)))
        return setError(FEEL_HIP_EINVAL, "bad argument");
    try {
        std::vector<int64_t> pos;
        TRY_CALL(positionsOf(A->g, nrows, rows, ncols, cols, pos));
        for (int64_t p : pos)
            if (p < 0) return setError(FEEL_HIP_EPATTERN, "entry outside the sparsity graph");
        std::vector<double> val(data, data + pos.size());
        return putEntries(A, pos, val, 0);
    } catch (const std::exception& e) {
        return setError(FEEL_HIP_ENOMEM, e.what());
    }
}

int feel_hip_mat_get_entries(const feel_hip_mat* A, int64_t m, const int32_t* rows, const int32_t* cols, double* out) {
    if (!A || m < 0 || (m > 0 && (!rows || !cols || !out))) return setError(FEEL_HIP_EINVAL, "bad argument");
    if (!A->v64 && !A->v32) return setError(FEEL_HIP_EINVAL, "matrix holds no values");
    feel_hip_ctx* c = ctxOf(A);
    const feel_hip_graph* g = A->g;
    for (int64_t k = 0; k < m; ++k) {
        if (rows[k] < 0 || rows[k] >= g->n) return setError(FEEL_HIP_EINVAL, "row out of range");
        const int32_t* b = g->hja.data() + g->hia[rows[k]];
        const int32_t* e = g->hja.data() + g->hia[rows[k] + 1];
        const int32_t* p = std::lower_bound(b, e, cols[k]);
        out[k] = 0.0;   // entries outside the pattern read as zero
        if (p == e || *p != cols[k]) continue;
        const size_t q = (size_t)(p - g->hja.data());
        if (A->v64) {
            FEEL_HIP_CHECK(hipMemcpyAsync(&out[k], A->v64 + q, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        } else {
            float f = 0.0f;
            FEEL_HIP_CHECK(hipMemcpyAsync(&f, A->v32 + q, sizeof(float), hipMemcpyDeviceToHost, c->stream));
            FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
            out[k] = f;
        }
    }
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    return FEEL_HIP_OK;
}

int feel_hip_vec_mult_add(feel_hip_vec* y, const feel_hip_mat* A, const feel_hip_vec* x) {
    if (!y || !A || !x) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_graph* g = A->g;
    if (x->n != g->n || y->n != g->n) return setError(FEEL_HIP_EINVAL, "size mismatch");
    if (x == y) return setError(FEEL_HIP_EINVAL, "x and y must be distinct vectors");
    feel_hip_ctx* c = ctxOf(A);
    if (g->n == 0) return FEEL_HIP_OK;
    const int L = lanesOf(g);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((g->n + (kBlock / L) - 1) / (kBlock / L), 8192));
    if (L == 16)
        hipLaunchKernelGGL(kMatMultAdd<16>, dim3(grid), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, A->v64, A->v32,
                           x->d, y->d);
    else if (L == 8)
        hipLaunchKernelGGL(kMatMultAdd<8>, dim3(grid), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, A->v64, A->v32,
                           x->d, y->d);
    else
        hipLaunchKernelGGL(kMatMultAdd<4>, dim3(grid), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, A->v64, A->v32,
                           x->d, y->d);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

int feel_hip_vec_set_constant(feel_hip_vec* v, double s) {
    if (!v) return setError(FEEL_HIP_EINVAL, "null vector");
    if (v->n) hipLaunchKernelGGL(kVecFill, dim3(gridOf(v->n)), dim3(kBlock), 0, v->ctx->stream, v->n, s, v->d);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

int feel_hip_vec_shift(feel_hip_vec* v, double s) {
    if (!v) return setError(FEEL_HIP_EINVAL, "null vector");
    if (v->n) hipLaunchKernelGGL(kVecShift, dim3(gridOf(v->n)), dim3(kBlock), 0, v->ctx->stream, v->n, s, v->d);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

int feel_hip_vec_reduce(const feel_hip_vec* v, int kind, double* out) {
    if (!v || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    if (kind < FEEL_HIP_VEC_SUM || kind > FEEL_HIP_VEC_LINF) return setError(FEEL_HIP_EINVAL, "unknown reduction");
    feel_hip_ctx* c = v->ctx;
    if (v->n == 0) {
        // sums and norms of an empty vector are 0; it has no minimum or maximum
        if (kind == FEEL_HIP_VEC_MIN || kind == FEEL_HIP_VEC_MAX) return setError(FEEL_HIP_EINVAL, "empty vector");
        *out = 0.0;
        return FEEL_HIP_OK;
    }
    hipLaunchKernelGGL(kVecReduce, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, v->n, v->d, kind, c->part);
    FEEL_HIP_CHECK(hipGetLastError());
    std::vector<double> h(kReduceBlocks);
    FEEL_HIP_CHECK(hipMemcpyAsync(h.data(), c->part, sizeof(double) * kReduceBlocks, hipMemcpyDeviceToHost, c->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    double s = h[0];
    if (kind <= FEEL_HIP_VEC_L2) {
        s = 0.0;
        for (double x : h) s += x;
        if (kind == FEEL_HIP_VEC_L2) s = std::sqrt(s);
    } else {
        for (double x : h) s = kind == FEEL_HIP_VEC_MIN ? std::min(s, x) : std::max(s, x);
    }
    *out = s;
    return FEEL_HIP_OK;
}

int feel_hip_vec_set_indexed(feel_hip_vec* v, int64_t m, const int32_t* idx, const double* vals) {
    try {
        return putVector(v, m, idx, vals, 0);
    } catch (const std::exception& e) {
        return setError(FEEL_HIP_ENOMEM, e.what());
    }
}

int feel_hip_vec_add_indexed(feel_hip_vec* v, int64_t m, const int32_t* idx, const double* vals) {
    try {
        return putVector(v, m, idx, vals, 1);
    } catch (const std::exception& e) {
        return setError(FEEL_HIP_ENOMEM, e.what());
    }
}

}  // extern "C"
