// Element-by-element FP32 operator of a P2 tetrahedral diffusion matrix, for
// the inner iterations of the mixed-precision solve.
//
// When a matrix's values are exactly form2 = integrate(elements(mesh),
// inner(k grad u, grad v)) on P2 tetrahedra (bilinearformcontext.hpp:455-628,
// assembled with add = 0 and a rule exact for the degree-2 integrand),
// followed only by symmetric eliminations with a unit diagonal
// (matrixpetsc.cpp:1764-1850, MatZeroRowsColumns), A p can be applied without
// the assembled matrix:
//   pass 1, one thread per element: gather p at the element's ten dofs
//     (Dirichlet columns read as 0), y_e = sum_q w_q G_q F_e G_q^T p_e with the
//     reference gradients G_q at the four points of a degree-2 rule (exact for
//     the quadratic integrand, so the same integrals as the assembly's rule)
//     and the element's factor F_e = k J K^{-1} K^{-T} (six FP32 numbers),
//     written to ten planes Y[l][e];
//   pass 2, one lane per solver row (64-row slices): the row's incident
//     entries of Y summed in a fixed order (a SELL image of Y indices); a
//     Dirichlet row returns p_i (unit diagonal, zero row and column).
// Per element 104 bytes are streamed (factors 24, dofs 40, Y 40) and per row
// about 8 bytes per incident element (index, Y entry), against about 6.4 bytes
// per nonzero for the assembled FP32 SELL SpMV (28 nonzeros per P2 row).
// Elements are ordered by their smallest solver-numbered dof, so consecutive
// elements gather nearby entries of p and a slice of rows reads nearby Y.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

using namespace feelpp_amd;

namespace feelpp_amd {

struct ElemOp {
    int64_t ne = 0, n = 0;
    // structure (solver numbering), built once per matrix
    int32_t* edof = nullptr;        // [10][ne] solver dofs; bit 31 marks a Dirichlet column
    int32_t* edof0 = nullptr;       // [10][ne] solver dofs, unflagged
    int32_t* eorder = nullptr;      // ne: element id at each position of the element order
    uint32_t* incPtr = nullptr;     // nslices + 1: first slot of each 64-row slice
    int32_t* inc = nullptr;         // nslots: Y index per slot, padding -> 10 ne (a zero entry)
    int64_t nslices = 0, nslots = 0;
    // values
    float* F = nullptr;             // [6][ne] k J (K^{-1} K^{-T}): 00 11 22 01 02 12, element order
    bool factorsValid = false;      // F holds the factors of the matrix's last assembly
    uint8_t* brow = nullptr;        // n, Dirichlet rows in solver numbering
    float* Y = nullptr;             // 10 ne + 1
    bool flagsDirty = true;
};

}  // namespace feelpp_amd

namespace {

constexpr int kNq = 4;      // degree-2 rule on the tetrahedron
constexpr int kN = 10;      // P2 local dofs

// P2 nodal basis on the reference tetrahedron [-1,1]^3 (vertices (-1,-1,-1),
// (1,-1,-1), (-1,1,-1), (-1,-1,1); local dofs: the vertices, then the edges
// (1,2) (0,2) (0,1) (0,3) (1,3) (2,3), the order of the P2 node set):
// phi_v = l_v (2 l_v - 1), phi_(a,b) = 4 l_a l_b, with l_k = (xi_k + 1) / 2 for
// k = 1..3 and l_0 = 1 - l_1 - l_2 - l_3, so grad l_0 = -(1,1,1)/2 and
// grad l_k = e_k / 2.  The degree-2 rule: the four points with barycentric
// coordinates (a, b, b, b) and permutations, weight |T_ref| / 4 = 1/3 each.
constexpr float kAlpha = 0.58541019662496845446f;   // (5 + 3 sqrt 5) / 20
constexpr float kBeta = 0.13819660112501051518f;    // (5 - sqrt 5) / 20
constexpr float kWq = 1.0f / 3.0f;
constexpr int kEdge[6][2] = {{1, 2}, {0, 2}, {0, 1}, {0, 3}, {1, 3}, {2, 3}};

#define EO_CHECK(expr)                                                                                   \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            rc = setError(FEEL_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));         \
            goto done;                                                                                   \
        }                                                                                                \
    } while (0)

#define TRY_CALL_EO(expr)                    \
    do {                                     \
        int rc_ = (expr);                    \
        if (rc_ != FEEL_HIP_OK) return rc_;  \
    } while (0)

int gridOf(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 65536)); }

__global__ __launch_bounds__(kBlock) void kEoInverse(int64_t n, const int32_t* __restrict__ perm,
                                                     int32_t* __restrict__ iperm) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        iperm[perm[i]] = (int32_t)i;
}

// sort key of each element: its smallest solver-numbered dof
__global__ __launch_bounds__(kBlock) void kEoKeys(int64_t ne, const int32_t* __restrict__ e2d,
                                                  const int32_t* __restrict__ iperm, int32_t* __restrict__ key,
                                                  int32_t* __restrict__ id) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        int32_t m = INT32_MAX;
        for (int l = 0; l < kN; ++l) m = min(m, iperm[e2d[e * kN + l]]);
        key[e] = m;
        id[e] = (int32_t)e;
    }
}

// edof0[l][pos] and, for the incidence sort, (dof, Y index) pairs in Y order
__global__ __launch_bounds__(kBlock) void kEoDofs(int64_t ne, const int32_t* __restrict__ e2d,
                                                  const int32_t* __restrict__ iperm,
                                                  const int32_t* __restrict__ eorder, int32_t* __restrict__ edof0,
                                                  int32_t* __restrict__ ykey, int32_t* __restrict__ yval) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ne; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = eorder[p];
        for (int l = 0; l < kN; ++l) {
            const int32_t d = iperm[e2d[e * kN + l]];
            edof0[(size_t)l * ne + p] = d;
            ykey[(size_t)l * ne + p] = d;
            yval[(size_t)l * ne + p] = (int32_t)((size_t)l * ne + p);
        }
    }
}

__global__ __launch_bounds__(kBlock) void kEoCount(int64_t m, const int32_t* __restrict__ key,
                                                   uint32_t* __restrict__ cnt) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[key[k]], 1u);
}

// slots of each slice: 64 x the largest incidence count of its rows
__global__ __launch_bounds__(kBlock) void kEoSliceSlots(int64_t n, int64_t nslices, const uint32_t* __restrict__ cnt,
                                                        uint64_t* __restrict__ slots) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nslices; s += nw) {
        const int64_t r = s * 64 + lane;
        uint32_t c = r < n ? cnt[r] : 0u;
        for (int o = 32; o > 0; o >>= 1) c = max(c, (uint32_t)__shfl_xor((int)c, o));
        if (lane == 0) slots[s] = (uint64_t)c * 64;
    }
}

__global__ __launch_bounds__(kBlock) void kEoSlicePtr(int64_t m, const uint64_t* __restrict__ off,
                                                      uint32_t* __restrict__ ptr) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < m; s += (int64_t)gridDim.x * blockDim.x)
        ptr[s] = (uint32_t)off[s];
}

// inc[base + k 64 + lane] = the k-th Y index of row s 64 + lane (sorted order), or the zero entry
__global__ __launch_bounds__(kBlock) void kEoFill(int64_t n, int64_t nslices, int64_t ne,
                                                  const uint32_t* __restrict__ incPtr,
                                                  const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ start,
                                                  const int32_t* __restrict__ ysorted, int32_t* __restrict__ inc) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nslices * 64;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = r >> 6;
        const int lane = (int)(r & 63);
        const uint32_t base = incPtr[s];
        const int width = (int)((incPtr[s + 1] - base) >> 6);
        const int c = r < n ? (int)cnt[r] : 0;
        const uint32_t b0 = r < n ? start[r] : 0u;
        for (int k = 0; k < width; ++k)
            inc[base + (size_t)k * 64 + lane] = k < c ? ysorted[b0 + k] : (int32_t)(kN * ne);
    }
}

// F[a][pos] from the space's FP64 factors [C00 C11 C22 C01 C02 C12 | m | pad] (stride 8)
__global__ __launch_bounds__(kBlock) void kEoFactors(int64_t ne, const int32_t* __restrict__ eorder,
                                                     const double* __restrict__ geo, float* __restrict__ F) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ne; p += (int64_t)gridDim.x * blockDim.x) {
        const double* f = geo + (size_t)eorder[p] * 8;
#pragma unroll
        for (int a = 0; a < 6; ++a) F[(size_t)a * ne + p] = (float)f[a];
    }
}

__global__ __launch_bounds__(kBlock) void kEoOrFlags(int64_t n, const uint8_t* __restrict__ flag,
                                                     uint8_t* __restrict__ bold) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (flag[i]) bold[i] = 1;
}

__global__ __launch_bounds__(kBlock) void kEoRowFlags(int64_t n, const int32_t* __restrict__ perm,
                                                      const uint8_t* __restrict__ bold, uint8_t* __restrict__ brow) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        brow[i] = bold[perm[i]];
}

__global__ __launch_bounds__(kBlock) void kEoColFlags(int64_t m, const int32_t* __restrict__ edof0,
                                                      const uint8_t* __restrict__ brow, int32_t* __restrict__ edof) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t d = edof0[k];
        edof[k] = brow[d] ? (int32_t)((uint32_t)d | 0x80000000u) : d;
    }
}

// pass 1: Y[l][pos] = (A_e p_e)_l, Dirichlet columns excluded.  At each
// point of the rule: grad u = sum_v s_v grad l_v with
// s_v = p_v (4 l_v - 1) + 4 sum_{edges (v,o)} p_e l_o, then h = w F grad u and
// (A_e p)_i += grad phi_i . h, i.e. (4 l_v - 1) t_v for a vertex and
// 4 (l_a t_b + l_b t_a) for an edge, t_v = grad l_v . h.  The barycentric
// coordinates are constants, so no table is loaded.
__global__ __launch_bounds__(kBlock) void kElemApplyP2(int64_t ne, const float* __restrict__ F,
                                                       const int32_t* __restrict__ edof, const float* __restrict__ p,
                                                       float* __restrict__ Y, const CgState* __restrict__ st) {
    if (st && blockDone(st)) return;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        float pe[kN];
#pragma unroll
        for (int l = 0; l < kN; ++l) {
            const int32_t d = edof[(size_t)l * ne + e];
            pe[l] = d < 0 ? 0.0f : p[d];
        }
        const float f00 = F[e], f11 = F[(size_t)ne + e], f22 = F[2 * (size_t)ne + e];
        const float f01 = F[3 * (size_t)ne + e], f02 = F[4 * (size_t)ne + e], f12 = F[5 * (size_t)ne + e];
        float y[kN];
#pragma unroll
        for (int l = 0; l < kN; ++l) y[l] = 0.0f;
#pragma unroll
        for (int q = 0; q < kNq; ++q) {
            float lam[4], sv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                lam[v] = v == q ? kAlpha : kBeta;
                sv[v] = pe[v] * (4.0f * lam[v] - 1.0f);
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int a = kEdge[k][0], b = kEdge[k][1];
                sv[a] += (4.0f * lam[b]) * pe[4 + k];
                sv[b] += (4.0f * lam[a]) * pe[4 + k];
            }
            const float g0 = 0.5f * (sv[1] - sv[0]), g1 = 0.5f * (sv[2] - sv[0]), g2 = 0.5f * (sv[3] - sv[0]);
            const float h0 = kWq * (f00 * g0 + f01 * g1 + f02 * g2);
            const float h1 = kWq * (f01 * g0 + f11 * g1 + f12 * g2);
            const float h2 = kWq * (f02 * g0 + f12 * g1 + f22 * g2);
            const float t[4] = {-0.5f * (h0 + h1 + h2), 0.5f * h0, 0.5f * h1, 0.5f * h2};
#pragma unroll
            for (int v = 0; v < 4; ++v) y[v] += (4.0f * lam[v] - 1.0f) * t[v];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int a = kEdge[k][0], b = kEdge[k][1];
                y[4 + k] += 4.0f * (lam[a] * t[b] + lam[b] * t[a]);
            }
        }
#pragma unroll
        for (int l = 0; l < kN; ++l) {
            const int32_t d = edof[(size_t)l * ne + e];
            Y[(size_t)l * ne + e] = d < 0 ? 0.0f : y[l];
        }
    }
}

// pass 2: w_i = sum of row i's Y entries (or p_i on a Dirichlet row), and the
// block partial of p.w
template <int BS>
__global__ __launch_bounds__(BS) void kElemGather(int64_t n, int64_t nslices, const uint32_t* __restrict__ incPtr,
                                                  const int32_t* __restrict__ inc, const float* __restrict__ Y,
                                                  const uint8_t* __restrict__ brow, const float* __restrict__ p,
                                                  float* __restrict__ w, double* __restrict__ part,
                                                  const CgState* __restrict__ st) {
    if (st && blockDone(st)) return;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * BS + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * BS) >> 6;
    double dot = 0.0;
    for (int64_t s = wave; s < nslices; s += nwaves) {
        const uint32_t base = incPtr[s];
        const int width = (int)((incPtr[s + 1] - base) >> 6);
        const int32_t* c = inc + base + lane;
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
        int k = 0;
        for (; k + 4 <= width; k += 4) {
            const int32_t i0 = c[(size_t)k * 64], i1 = c[(size_t)(k + 1) * 64];
            const int32_t i2 = c[(size_t)(k + 2) * 64], i3 = c[(size_t)(k + 3) * 64];
            a0 += Y[i0];
            a1 += Y[i1];
            a2 += Y[i2];
            a3 += Y[i3];
        }
        for (; k < width; ++k) a0 += Y[c[(size_t)k * 64]];
        const int64_t r = s * 64 + lane;
        if (r < n) {
            const float pr = p[r];
            const float a = brow[r] ? pr : (a0 + a1) + (a2 + a3);
            w[r] = a;
            if (part) dot += (double)pr * (double)a;
        }
    }
    if (part) {
        const double t = blockReduceSumN<BS>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

__global__ void kEoZeroTail(float* __restrict__ y) {
    if (blockIdx.x == 0 && threadIdx.x == 0) y[0] = 0.0f;
}

constexpr int kGatherBlock = 256;
constexpr int kGatherGrid = 2048;

}  // namespace

namespace feelpp_amd {

bool elemOpCandidate(const feel_hip_mat* A) {
    const feel_hip_space* s = A->g->space;
    return s->dim == 3 && s->order == 2 && s->ncomp == 1 && !A->g->block && A->g->perm &&
           (A->precision & FEEL_HIP_FP32);
}

// Structure of the element operator from the space's elem2dof and the
// solver numbering of the graph (one-time cost per matrix: two radix sorts).
static int buildStructure(feel_hip_mat* A, ElemOp* E) {
    feel_hip_graph* g = A->g;
    feel_hip_space* s = g->space;
    feel_hip_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    const int64_t ne = s->ne, n = g->n, m = ne * kN;
    int rc = FEEL_HIP_OK;
    int32_t *iperm = nullptr, *key = nullptr, *key2 = nullptr, *id = nullptr, *ykey = nullptr, *ykey2 = nullptr,
            *yval = nullptr, *ysorted = nullptr;
    uint32_t *cnt = nullptr, *start = nullptr;
    uint64_t *slots = nullptr, *off = nullptr;
    void* tmp = nullptr;
    size_t need = 0, need2 = 0;
    uint64_t total = 0;
    if (m >= (int64_t)INT32_MAX) return setError(FEEL_HIP_EUNSUPPORTED, "element operator: too many element entries");
    E->ne = ne;
    E->n = n;
    E->nslices = (n + 63) / 64;
    EO_CHECK(hipMalloc(&iperm, sizeof(int32_t) * n));
    EO_CHECK(hipMalloc(&key, sizeof(int32_t) * ne));
    EO_CHECK(hipMalloc(&key2, sizeof(int32_t) * ne));
    EO_CHECK(hipMalloc(&id, sizeof(int32_t) * ne));
    rc = c->alloc((void**)&E->eorder, sizeof(int32_t) * ne);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&E->edof0, sizeof(int32_t) * m);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&E->edof, sizeof(int32_t) * m);
    if (rc != FEEL_HIP_OK) goto done;
    hipLaunchKernelGGL(kEoInverse, dim3(gridOf(n)), dim3(kBlock), 0, st, n, g->perm, iperm);
    hipLaunchKernelGGL(kEoKeys, dim3(gridOf(ne)), dim3(kBlock), 0, st, ne, s->e2d, iperm, key, id);
    EO_CHECK(hipGetLastError());
    EO_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, key, key2, id, E->eorder, (int)ne, 0, 32, st));
    EO_CHECK(hipMalloc(&ykey, sizeof(int32_t) * m));
    EO_CHECK(hipMalloc(&ykey2, sizeof(int32_t) * m));
    EO_CHECK(hipMalloc(&yval, sizeof(int32_t) * m));
    EO_CHECK(hipMalloc(&ysorted, sizeof(int32_t) * m));
    EO_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need2, ykey, ykey2, yval, ysorted, (int)m, 0, 32, st));
    need = std::max(need, need2);
    EO_CHECK(hipMalloc(&tmp, need));
    EO_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, need, key, key2, id, E->eorder, (int)ne, 0, 32, st));
    hipLaunchKernelGGL(kEoDofs, dim3(gridOf(ne)), dim3(kBlock), 0, st, ne, s->e2d, iperm, E->eorder, E->edof0, ykey,
                       yval);
    EO_CHECK(hipGetLastError());
    EO_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, need, ykey, ykey2, yval, ysorted, (int)m, 0, 32, st));
    // incidence counts, row starts, slice widths
    EO_CHECK(hipMalloc(&cnt, sizeof(uint32_t) * (n + 1)));
    EO_CHECK(hipMalloc(&start, sizeof(uint32_t) * (n + 1)));
    EO_CHECK(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * (n + 1), st));
    hipLaunchKernelGGL(kEoCount, dim3(gridOf(m)), dim3(kBlock), 0, st, m, ykey2, cnt);
    EO_CHECK(hipGetLastError());
    EO_CHECK(hipFree(tmp));
    tmp = nullptr;
    need = 0;
    EO_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, cnt, start, (int)(n + 1), st));
    {
        size_t needS = 0;
        EO_CHECK(hipMalloc(&slots, sizeof(uint64_t) * (E->nslices + 1)));
        EO_CHECK(hipMalloc(&off, sizeof(uint64_t) * (E->nslices + 1)));
        EO_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, needS, slots, off, (int)(E->nslices + 1), st));
        need = std::max(need, needS);
    }
    EO_CHECK(hipMalloc(&tmp, need));
    EO_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, need, cnt, start, (int)(n + 1), st));
    EO_CHECK(hipMemsetAsync(slots, 0, sizeof(uint64_t) * (E->nslices + 1), st));
    hipLaunchKernelGGL(kEoSliceSlots, dim3(gridOf(E->nslices * 64)), dim3(kBlock), 0, st, n, E->nslices, cnt, slots);
    EO_CHECK(hipGetLastError());
    EO_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, need, slots, off, (int)(E->nslices + 1), st));
    EO_CHECK(hipMemcpyAsync(&total, off + E->nslices, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    EO_CHECK(hipStreamSynchronize(st));
    if (total >= (1ULL << 32)) { rc = setError(FEEL_HIP_EUNSUPPORTED, "element operator: incidence image too large"); goto done; }
    E->nslots = (int64_t)total;
    rc = c->alloc((void**)&E->incPtr, sizeof(uint32_t) * (E->nslices + 1));
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&E->inc, sizeof(int32_t) * std::max<int64_t>(1, E->nslots));
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&E->F, sizeof(float) * 6 * ne);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&E->brow, n);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&E->Y, sizeof(float) * (m + 1));
    if (rc != FEEL_HIP_OK) goto done;
    hipLaunchKernelGGL(kEoSlicePtr, dim3(gridOf(E->nslices + 1)), dim3(kBlock), 0, st, E->nslices + 1, off, E->incPtr);
    hipLaunchKernelGGL(kEoFill, dim3(gridOf(E->nslices * 64)), dim3(kBlock), 0, st, n, E->nslices, ne, E->incPtr, cnt,
                       start, ysorted, E->inc);
    hipLaunchKernelGGL(kEoZeroTail, dim3(1), dim3(64), 0, st, E->Y + m);
    EO_CHECK(hipGetLastError());
    EO_CHECK(hipStreamSynchronize(st));
done:
    for (void* p : {(void*)iperm, (void*)key, (void*)key2, (void*)id, (void*)ykey, (void*)ykey2, (void*)yval,
                    (void*)ysorted, (void*)cnt, (void*)start, (void*)slots, (void*)off, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

// the operator's arrays (not the matrix's row set); sizes are zero for arrays
// a failed build did not reach
static void releaseOp(feel_hip_mat* A) {
    feel_hip_ctx* c = A->g->space->ctx;
    ElemOp* E = A->eop;
    if (!E) return;
    const int64_t m = E->ne * kN;
    c->release(E->edof, sizeof(int32_t) * m);
    c->release(E->edof0, sizeof(int32_t) * m);
    c->release(E->eorder, sizeof(int32_t) * E->ne);
    c->release(E->incPtr, sizeof(uint32_t) * (E->nslices + 1));
    c->release(E->inc, sizeof(int32_t) * std::max<int64_t>(1, E->nslots));
    c->release(E->F, sizeof(float) * 6 * E->ne);
    c->release(E->brow, E->n);
    c->release(E->Y, sizeof(float) * (m + 1));
    delete E;
    A->eop = nullptr;
}

void elemOpFree(feel_hip_mat* A) {
    A->g->space->ctx->release(A->mfRows, A->g->n);
    A->mfRows = nullptr;
    A->mfState = 0;
    releaseOp(A);
}

void elemOpInvalidate(feel_hip_mat* A) { A->mfState = 0; }

// the factors of the matrix's last assembly, still in the space's scratch, into F
static int snapshotFactors(feel_hip_mat* A) {
    feel_hip_space* s = A->g->space;
    ElemOp* E = A->eop;
    hipLaunchKernelGGL(kEoFactors, dim3(gridOf(E->ne)), dim3(kBlock), 0, s->ctx->stream, E->ne, E->eorder, s->geo,
                       E->F);
    FEEL_HIP_CHECK(hipGetLastError());
    E->factorsValid = true;
    return FEEL_HIP_OK;
}

// After a bilinear assembly: the operator stays available when the values are
// a P2 diffusion assembly (qualifies).  Nothing is copied until an
// element-operator solve asks for it; from then on each qualifying assembly
// of the matrix copies its factors.
int elemOpAssembled(feel_hip_mat* A, bool qualifies) {
    feel_hip_space* s = A->g->space;
    A->mfState = qualifies && elemOpCandidate(A) ? 1 : 0;
    A->geoStamp = s->geoStamp;
    if (A->mfRows) FEEL_HIP_CHECK(hipMemsetAsync(A->mfRows, 0, A->g->n, s->ctx->stream));
    if (A->eop) {
        A->eop->factorsValid = false;
        A->eop->flagsDirty = true;
        if (A->mfState == 1) return snapshotFactors(A);
    }
    return FEEL_HIP_OK;
}

int elemOpZeroRows(feel_hip_mat* A, const feel_hip_rows* rows, int mode, double diag) {
    if (A->mfState != 1) return FEEL_HIP_OK;
    if (mode != FEEL_HIP_ON_ELIMINATION_SYMMETRIC || diag != 1.0) {
        A->mfState = 0;
        return FEEL_HIP_OK;
    }
    feel_hip_ctx* c = A->g->space->ctx;
    const int64_t n = A->g->n;
    if (!A->mfRows) {
        if (c->alloc((void**)&A->mfRows, n) != FEEL_HIP_OK) {   // without the row set the operator is unavailable
            A->mfState = 0;
            return FEEL_HIP_OK;
        }
        FEEL_HIP_CHECK(hipMemsetAsync(A->mfRows, 0, n, c->stream));
    }
    hipLaunchKernelGGL(kEoOrFlags, dim3(gridOf(n)), dim3(kBlock), 0, c->stream, n, rows->flag, A->mfRows);
    FEEL_HIP_CHECK(hipGetLastError());
    if (A->eop) A->eop->flagsDirty = true;
    return FEEL_HIP_OK;
}

bool elemOpReady(const feel_hip_mat* A) { return A->mfState == 1; }

// Before an element-operator solve: builds the structure on first use and
// copies the factors if the space's scratch still holds this matrix's last
// assembly (no other assembly on the space since); otherwise the matrix must
// be assembled again.
int elemOpPrepare(feel_hip_mat* A) {
    feel_hip_space* s = A->g->space;
    feel_hip_ctx* c = s->ctx;
    if (A->mfState != 1)
        return setError(FEEL_HIP_EUNSUPPORTED,
                        "element-by-element inner operator needs a P2 diffusion assembly (symmetric on() only)");
    if (!A->eop) {
        A->eop = new ElemOp();
        int rc = buildStructure(A, A->eop);
        if (rc != FEEL_HIP_OK) {
            releaseOp(A);   // buildStructure frees its temporaries, this the arrays it kept
            return rc;
        }
    }
    if (!A->eop->factorsValid) {
        if (s->geoStamp != A->geoStamp)
            return setError(FEEL_HIP_EUNSUPPORTED,
                            "element-by-element inner operator: another assembly on this space replaced the matrix's "
                            "element factors; assemble the matrix again");
        TRY_CALL_EO(snapshotFactors(A));
    }
    if (!A->mfRows) {   // no on() since the assembly: no Dirichlet rows
        TRY_CALL_EO(c->alloc((void**)&A->mfRows, A->g->n));
        FEEL_HIP_CHECK(hipMemsetAsync(A->mfRows, 0, A->g->n, c->stream));
        A->eop->flagsDirty = true;
    }
    return FEEL_HIP_OK;
}

// w = A p in the solver numbering (FP32), partials of p.w in part[0..grid)
int elemOpApply(feel_hip_mat* A, const float* p, float* w, double* part, const CgState* st, int* npart) {
    ElemOp* E = A->eop;
    feel_hip_ctx* c = A->g->space->ctx;
    if (E->flagsDirty) {
        hipLaunchKernelGGL(kEoRowFlags, dim3(gridOf(E->n)), dim3(kBlock), 0, c->stream, E->n, A->g->perm, A->mfRows,
                           E->brow);
        hipLaunchKernelGGL(kEoColFlags, dim3(gridOf(E->ne * kN)), dim3(kBlock), 0, c->stream, E->ne * kN, E->edof0,
                           E->brow, E->edof);
        FEEL_HIP_CHECK(hipGetLastError());
        E->flagsDirty = false;
    }
    const int g1 = (int)std::min<int64_t>((E->ne + kBlock - 1) / kBlock, 16384);
    hipLaunchKernelGGL(kElemApplyP2, dim3(g1), dim3(kBlock), 0, c->stream, E->ne, E->F, E->edof, p, E->Y, st);
    hipLaunchKernelGGL(kElemGather<kGatherBlock>, dim3(kGatherGrid), dim3(kGatherBlock), 0, c->stream, E->n,
                       E->nslices, E->incPtr, E->inc, E->Y, E->brow, p, w, part, st);
    FEEL_HIP_CHECK(hipGetLastError());
    if (npart) *npart = kGatherGrid;
    return FEEL_HIP_OK;
}

}  // namespace feelpp_amd
