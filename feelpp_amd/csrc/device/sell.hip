// Device-side construction of the solver's SELL-C-sigma image (C = 64 rows
// per slice) from the graph's device ia/ja and a row order computed on the
// host (reverse Cuthill-McKee, then rows sorted by length inside windows of
// sigma rows).  Per slice: width = longest row rounded up to a multiple of 4;
// the lane's k-th entry (permuted column, CSR position) sits at sellSlot(base,
// k, lane) (quads of 4 consecutive entries per lane) or, past the row's end,
// a padding entry pointing at the row itself with no CSR position.  The packed
// column stream stores a slice whose columns span fewer than 2^16 as 16-bit
// offsets from its smallest column (colBase), any other slice as int32 pairs
// of uint16 (colBase = -1); it is kept when at least a quarter of the slots
// are narrow.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

using namespace feelpp_amd;

namespace {

constexpr int kSlBlock = 256;
constexpr int kC = 64;

int slGrid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kSlBlock - 1) / kSlBlock, 65536)); }

__global__ __launch_bounds__(kSlBlock) void kInversePerm(int64_t n, const int32_t* __restrict__ perm,
                                                        int32_t* __restrict__ iperm) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        iperm[perm[r]] = (int32_t)r;
}

// slots of every slice (width x 64), one wavefront per slice
__global__ __launch_bounds__(kSlBlock) void kSliceSlots(int64_t n, int64_t nslices, const int32_t* __restrict__ ia,
                                                       const int32_t* __restrict__ perm, uint64_t* __restrict__ slots) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nslices; s += nw) {
        const int64_t r = s * kC + lane;
        int len = 0;
        if (r < n) len = ia[perm[r] + 1] - ia[perm[r]];
        for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o));
        len = (len + kSellQuad - 1) / kSellQuad * kSellQuad;
        if (lane == 0) slots[s] = (uint64_t)len * kC;
    }
}

__global__ __launch_bounds__(kSlBlock) void kSlicePtr(int64_t m, const uint64_t* __restrict__ off,
                                                     uint32_t* __restrict__ ptr) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < m; s += (int64_t)gridDim.x * blockDim.x)
        ptr[s] = (uint32_t)off[s];
}

// one thread per (slice, lane) row of the image
__global__ __launch_bounds__(kSlBlock) void kSellFill(int64_t n, int64_t nrowsPadded, const int32_t* __restrict__ ia,
                                                     const int32_t* __restrict__ ja, const int32_t* __restrict__ perm,
                                                     const int32_t* __restrict__ iperm,
                                                     const uint32_t* __restrict__ slicePtr, int32_t* __restrict__ col,
                                                     int32_t* __restrict__ slot2csr, int32_t* __restrict__ diagSlot) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrowsPadded;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = r / kC;
        const int lane = (int)(r - s * kC);
        const uint32_t b = slicePtr[s];
        const int width = (int)((slicePtr[s + 1] - b) / kC);
        const int64_t self = r < n ? r : n - 1;   // padding rows point inside the slice's range
        const int32_t old = r < n ? perm[r] : 0;
        const int32_t p0 = r < n ? ia[old] : 0;
        const int len = r < n ? ia[old + 1] - p0 : 0;
        for (int k = 0; k < width; ++k) {
            const int64_t slot = (int64_t)sellSlot(b, k, lane);
            if (k < len) {
                const int32_t c = ja[p0 + k];
                col[slot] = iperm[c];
                slot2csr[slot] = p0 + k;
                if (c == old) diagSlot[r] = (int32_t)slot;
            } else {
                col[slot] = (int32_t)self;
                slot2csr[slot] = -1;
            }
        }
    }
}

// per slice: colBase (smallest column, or -1 when the span needs 32 bits),
// packed units and narrow slot count; one wavefront per slice
__global__ __launch_bounds__(kSlBlock) void kSliceSpan(int64_t nslices, const uint32_t* __restrict__ slicePtr,
                                                      const int32_t* __restrict__ col, int32_t* __restrict__ colBase,
                                                      int64_t* __restrict__ units, int64_t* __restrict__ narrow) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nslices; s += nw) {
        const uint32_t b = slicePtr[s], e = slicePtr[s + 1];
        int32_t lo = INT32_MAX, hi = INT32_MIN;
        for (uint32_t q = b + lane; q < e; q += kC) { lo = min(lo, col[q]); hi = max(hi, col[q]); }
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o));
            hi = max(hi, __shfl_xor(hi, o));
        }
        if (lane == 0) {
            const int64_t w = (int64_t)e - b;
            const bool nar = w == 0 || (int64_t)hi - lo < 65536;
            colBase[s] = w == 0 ? 0 : (nar ? lo : -1);
            units[s] = nar ? w : 2 * w;
            narrow[s] = nar ? w : 0;
        }
    }
}

__global__ __launch_bounds__(kSlBlock) void kPackFill(int64_t nslices, const uint32_t* __restrict__ slicePtr,
                                                     const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ colBase,
                                                     const int64_t* __restrict__ cptr, uint16_t* __restrict__ packed) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nslices; s += nw) {
        const uint32_t b = slicePtr[s], e = slicePtr[s + 1];
        uint16_t* out = packed + cptr[s];
        const int32_t cb = colBase[s];
        if (cb >= 0) {
            for (uint32_t q = b + lane; q < e; q += kC) out[q - b] = (uint16_t)(col[q] - cb);
        } else {
            int32_t* out32 = reinterpret_cast<int32_t*>(out);   // cptr is a multiple of 64 units
            for (uint32_t q = b + lane; q < e; q += kC) out32[q - b] = col[q];
        }
    }
}

#define SL_CHECK(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) { rc = setError(FEEL_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); goto done; } \
    } while (0)

}  // namespace

namespace feelpp_amd {

// Fills g->perm, slicePtr, scol, slot2csr, sdiag, nslices, nslots and, when
// worth it, the packed column stream, from g->ia/g->ja (device) and the host
// row order perm (new -> old).
int buildSellDevice(feel_hip_graph* g, const std::vector<int32_t>& perm) {
    feel_hip_ctx* c = g->space->ctx;
    hipStream_t st = c->stream;
    const int64_t n = g->n;
    const int64_t nslices = (n + kC - 1) / kC;
    int rc = FEEL_HIP_OK;
    int32_t* iperm = nullptr;
    uint64_t *cnt = nullptr, *off = nullptr;
    int64_t *units = nullptr, *uoff = nullptr, *narrow = nullptr, *sums = nullptr;
    void* tmp = nullptr;
    size_t need = 0, tmpBytes = 0;
    uint64_t total = 0;
    int64_t hs[2] = {0, 0};
    g->nslices = nslices;
    rc = c->alloc((void**)&g->perm, sizeof(int32_t) * n);
    if (rc != FEEL_HIP_OK) return rc;
    SL_CHECK(hipMemcpyAsync(g->perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    SL_CHECK(hipMalloc(&iperm, sizeof(int32_t) * n));
    hipLaunchKernelGGL(kInversePerm, dim3(slGrid(n)), dim3(kSlBlock), 0, st, n, g->perm, iperm);
    // slice offsets: exclusive scan of the slice slot counts (64-bit, checked below 2^32)
    SL_CHECK(hipMalloc(&cnt, sizeof(uint64_t) * (nslices + 1)));
    SL_CHECK(hipMalloc(&off, sizeof(uint64_t) * (nslices + 1)));
    SL_CHECK(hipMemsetAsync(cnt, 0, sizeof(uint64_t) * (nslices + 1), st));
    hipLaunchKernelGGL(kSliceSlots, dim3(slGrid(nslices * kC)), dim3(kSlBlock), 0, st, n, nslices, g->ia, g->perm, cnt);
    SL_CHECK(hipGetLastError());
    SL_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, cnt, off, (int)(nslices + 1), st));
    tmpBytes = need;
    SL_CHECK(hipMalloc(&tmp, tmpBytes));
    SL_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmpBytes, cnt, off, (int)(nslices + 1), st));
    SL_CHECK(hipMemcpyAsync(&total, off + nslices, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    SL_CHECK(hipStreamSynchronize(st));
    if (total >= (1ULL << 32)) { rc = setError(FEEL_HIP_EUNSUPPORTED, "SELL image exceeds 2^32 slots"); goto done; }
    g->nslots = (int64_t)total;
    rc = c->alloc((void**)&g->slicePtr, sizeof(uint32_t) * (nslices + 1));
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->scol, sizeof(int32_t) * g->nslots);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->slot2csr, sizeof(int32_t) * g->nslots);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->sdiag, sizeof(int32_t) * n);
    if (rc != FEEL_HIP_OK) goto done;
    hipLaunchKernelGGL(kSlicePtr, dim3(slGrid(nslices + 1)), dim3(kSlBlock), 0, st, nslices + 1, off, g->slicePtr);
    hipLaunchKernelGGL(kSellFill, dim3(slGrid(nslices * kC)), dim3(kSlBlock), 0, st, n, nslices * kC, g->ia, g->ja,
                       g->perm, iperm, g->slicePtr, g->scol, g->slot2csr, g->sdiag);
    SL_CHECK(hipGetLastError());
    // packed column stream
    SL_CHECK(hipMalloc(&units, sizeof(int64_t) * (nslices + 1)));
    SL_CHECK(hipMalloc(&uoff, sizeof(int64_t) * (nslices + 1)));
    SL_CHECK(hipMalloc(&narrow, sizeof(int64_t) * nslices));
    SL_CHECK(hipMalloc(&sums, sizeof(int64_t)));
    SL_CHECK(hipMemsetAsync(units + nslices, 0, sizeof(int64_t), st));
    rc = c->alloc((void**)&g->colBase, sizeof(int32_t) * nslices);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->ccptr, sizeof(int64_t) * nslices);
    if (rc != FEEL_HIP_OK) goto done;
    hipLaunchKernelGGL(kSliceSpan, dim3(slGrid(nslices * kC)), dim3(kSlBlock), 0, st, nslices, g->slicePtr, g->scol,
                       g->colBase, units, narrow);
    SL_CHECK(hipGetLastError());
    need = 0;
    SL_CHECK(hipcub::DeviceReduce::Sum(nullptr, need, narrow, sums, (int)nslices, st));
    if (need > tmpBytes) {
        SL_CHECK(hipFree(tmp));
        tmp = nullptr;
        tmpBytes = need;
        SL_CHECK(hipMalloc(&tmp, tmpBytes));
    }
    SL_CHECK(hipcub::DeviceReduce::Sum(tmp, tmpBytes, narrow, sums, (int)nslices, st));
    need = 0;
    SL_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, units, uoff, (int)(nslices + 1), st));
    if (need > tmpBytes) {
        SL_CHECK(hipFree(tmp));
        tmp = nullptr;
        tmpBytes = need;
        SL_CHECK(hipMalloc(&tmp, tmpBytes));
    }
    SL_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmpBytes, units, uoff, (int)(nslices + 1), st));
    SL_CHECK(hipMemcpyAsync(&hs[0], sums, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SL_CHECK(hipMemcpyAsync(&hs[1], uoff + nslices, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SL_CHECK(hipMemcpyAsync(g->ccptr, uoff, sizeof(int64_t) * nslices, hipMemcpyDeviceToDevice, st));
    SL_CHECK(hipStreamSynchronize(st));
    // below a quarter of narrow slots the packed stream saves too little to pay
    // for the second copy of the columns
    if (g->nslots > 0 && (double)hs[0] / (double)g->nslots >= 0.25) {
        g->npacked = hs[1];
        rc = c->alloc((void**)&g->packed, sizeof(uint16_t) * g->npacked);
        if (rc != FEEL_HIP_OK) goto done;
        g->compact = true;
        hipLaunchKernelGGL(kPackFill, dim3(slGrid(nslices * kC)), dim3(kSlBlock), 0, st, nslices, g->slicePtr, g->scol,
                           g->colBase, g->ccptr, g->packed);
        SL_CHECK(hipGetLastError());
    } else {
        c->release(g->colBase, sizeof(int32_t) * nslices);
        c->release(g->ccptr, sizeof(int64_t) * nslices);
        g->colBase = nullptr;
        g->ccptr = nullptr;
    }
    SL_CHECK(hipStreamSynchronize(st));
done:
    for (void* p : {(void*)iperm, (void*)cnt, (void*)off, (void*)units, (void*)uoff, (void*)narrow, (void*)sums, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace feelpp_amd

// ------------------------------------------------------------------------
// 3 x 3 block image for vector spaces (ncomp = 3, dofs interleaved node * 3 +
// c, COUPLED pattern: the rows of a node share their columns, which come in
// triples).  Slice s holds the node rows sliceRow[s] .. sliceRow[s+1] - 1 with
// L = 2^sliceShift[s] lanes per row (64 / L rows): the row's k-th block goes to
// step k / L and lane L (row - sliceRow[s]) + k % L, so a slice of S steps
// holds 64 S block slots; the block slot's permuted node column sits at
// base + 64 step + lane and its nine values in planes, value (c, d) at
// 9 base + (9 step + 3 c + d) 64 + lane.  One column index serves nine values,
// the x gathers read three consecutive entries, and the L lanes of a row add
// their sums at the end.  Rows are ordered longest first inside windows, and a
// slice whose longest row would take more than kBlockStepCap steps with two
// lanes takes 4, 8 or 16, so no wavefront streams far longer than the rest
// (P4 tetrahedra: rows of 35 to 369 blocks).

namespace {

// slots of slice s: 64 x the steps of its longest row, a row of len blocks
// taking ceil(len / L) steps
__global__ __launch_bounds__(kSlBlock) void kBlockSliceSlots(int64_t nslices, const int32_t* __restrict__ sliceRow,
                                                            const uint8_t* __restrict__ sliceShift,
                                                            const int32_t* __restrict__ ia,
                                                            const int32_t* __restrict__ nperm,
                                                            uint64_t* __restrict__ slots) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nslices; s += nw) {
        const int sh = sliceShift[s], L = 1 << sh;
        const int64_t r = (int64_t)sliceRow[s] + (lane >> sh);
        int len = 0;
        if ((lane & (L - 1)) == 0 && r < sliceRow[s + 1])
            len = (ia[3 * (int64_t)nperm[r] + 1] - ia[3 * (int64_t)nperm[r]]) / 3;
        for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o));
        if (lane == 0) slots[s] = (uint64_t)((len + L - 1) / L) * kC;
    }
}

// One wavefront per slice, one thread per lane of the slice: lane L i + j
// fills the blocks k = m L + j (m = 0 .. steps - 1) of the slice's row i.
__global__ __launch_bounds__(kSlBlock) void kBlockSellFill(int64_t nslices, const int32_t* __restrict__ sliceRow,
                                                          const uint8_t* __restrict__ sliceShift,
                                                          const int32_t* __restrict__ ia,
                                                          const int32_t* __restrict__ ja,
                                                          const int32_t* __restrict__ nperm,
                                                          const int32_t* __restrict__ niperm,
                                                          const uint32_t* __restrict__ slicePtr,
                                                          int32_t* __restrict__ bcol, int32_t* __restrict__ slot2csr,
                                                          int32_t* __restrict__ diagSlot) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nslices; s += nw) {
        const int sh = sliceShift[s], L = 1 << sh, j = lane & (L - 1);
        const int64_t r0 = sliceRow[s], r1 = sliceRow[s + 1];
        const int64_t r = r0 + (lane >> sh);
        const bool real = r < r1;
        const uint32_t b = slicePtr[s];
        const int steps = (int)((slicePtr[s + 1] - b) / kC);
        const int32_t self = (int32_t)(real ? r : r1 - 1);   // padding reads a row of this slice
        const int32_t old = real ? nperm[r] : 0;
        const int len = real ? (ia[3 * (int64_t)old + 1] - ia[3 * (int64_t)old]) / 3 : 0;
        for (int m = 0; m < steps; ++m) {
            const int k = m * L + j;
            const int64_t bs = b + (int64_t)m * kC + lane;
            const size_t v0 = 9 * (size_t)b + (size_t)m * 9 * kC + lane;
            if (k < len) {
                const int32_t bj = ja[ia[3 * (int64_t)old] + 3 * k] / 3;
                bcol[bs] = niperm[bj];
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d) {
                        const int32_t p = ia[3 * (int64_t)old + c] + 3 * k + d;
                        const size_t vs = v0 + (size_t)(3 * c + d) * kC;
                        slot2csr[vs] = p;
                        if (bj == old && c == d) diagSlot[3 * r + c] = (int32_t)vs;
                    }
            } else {
                bcol[bs] = self;
                for (int q = 0; q < 9; ++q) slot2csr[v0 + (size_t)q * kC] = -1;
            }
        }
    }
}

__global__ __launch_bounds__(kSlBlock) void kExpandPerm(int64_t nb, const int32_t* __restrict__ nperm,
                                                       int32_t* __restrict__ perm) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nb; r += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) perm[3 * r + c] = 3 * nperm[r] + c;
}

}  // namespace

namespace feelpp_amd {

// Slices of the block image over the node order nperm: each takes the next
// 64 / L rows, L = 2 unless the longest of them would take more than
// kBlockStepCap steps (then 4, 8, 16).  Fills the first row of each slice
// (and nb at the end) and log2 L per slice.
static void blockSlices(const std::vector<int32_t>& nperm, const std::vector<int32_t>& nodeLen, std::vector<int32_t>& row0,
                 std::vector<uint8_t>& shift) {
    const int64_t nb = (int64_t)nperm.size();
    row0.clear();
    shift.clear();
    for (int64_t r = 0; r < nb;) {
        int sh = kBlockMinShift;
        for (;; ++sh) {
            const int L = 1 << sh;
            const int64_t r1 = std::min<int64_t>(nb, r + kC / L);
            int maxlen = 0;
            for (int64_t i = r; i < r1; ++i) maxlen = std::max(maxlen, nodeLen[nperm[i]]);
            if ((maxlen + L - 1) / L <= kBlockStepCap || sh == kBlockMaxShift) break;
        }
        row0.push_back((int32_t)r);
        shift.push_back((uint8_t)sh);
        r += kC >> sh;
    }
    row0.push_back((int32_t)nb);
}

// Fills the block image (g->block, perm, sliceRow/sliceShift, slicePtr in
// block slots, scol = block columns, slot2csr and sdiag over value slots,
// nslots = value slots) from the device ia/ja, the host node order nperm
// (new -> old node) and the node row lengths in blocks (old numbering).
int buildBlockSellDevice(feel_hip_graph* g, const std::vector<int32_t>& nperm, const std::vector<int32_t>& nodeLen) {
    feel_hip_ctx* c = g->space->ctx;
    hipStream_t st = c->stream;
    const int64_t nb = g->n / 3;
    std::vector<int32_t> row0;
    std::vector<uint8_t> shift;
    blockSlices(nperm, nodeLen, row0, shift);
    const int64_t nslices = (int64_t)shift.size();
    int rc = FEEL_HIP_OK;
    int32_t *dnperm = nullptr, *niperm = nullptr;
    uint64_t *cnt = nullptr, *off = nullptr;
    void* tmp = nullptr;
    size_t need = 0;
    uint64_t total = 0;
    g->block = true;
    g->nslices = nslices;
    rc = c->alloc((void**)&g->perm, sizeof(int32_t) * g->n);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->sliceRow, sizeof(int32_t) * (nslices + 1));
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->sliceShift, sizeof(uint8_t) * nslices);
    if (rc != FEEL_HIP_OK) return rc;
    SL_CHECK(hipMemcpyAsync(g->sliceRow, row0.data(), sizeof(int32_t) * (nslices + 1), hipMemcpyHostToDevice, st));
    SL_CHECK(hipMemcpyAsync(g->sliceShift, shift.data(), sizeof(uint8_t) * nslices, hipMemcpyHostToDevice, st));
    SL_CHECK(hipMalloc(&dnperm, sizeof(int32_t) * nb));
    SL_CHECK(hipMemcpyAsync(dnperm, nperm.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
    SL_CHECK(hipMalloc(&niperm, sizeof(int32_t) * nb));
    hipLaunchKernelGGL(kInversePerm, dim3(slGrid(nb)), dim3(kSlBlock), 0, st, nb, dnperm, niperm);
    hipLaunchKernelGGL(kExpandPerm, dim3(slGrid(nb)), dim3(kSlBlock), 0, st, nb, dnperm, g->perm);
    SL_CHECK(hipMalloc(&cnt, sizeof(uint64_t) * (nslices + 1)));
    SL_CHECK(hipMalloc(&off, sizeof(uint64_t) * (nslices + 1)));
    SL_CHECK(hipMemsetAsync(cnt, 0, sizeof(uint64_t) * (nslices + 1), st));
    hipLaunchKernelGGL(kBlockSliceSlots, dim3(slGrid(nslices * kC)), dim3(kSlBlock), 0, st, nslices, g->sliceRow,
                       g->sliceShift, g->ia, dnperm, cnt);
    SL_CHECK(hipGetLastError());
    SL_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, cnt, off, (int)(nslices + 1), st));
    SL_CHECK(hipMalloc(&tmp, need));
    SL_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, need, cnt, off, (int)(nslices + 1), st));
    SL_CHECK(hipMemcpyAsync(&total, off + nslices, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    SL_CHECK(hipStreamSynchronize(st));
    if (9 * total >= (1ULL << 31)) { rc = setError(FEEL_HIP_EUNSUPPORTED, "block SELL image exceeds 2^31 values"); goto done; }
    g->nbslots = (int64_t)total;
    g->nslots = 9 * (int64_t)total;
    rc = c->alloc((void**)&g->slicePtr, sizeof(uint32_t) * (nslices + 1));
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->scol, sizeof(int32_t) * g->nbslots);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->slot2csr, sizeof(int32_t) * g->nslots);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->sdiag, sizeof(int32_t) * g->n);
    if (rc != FEEL_HIP_OK) goto done;
    hipLaunchKernelGGL(kSlicePtr, dim3(slGrid(nslices + 1)), dim3(kSlBlock), 0, st, nslices + 1, off, g->slicePtr);
    hipLaunchKernelGGL(kBlockSellFill, dim3(slGrid(nslices * kC)), dim3(kSlBlock), 0, st, nslices, g->sliceRow,
                       g->sliceShift, g->ia, g->ja, dnperm, niperm, g->slicePtr, g->scol, g->slot2csr, g->sdiag);
    SL_CHECK(hipGetLastError());
    SL_CHECK(hipStreamSynchronize(st));
done:
    for (void* p : {(void*)dnperm, (void*)niperm, (void*)cnt, (void*)off, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace feelpp_amd
