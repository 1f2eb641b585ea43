// Device-side construction of the scatter's gather map for a (space, graph)
// pair: for every element-matrix entry k = e*nldof^2 + i*nldof + j the CSR
// slot of (elem2dof[e][i], elem2dof[e][j]) in Feel++'s graph (binary search
// in the row), a stable radix sort of the entries by slot (so each slot lists
// its entries in ascending element order, the order MatSetValues ADD applies
// them), end-of-slot flags in bit 31, the segment -> slot map when some slot
// receives no entry, and the scatter tasks (runs of whole slots, a cut after
// the first slot end crossing each multiple of kSegTask entries).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "common.hpp"

using namespace feelpp_amd;

namespace {

constexpr int kGmBlock = 256;

int gmGrid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kGmBlock - 1) / kGmBlock, 65536)); }

// slot of every element entry; nnz when the entry lies outside the pattern
__global__ __launch_bounds__(kGmBlock) void kEntrySlot(int64_t tot, int nl, const int32_t* __restrict__ e2d,
                                                      const int32_t* __restrict__ ia, const int32_t* __restrict__ ja,
                                                      uint32_t nnz, uint32_t* __restrict__ key,
                                                      uint32_t* __restrict__ val) {
    const int64_t nn = (int64_t)nl * nl;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < tot; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = k / nn;
        const int ij = (int)(k - e * nn);
        const int i = ij / nl, j = ij - i * nl;
        const int32_t row = e2d[e * nl + i], col = e2d[e * nl + j];
        int32_t lo = ia[row], hi = ia[row + 1];
        while (lo < hi) {
            const int32_t mid = lo + ((hi - lo) >> 1);
            if (ja[mid] < col) lo = mid + 1;
            else hi = mid;
        }
        key[k] = (lo < ia[row + 1] && ja[lo] == col) ? (uint32_t)lo : nnz;
        val[k] = (uint32_t)k;
    }
}

// end flags (bit 31 of the entry, and 0/1 for the segment scan)
__global__ __launch_bounds__(kGmBlock) void kMarkEnds(int64_t tot, const uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ entry, uint32_t* __restrict__ endFlag) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const bool end = q == tot - 1 || key[q + 1] != key[q];
        if (end) entry[q] |= 0x80000000u;
        endFlag[q] = end ? 1u : 0u;
    }
}

// per segment (slot with entries): its last entry position and its slot
__global__ __launch_bounds__(kGmBlock) void kSegments(int64_t tot, const uint32_t* __restrict__ key,
                                                     const uint32_t* __restrict__ endFlag,
                                                     const uint32_t* __restrict__ segIdx, uint32_t* __restrict__ segEnd,
                                                     int32_t* __restrict__ segSlot) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        if (!endFlag[q]) continue;
        const uint32_t s = segIdx[q];
        segEnd[s] = (uint32_t)q;
        if (segSlot) segSlot[s] = (int32_t)key[q];
    }
}

// cut after segment s when a multiple of kSegTask entries lies in
// (end of segment s-1 + 1, end of segment s + 1], and after the last segment
__global__ __launch_bounds__(kGmBlock) void kCuts(int64_t nseg, const uint32_t* __restrict__ segEnd,
                                                 uint32_t* __restrict__ cut) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = s > 0 ? (int64_t)segEnd[s - 1] + 1 : 0, b = (int64_t)segEnd[s] + 1;
        cut[s] = (s == nseg - 1 || b / kSegTask > a / kSegTask) ? 1u : 0u;
    }
}

// task t ends after the t-th cut segment: taskSeg[t+1] = s + 1, taskK[t+1] = end + 1
__global__ __launch_bounds__(kGmBlock) void kTasks(int64_t nseg, const uint32_t* __restrict__ segEnd,
                                                  const uint32_t* __restrict__ cut, const uint32_t* __restrict__ cutIdx,
                                                  uint32_t* __restrict__ taskSeg, uint32_t* __restrict__ taskK) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (int64_t)gridDim.x * blockDim.x) {
        if (s == 0) { taskSeg[0] = 0; taskK[0] = 0; }
        if (!cut[s]) continue;
        const uint32_t t = cutIdx[s] + 1;
        taskSeg[t] = (uint32_t)(s + 1);
        taskK[t] = segEnd[s] + 1;
    }
}

#define GM_CHECK(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) { rc = setError(FEEL_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); goto done; } \
    } while (0)

}  // namespace

namespace feelpp_amd {

// Fills g->centry, g->ncontrib, g->taskSeg, g->taskK, g->ntasks, g->segSlot,
// g->nseg from the space's element-to-dof map and the graph's ia/ja (device).
int buildGatherMapDevice(feel_hip_space* s, feel_hip_graph* g) {
    feel_hip_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    const int nl = s->nldof;
    const int64_t tot = s->ne * (int64_t)nl * nl;
    if (tot >= (1LL << 31)) return setError(FEEL_HIP_EUNSUPPORTED, "space too large for 31-bit element-entry ids");
    if (g->nnz >= (1LL << 31)) return setError(FEEL_HIP_EUNSUPPORTED, "graph too large");
    const uint32_t nnz = (uint32_t)g->nnz;
    int bits = 1;
    while (bits < 32 && (1ull << bits) <= nnz) ++bits;
    int rc = FEEL_HIP_OK;
    uint32_t *keyA = nullptr, *keyB = nullptr, *valA = nullptr, *valB = nullptr, *flag = nullptr, *idx = nullptr;
    uint32_t *segEnd = nullptr, *cut = nullptr;
    void* tmp = nullptr;
    size_t tmpBytes = 0, need = 0;
    uint32_t lastKey = 0, hv[2] = {0, 0};
    int64_t nseg = 0, ntasks = 0;
    hipcub::DoubleBuffer<uint32_t> keys, vals;
    GM_CHECK(hipMalloc(&keyA, sizeof(uint32_t) * tot));
    GM_CHECK(hipMalloc(&keyB, sizeof(uint32_t) * tot));
    GM_CHECK(hipMalloc(&valA, sizeof(uint32_t) * tot));
    GM_CHECK(hipMalloc(&valB, sizeof(uint32_t) * tot));
    hipLaunchKernelGGL(kEntrySlot, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, nl, s->e2d, g->ia, g->ja, nnz, keyA,
                       valA);
    GM_CHECK(hipGetLastError());
    keys = hipcub::DoubleBuffer<uint32_t>(keyA, keyB);
    vals = hipcub::DoubleBuffer<uint32_t>(valA, valB);
    // radix sort is stable: equal slots keep ascending entry (element) order
    GM_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys, vals, (int)tot, 0, bits, st));
    tmpBytes = need;
    GM_CHECK(hipMalloc(&tmp, tmpBytes));
    GM_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmpBytes, keys, vals, (int)tot, 0, bits, st));
    GM_CHECK(hipMemcpyAsync(&lastKey, keys.Current() + (tot - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GM_CHECK(hipStreamSynchronize(st));
    if (lastKey >= nnz) { rc = setError(FEEL_HIP_EPATTERN, "element entry outside the sparsity graph"); goto done; }
    // the sorted values become the gather map entries; the other value buffer holds the end flags
    flag = vals.Alternate();
    hipLaunchKernelGGL(kMarkEnds, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, keys.Current(), vals.Current(), flag);
    GM_CHECK(hipGetLastError());
    idx = keys.Alternate();
    need = 0;
    GM_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, flag, idx, (int)tot, st));
    if (need > tmpBytes) {
        GM_CHECK(hipFree(tmp));
        tmp = nullptr;
        tmpBytes = need;
        GM_CHECK(hipMalloc(&tmp, tmpBytes));
    }
    GM_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmpBytes, flag, idx, (int)tot, st));
    GM_CHECK(hipMemcpyAsync(&hv[0], idx + (tot - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GM_CHECK(hipMemcpyAsync(&hv[1], flag + (tot - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GM_CHECK(hipStreamSynchronize(st));
    nseg = (int64_t)hv[0] + hv[1];
    g->nseg = nseg;
    GM_CHECK(hipMalloc(&segEnd, sizeof(uint32_t) * nseg));
    if (nseg < (int64_t)nnz) {
        rc = c->alloc((void**)&g->segSlot, sizeof(int32_t) * nseg);
        if (rc != FEEL_HIP_OK) goto done;
    }
    hipLaunchKernelGGL(kSegments, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, keys.Current(), flag, idx, segEnd,
                       g->segSlot);
    GM_CHECK(hipGetLastError());
    // tasks: cut flags over segments, their exclusive scan, task arrays
    GM_CHECK(hipMalloc(&cut, sizeof(uint32_t) * nseg));
    hipLaunchKernelGGL(kCuts, dim3(gmGrid(nseg)), dim3(kGmBlock), 0, st, nseg, segEnd, cut);
    GM_CHECK(hipGetLastError());
    need = 0;
    GM_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, cut, idx, (int)nseg, st));
    if (need > tmpBytes) {
        GM_CHECK(hipFree(tmp));
        tmp = nullptr;
        tmpBytes = need;
        GM_CHECK(hipMalloc(&tmp, tmpBytes));
    }
    GM_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmpBytes, cut, idx, (int)nseg, st));
    GM_CHECK(hipMemcpyAsync(&hv[0], idx + (nseg - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GM_CHECK(hipStreamSynchronize(st));
    ntasks = (int64_t)hv[0] + 1;   // the last segment always cuts
    g->ntasks = ntasks;
    rc = c->alloc((void**)&g->taskSeg, sizeof(uint32_t) * (ntasks + 1));
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->taskK, sizeof(uint32_t) * (ntasks + 1));
    if (rc != FEEL_HIP_OK) goto done;
    hipLaunchKernelGGL(kTasks, dim3(gmGrid(nseg)), dim3(kGmBlock), 0, st, nseg, segEnd, cut, idx, g->taskSeg, g->taskK);
    GM_CHECK(hipGetLastError());
    // keep the sorted entries: copy into a context-counted allocation
    g->ncontrib = tot;
    rc = c->alloc((void**)&g->centry, sizeof(uint32_t) * tot);
    if (rc != FEEL_HIP_OK) goto done;
    GM_CHECK(hipMemcpyAsync(g->centry, vals.Current(), sizeof(uint32_t) * tot, hipMemcpyDeviceToDevice, st));
    GM_CHECK(hipStreamSynchronize(st));
done:
    for (void* p : {(void*)keyA, (void*)keyB, (void*)valA, (void*)valB, (void*)segEnd, (void*)cut, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace feelpp_amd
