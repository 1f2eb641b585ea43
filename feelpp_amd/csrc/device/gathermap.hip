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

// (row << 32 | col) key and entry id of every element-matrix entry
__global__ __launch_bounds__(kGmBlock) void kPairKeys(int64_t tot, int nl, const int32_t* __restrict__ e2d,
                                                     uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    const int64_t nn = (int64_t)nl * nl;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < tot; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = k / nn;
        const int ij = (int)(k - e * nn);
        const int i = ij / nl, j = ij - i * nl;
        key[k] = ((uint64_t)(uint32_t)e2d[e * nl + i] << 32) | (uint32_t)e2d[e * nl + j];
        val[k] = (uint32_t)k;
    }
}

__global__ __launch_bounds__(kGmBlock) void kHeads(int64_t tot, const uint64_t* __restrict__ key,
                                                  uint32_t* __restrict__ head) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x)
        head[q] = (q == 0 || key[q] != key[q - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(kGmBlock) void kDecrement(int64_t n, uint32_t* __restrict__ a) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        a[q] -= 1u;
}

// ja[slot] = col of every distinct key; ia[row] = slot of the row's first key
// (every row holds its diagonal, so no row is empty); ia[n] = nnz
__global__ __launch_bounds__(kGmBlock) void kPattern(int64_t tot, const uint64_t* __restrict__ key,
                                                    const uint32_t* __restrict__ head,
                                                    const uint32_t* __restrict__ slot, int64_t n, int32_t nnz,
                                                    int32_t* __restrict__ ia, int32_t* __restrict__ ja) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        if (q == 0) ia[n] = nnz;
        if (!head[q]) continue;
        const uint64_t k = key[q];
        ja[slot[q]] = (int32_t)(uint32_t)k;
        if (q == 0 || (key[q - 1] >> 32) != (k >> 32)) ia[k >> 32] = (int32_t)slot[q];
    }
}

#define GM_CHECK(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) { rc = setError(FEEL_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); goto done; } \
    } while (0)


// Gather map tail: from the entries sorted by slot (key = slot, ascending,
// entries in ascending element-entry order within a slot) to the end flags,
// the segment -> slot map, the scatter tasks and g->centry.  spareA/spareB:
// two scratch buffers of tot uint32; tmp/tmpBytes: the sort's temp storage,
// grown here when a scan needs more.
int gatherTail(feel_hip_space* s, feel_hip_graph* g, int64_t tot, const uint32_t* key, uint32_t* entry,
               uint32_t* spareA, uint32_t* spareB, void*& tmp, size_t& tmpBytes) {
    feel_hip_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    const uint32_t nnz = (uint32_t)g->nnz;
    int rc = FEEL_HIP_OK;
    uint32_t *segEnd = nullptr, *cut = nullptr;
    uint32_t hv[2] = {0, 0};
    int64_t nseg = 0, ntasks = 0;
    size_t need = 0;
    uint32_t* flag = spareA;
    uint32_t* idx = spareB;
    hipLaunchKernelGGL(kMarkEnds, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, key, entry, flag);
    GM_CHECK(hipGetLastError());
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
    hipLaunchKernelGGL(kSegments, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, key, flag, idx, segEnd, g->segSlot);
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
    GM_CHECK(hipMemcpyAsync(g->centry, entry, sizeof(uint32_t) * tot, hipMemcpyDeviceToDevice, st));
    GM_CHECK(hipStreamSynchronize(st));
done:
    for (void* p : {(void*)segEnd, (void*)cut})
        if (p) (void)hipFree(p);
    return rc;
}

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
    uint32_t *keyA = nullptr, *keyB = nullptr, *valA = nullptr, *valB = nullptr;
    void* tmp = nullptr;
    size_t tmpBytes = 0, need = 0;
    uint32_t lastKey = 0;
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
    rc = gatherTail(s, g, tot, keys.Current(), vals.Current(), vals.Alternate(), keys.Alternate(), tmp, tmpBytes);
done:
    for (void* p : {(void*)keyA, (void*)keyB, (void*)valA, (void*)valB, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

// The COUPLED sparsity pattern of the space (Stencil::computeGraph,
// feeldiscr/stencil.hpp:1502-1692; GraphCSR::close, feelalg/graphcsr.cpp:663-1021):
// row i holds the union of the dofs of the elements containing dof i, columns
// ascending.  Every element-matrix entry k = e*nldof^2 + i*nldof + j becomes
// the 64-bit key (row << 32 | col); one stable radix sort of (key, k) gives
// the pattern (the distinct keys, in row-major order) and, with the slot of a
// sorted entry = the number of distinct keys before it, the gather map of the
// scatter as buildGatherMapDevice would build it from that pattern.
// Fills g->n, g->nnz, g->ia, g->ja (device) and the gather map fields.
int buildGraphDevice(feel_hip_space* s, feel_hip_graph* g) {
    feel_hip_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    const int nl = s->nldof;
    const int64_t tot = s->ne * (int64_t)nl * nl;
    if (tot >= (1LL << 31)) return setError(FEEL_HIP_EUNSUPPORTED, "space too large for 31-bit element-entry ids");
    int rbits = 1;
    while ((1LL << rbits) < s->ndof) ++rbits;
    g->n = s->ndof;
    int rc = FEEL_HIP_OK;
    uint64_t *keyA = nullptr, *keyB = nullptr;
    uint32_t *valA = nullptr, *valB = nullptr, *head = nullptr, *slot = nullptr;
    void* tmp = nullptr;
    size_t tmpBytes = 0, need = 0;
    uint32_t hv[2] = {0, 0};
    hipcub::DoubleBuffer<uint64_t> keys;
    hipcub::DoubleBuffer<uint32_t> vals;
    GM_CHECK(hipMalloc(&keyA, sizeof(uint64_t) * tot));
    GM_CHECK(hipMalloc(&keyB, sizeof(uint64_t) * tot));
    GM_CHECK(hipMalloc(&valA, sizeof(uint32_t) * tot));
    GM_CHECK(hipMalloc(&valB, sizeof(uint32_t) * tot));
    hipLaunchKernelGGL(kPairKeys, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, nl, s->e2d, keyA, valA);
    GM_CHECK(hipGetLastError());
    keys = hipcub::DoubleBuffer<uint64_t>(keyA, keyB);
    vals = hipcub::DoubleBuffer<uint32_t>(valA, valB);
    GM_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys, vals, (int)tot, 0, 32 + rbits, st));
    tmpBytes = need;
    GM_CHECK(hipMalloc(&tmp, tmpBytes));
    GM_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmpBytes, keys, vals, (int)tot, 0, 32 + rbits, st));
    // distinct keys: head flags; slot of a sorted entry = (heads up to and including it) - 1
    head = vals.Alternate();
    slot = reinterpret_cast<uint32_t*>(keys.Alternate());
    hipLaunchKernelGGL(kHeads, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, keys.Current(), head);
    GM_CHECK(hipGetLastError());
    need = 0;
    GM_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, need, head, slot, (int)tot, st));
    if (need > tmpBytes) {
        GM_CHECK(hipFree(tmp));
        tmp = nullptr;
        tmpBytes = need;
        GM_CHECK(hipMalloc(&tmp, tmpBytes));
    }
    GM_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, tmpBytes, head, slot, (int)tot, st));
    GM_CHECK(hipMemcpyAsync(&hv[0], slot + (tot - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(kDecrement, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, slot);
    GM_CHECK(hipGetLastError());
    GM_CHECK(hipStreamSynchronize(st));
    g->nnz = (int64_t)hv[0];
    rc = c->alloc((void**)&g->ia, sizeof(int32_t) * (g->n + 1));
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&g->ja, sizeof(int32_t) * g->nnz);
    if (rc != FEEL_HIP_OK) goto done;
    hipLaunchKernelGGL(kPattern, dim3(gmGrid(tot)), dim3(kGmBlock), 0, st, tot, keys.Current(), head, slot, g->n,
                       (int32_t)g->nnz, g->ia, g->ja);
    GM_CHECK(hipGetLastError());
    // the head buffer is free again: the gather map tail needs two scratch buffers
    {
        // the sorted 64-bit keys are no longer needed: their buffer holds the slot keys' scratch
        uint32_t* spare = reinterpret_cast<uint32_t*>(keys.Current());
        rc = gatherTail(s, g, tot, slot, vals.Current(), head, spare, tmp, tmpBytes);
    }
done:
    for (void* p : {(void*)keyA, (void*)keyB, (void*)valA, (void*)valB, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace feelpp_amd
