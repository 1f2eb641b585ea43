// Device-side DofTable: Feel++'s first-touch global numbering of continuous
// Lagrange Pk dofs on a simplex mesh (DofTable::build / buildDofMap,
// feeldiscr/doftable.hpp:1802-1890,2364-2455; DofFromElement::add,
// feeldiscr/doffromelement.hpp:492-530; insertDof, doftable.hpp:1213-1300).
//
// The reference walks the elements in id order and, inside an element, its
// local dofs in local order (vertices, edges, faces, interior), giving the
// next free number to every dof not seen before; vector spaces take ncomp
// consecutive numbers per node.  The walk is sequential, but its result is a
// prefix count: visit position p = e * nloc + l; a position is new when it is
// the first position of its entity dof (interior dofs are always new); its
// number is ncomp x (new positions before p), and every other position takes
// the number of its entity dof's first position.  Here:
//   1. edges and faces get ids by a sort of their vertex keys (the ids are
//      only names; the numbering does not depend on them);
//   2. each position's entity dof: vertex id, edge id x nde + slot (slots
//      reversed when the element sees the edge from its larger vertex,
//      mesh3d.hpp:733-734,759-760), face id x ndf + rank of the vertex the dof
//      sits next to;
//   3. first position of every entity dof by atomicMin (order independent);
//   4. exclusive scan of the new flags, then every position reads its number.
// The dof -> (element, local) incidence used by the linear-form gather is a
// stable sort of the element-to-dof entries by dof.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "common.hpp"

using namespace feelpp_amd;

namespace {

constexpr int kDtBlock = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;

int dtGrid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kDtBlock - 1) / kDtBlock, 65536)); }

// local entity tables of the reference simplex (refsimplex.hpp:139-164 vertex
// order; edge i joins the vertices listed, face f is opposite vertex f)
__constant__ int cTriEdge[3][2] = {{1, 2}, {2, 0}, {0, 1}};
__constant__ int cTetEdge[6][2] = {{1, 2}, {2, 0}, {0, 1}, {0, 3}, {1, 3}, {2, 3}};
__constant__ int cTetFace[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

struct DtShape {
    int dim, nvx, nedge, nde, ndf, ndv, nloc, ncomp;
    int64_t ne, nv;
};

__device__ inline uint64_t edgeKey(const int32_t* el, int i, int dim) {
    const int a = dim == 2 ? el[cTriEdge[i][0]] : el[cTetEdge[i][0]];
    const int b = dim == 2 ? el[cTriEdge[i][1]] : el[cTetEdge[i][1]];
    return ((uint64_t)min(a, b) << 32) | (uint32_t)max(a, b);
}

__device__ inline void sort3(int32_t* v) {
    if (v[0] > v[1]) { const int32_t t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[1] > v[2]) { const int32_t t = v[1]; v[1] = v[2]; v[2] = t; }
    if (v[0] > v[1]) { const int32_t t = v[0]; v[0] = v[1]; v[1] = t; }
}

__device__ inline uint64_t faceKey(const int32_t* el, int f) {
    int32_t v[3] = {el[cTetFace[f][0]], el[cTetFace[f][1]], el[cTetFace[f][2]]};
    sort3(v);
    return ((uint64_t)v[0] << 42) | ((uint64_t)v[1] << 21) | (uint64_t)v[2];
}

// keys of all element edges (nedge per element) or faces (4 per element)
__global__ __launch_bounds__(kDtBlock) void kEntityKeys(DtShape S, const int32_t* __restrict__ elems, int faces,
                                                       uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const int per = faces ? 4 : S.nedge;
    const int64_t tot = S.ne * per;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = q / per;
        const int i = (int)(q - e * per);
        const int32_t* el = elems + e * S.nvx;
        key[q] = faces ? faceKey(el, i) : edgeKey(el, i, S.dim);
        idx[q] = (uint32_t)q;
    }
}

__global__ __launch_bounds__(kDtBlock) void kHeads64(int64_t n, const uint64_t* __restrict__ key,
                                                    uint32_t* __restrict__ head) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        head[q] = (q == 0 || key[q] != key[q - 1]) ? 1u : 0u;
}

// id of the entity at every (element, local entity): the inclusive count of
// distinct keys up to its sorted position, minus one
__global__ __launch_bounds__(kDtBlock) void kScatterIds(int64_t n, const uint32_t* __restrict__ sortedIdx,
                                                       const uint32_t* __restrict__ incl, uint32_t* __restrict__ id) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        id[sortedIdx[q]] = incl[q] - 1;
}

// entity dof of every visit position (kNone for interior dofs) and the first
// position of every entity dof.  Entity dofs: [vertices | edges x nde | faces x ndf].
__global__ __launch_bounds__(kDtBlock) void kEntityDofs(DtShape S, const int32_t* __restrict__ elems,
                                                       const uint32_t* __restrict__ edgeId,
                                                       const uint32_t* __restrict__ faceId, int64_t nedges,
                                                       uint32_t* __restrict__ ent, uint32_t* __restrict__ first) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < S.ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t* el = elems + e * S.nvx;
        const uint32_t p0 = (uint32_t)(e * S.nloc);
        int l = 0;
        auto visit = [&](uint32_t d) {
            ent[p0 + l] = d;
            if (d != kNone) atomicMin(first + d, p0 + l);
            ++l;
        };
        for (int i = 0; i < S.nvx; ++i) visit((uint32_t)el[i]);
        for (int i = 0; i < S.nedge && S.nde > 0; ++i) {
            const int a = S.dim == 2 ? el[cTriEdge[i][0]] : el[cTetEdge[i][0]];
            const int b = S.dim == 2 ? el[cTriEdge[i][1]] : el[cTetEdge[i][1]];
            const uint32_t base = (uint32_t)S.nv + edgeId[e * S.nedge + i] * (uint32_t)S.nde;
            for (int k = 0; k < S.nde; ++k) visit(base + (uint32_t)(a > b ? S.nde - 1 - k : k));
        }
        if (S.ndf > 0)
            for (int f = 0; f < 4; ++f) {
                const int32_t fv[3] = {el[cTetFace[f][0]], el[cTetFace[f][1]], el[cTetFace[f][2]]};
                int32_t v[3] = {fv[0], fv[1], fv[2]};
                sort3(v);
                const uint32_t base = (uint32_t)(S.nv + nedges * S.nde) + faceId[e * 4 + f] * (uint32_t)S.ndf;
                for (int k = 0; k < S.ndf; ++k) {
                    // several dofs per face (P4): slot by the mesh vertex the dof sits next to
                    const int rank = S.ndf == 1 ? 0 : (fv[k] == v[0] ? 0 : (fv[k] == v[1] ? 1 : 2));
                    visit(base + (uint32_t)rank);
                }
            }
        for (int k = 0; k < S.ndv; ++k) visit(kNone);
    }
}

__global__ __launch_bounds__(kDtBlock) void kNewFlags(int64_t n, const uint32_t* __restrict__ ent,
                                                     const uint32_t* __restrict__ first, uint32_t* __restrict__ isNew) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        isNew[p] = (ent[p] == kNone || first[ent[p]] == (uint32_t)p) ? 1u : 0u;
}

// elem2dof[e][c * nloc + l] = ncomp x rank(first position) + c
__global__ __launch_bounds__(kDtBlock) void kNumber(DtShape S, const uint32_t* __restrict__ ent,
                                                   const uint32_t* __restrict__ first,
                                                   const uint32_t* __restrict__ rank, int32_t* __restrict__ e2d) {
    const int64_t n = S.ne * S.nloc;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t f = ent[p] == kNone ? (uint32_t)p : first[ent[p]];
        const int32_t base = (int32_t)(rank[f] * (uint32_t)S.ncomp);
        const int64_t e = p / S.nloc;
        const int l = (int)(p - e * S.nloc);
        int32_t* out = e2d + e * (int64_t)S.nloc * S.ncomp;
        for (int c = 0; c < S.ncomp; ++c) out[c * S.nloc + l] = base + c;
    }
}

__global__ __launch_bounds__(kDtBlock) void kFill32(int64_t n, uint32_t v, uint32_t* __restrict__ a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = v;
}

__global__ __launch_bounds__(kDtBlock) void kIota(int64_t n, uint32_t* __restrict__ a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = (uint32_t)i;
}

// rptr[d] = first position of dof d in the dof-sorted entries (every dof occurs)
__global__ __launch_bounds__(kDtBlock) void kRowStarts(int64_t n, int64_t ndof, const uint32_t* __restrict__ key,
                                                      uint32_t* __restrict__ ptr) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        if (q == 0 || key[q] != key[q - 1]) ptr[key[q]] = (uint32_t)q;
        if (q == n - 1) ptr[ndof] = (uint32_t)n;
    }
}

#define DT_CHECK(expr)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) { rc = setError(FEEL_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); goto done; } \
    } while (0)

// ids of the distinct keys among n entity keys: id[q] for entity q (element-major)
int entityIds(hipStream_t st, const DtShape& S, const int32_t* elems, int faces, int keyBits, uint32_t* id,
              int64_t* count) {
    const int64_t n = S.ne * (faces ? 4 : S.nedge);
    int rc = FEEL_HIP_OK;
    uint64_t *kA = nullptr, *kB = nullptr;
    uint32_t *vA = nullptr, *vB = nullptr, *head = nullptr, *rank = nullptr;
    void* tmp = nullptr;
    size_t need = 0, tmpBytes = 0;
    uint32_t hv[2] = {0, 0};
    hipcub::DoubleBuffer<uint64_t> keys;
    hipcub::DoubleBuffer<uint32_t> vals;
    DT_CHECK(hipMalloc(&kA, sizeof(uint64_t) * n));
    DT_CHECK(hipMalloc(&kB, sizeof(uint64_t) * n));
    DT_CHECK(hipMalloc(&vA, sizeof(uint32_t) * n));
    DT_CHECK(hipMalloc(&vB, sizeof(uint32_t) * n));
    hipLaunchKernelGGL(kEntityKeys, dim3(dtGrid(n)), dim3(kDtBlock), 0, st, S, elems, faces, kA, vA);
    DT_CHECK(hipGetLastError());
    keys = hipcub::DoubleBuffer<uint64_t>(kA, kB);
    vals = hipcub::DoubleBuffer<uint32_t>(vA, vB);
    DT_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys, vals, (int)n, 0, keyBits, st));
    tmpBytes = need;
    DT_CHECK(hipMalloc(&tmp, tmpBytes));
    DT_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmpBytes, keys, vals, (int)n, 0, keyBits, st));
    head = vals.Alternate();
    rank = reinterpret_cast<uint32_t*>(keys.Alternate());
    hipLaunchKernelGGL(kHeads64, dim3(dtGrid(n)), dim3(kDtBlock), 0, st, n, keys.Current(), head);
    DT_CHECK(hipGetLastError());
    need = 0;
    DT_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, need, head, rank, (int)n, st));
    if (need > tmpBytes) {
        DT_CHECK(hipFree(tmp));
        tmp = nullptr;
        tmpBytes = need;
        DT_CHECK(hipMalloc(&tmp, tmpBytes));
    }
    DT_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, tmpBytes, head, rank, (int)n, st));
    hipLaunchKernelGGL(kScatterIds, dim3(dtGrid(n)), dim3(kDtBlock), 0, st, n, vals.Current(), rank, id);
    DT_CHECK(hipGetLastError());
    DT_CHECK(hipMemcpyAsync(&hv[0], rank + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DT_CHECK(hipStreamSynchronize(st));
    *count = hv[0];
done:
    for (void* p : {(void*)kA, (void*)kB, (void*)vA, (void*)vB, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace

namespace feelpp_amd {

// elem2dof (device, ne x nloc*ncomp) of the space's mesh and its dof count
int buildDofTableDevice(feel_hip_space* s) {
    feel_hip_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    DtShape S{};
    S.dim = s->dim;
    S.nvx = s->dim + 1;
    S.nedge = s->dim == 2 ? 3 : 6;
    S.nde = s->order - 1;
    S.ndf = s->dim == 3 ? (s->order - 1) * (s->order - 2) / 2 : 0;
    S.ndv = s->dim == 3 ? (s->order - 1) * (s->order - 2) * (s->order - 3) / 6
                        : (s->order >= 3 ? (s->order - 1) * (s->order - 2) / 2 : 0);
    S.nloc = s->nloc;
    S.ncomp = s->ncomp;
    S.ne = s->ne;
    S.nv = s->nv;
    const int64_t npos = S.ne * S.nloc;
    if (npos >= (1LL << 32) - 1) return setError(FEEL_HIP_EUNSUPPORTED, "space too large for 32-bit positions");
    if (S.ndf > 0 && S.nv >= (1LL << 21)) return setError(FEEL_HIP_EUNSUPPORTED, "face keys need < 2^21 vertices");
    int vbits = 1;
    while ((1LL << vbits) < S.nv) ++vbits;
    int rc = FEEL_HIP_OK;
    uint32_t *edgeId = nullptr, *faceId = nullptr, *ent = nullptr, *first = nullptr, *isNew = nullptr, *rank = nullptr;
    void* tmp = nullptr;
    size_t need = 0;
    int64_t nedges = 0, nfaces = 0, nent = 0;
    uint32_t hv[2] = {0, 0};
    if (S.nde > 0) {
        DT_CHECK(hipMalloc(&edgeId, sizeof(uint32_t) * S.ne * S.nedge));
        rc = entityIds(st, S, s->elems, 0, 32 + vbits, edgeId, &nedges);
        if (rc != FEEL_HIP_OK) goto done;
    }
    if (S.ndf > 0) {
        DT_CHECK(hipMalloc(&faceId, sizeof(uint32_t) * S.ne * 4));
        rc = entityIds(st, S, s->elems, 1, 63, faceId, &nfaces);
        if (rc != FEEL_HIP_OK) goto done;
    }
    nent = S.nv + nedges * S.nde + nfaces * S.ndf;
    if (nent >= (1LL << 32) - 1) { rc = setError(FEEL_HIP_EUNSUPPORTED, "too many entity dofs"); goto done; }
    DT_CHECK(hipMalloc(&ent, sizeof(uint32_t) * npos));
    DT_CHECK(hipMalloc(&first, sizeof(uint32_t) * nent));
    DT_CHECK(hipMalloc(&isNew, sizeof(uint32_t) * npos));
    DT_CHECK(hipMalloc(&rank, sizeof(uint32_t) * npos));
    hipLaunchKernelGGL(kFill32, dim3(dtGrid(nent)), dim3(kDtBlock), 0, st, nent, kNone, first);
    hipLaunchKernelGGL(kEntityDofs, dim3(dtGrid(S.ne)), dim3(kDtBlock), 0, st, S, s->elems, edgeId, faceId, nedges, ent,
                       first);
    hipLaunchKernelGGL(kNewFlags, dim3(dtGrid(npos)), dim3(kDtBlock), 0, st, npos, ent, first, isNew);
    DT_CHECK(hipGetLastError());
    DT_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, isNew, rank, (int)npos, st));
    DT_CHECK(hipMalloc(&tmp, need));
    DT_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, need, isNew, rank, (int)npos, st));
    DT_CHECK(hipMemcpyAsync(&hv[0], rank + (npos - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DT_CHECK(hipMemcpyAsync(&hv[1], isNew + (npos - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DT_CHECK(hipStreamSynchronize(st));
    s->ndof = ((int64_t)hv[0] + hv[1]) * S.ncomp;
    if (s->ndof >= (1LL << 31)) { rc = setError(FEEL_HIP_EUNSUPPORTED, "more than 2^31 dofs"); goto done; }
    rc = c->alloc((void**)&s->e2d, sizeof(int32_t) * npos * S.ncomp);
    if (rc != FEEL_HIP_OK) goto done;
    hipLaunchKernelGGL(kNumber, dim3(dtGrid(npos)), dim3(kDtBlock), 0, st, S, ent, first, rank, s->e2d);
    DT_CHECK(hipGetLastError());
    DT_CHECK(hipStreamSynchronize(st));
done:
    for (void* p : {(void*)edgeId, (void*)faceId, (void*)ent, (void*)first, (void*)isNew, (void*)rank, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

// dof -> (element, local) incidence: rptr[ndof+1], rentry[ne*nldof] with the
// entries of each dof in ascending element-entry order
int buildIncidenceDevice(feel_hip_space* s) {
    feel_hip_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    const int64_t n = s->ne * s->nldof;
    if (n >= (1LL << 31)) return setError(FEEL_HIP_EUNSUPPORTED, "space too large");
    int bits = 1;
    while ((1LL << bits) <= s->ndof) ++bits;
    int rc = FEEL_HIP_OK;
    uint32_t *kA = nullptr, *kB = nullptr, *vA = nullptr;
    void* tmp = nullptr;
    size_t need = 0;
    hipcub::DoubleBuffer<uint32_t> keys, vals;
    DT_CHECK(hipMalloc(&kA, sizeof(uint32_t) * n));
    DT_CHECK(hipMalloc(&kB, sizeof(uint32_t) * n));
    DT_CHECK(hipMalloc(&vA, sizeof(uint32_t) * n));
    rc = c->alloc((void**)&s->rentry, sizeof(uint32_t) * n);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&s->rptr, sizeof(uint32_t) * (s->ndof + 1));
    if (rc != FEEL_HIP_OK) goto done;
    DT_CHECK(hipMemcpyAsync(kA, s->e2d, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(kIota, dim3(dtGrid(n)), dim3(kDtBlock), 0, st, n, vA);
    DT_CHECK(hipGetLastError());
    keys = hipcub::DoubleBuffer<uint32_t>(kA, kB);
    vals = hipcub::DoubleBuffer<uint32_t>(vA, s->rentry);
    DT_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys, vals, (int)n, 0, bits, st));
    DT_CHECK(hipMalloc(&tmp, need));
    DT_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, need, keys, vals, (int)n, 0, bits, st));
    if (vals.Current() != s->rentry)
        DT_CHECK(hipMemcpyAsync(s->rentry, vals.Current(), sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(kRowStarts, dim3(dtGrid(n)), dim3(kDtBlock), 0, st, n, s->ndof, keys.Current(), s->rptr);
    DT_CHECK(hipGetLastError());
    DT_CHECK(hipStreamSynchronize(st));
done:
    for (void* p : {(void*)kA, (void*)kB, (void*)vA, tmp})
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace feelpp_amd
