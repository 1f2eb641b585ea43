// C-ABI implementation of the MI355X device layer: handles, uploads, assembly,
// Dirichlet elimination and the Jacobi-preconditioned CG (FP64, and FP32 inside
// FP64 iterative refinement).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../host/fe.hpp"
#include "../host/space.hpp"
#include "common.hpp"
#include "kernels.hpp"

using namespace feelpp_amd;

namespace feelpp_amd {
thread_local std::string gLastError;
int setError(int code, const std::string& msg) {
    gLastError = msg;
    return code;
}
}  // namespace feelpp_amd

namespace {

int gridFor(int64_t n, int cap = 1 << 16) {
    int64_t g = (n + kBlock - 1) / kBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

// Number of scalar reference-tensor slots for the symmetric gradient form
int numSym(int dim) { return dim == 3 ? 6 : 3; }

#define TRY_CALL(expr)                       \
    do {                                     \
        int rc_ = (expr);                    \
        if (rc_ != FEEL_HIP_OK) return rc_;  \
    } while (0)

#define GUARD_BEGIN try {
#define GUARD_END                                                         \
    }                                                                     \
    catch (const std::bad_alloc&) {                                       \
        return setError(FEEL_HIP_ENOMEM, "host allocation failed");       \
    }                                                                     \
    catch (const std::invalid_argument& e) {                              \
        return setError(FEEL_HIP_EUNSUPPORTED, e.what());                 \
    }                                                                     \
    catch (const std::exception& e) {                                     \
        return setError(FEEL_HIP_EINVAL, e.what());                       \
    }

template <typename T>
int devAlloc(feel_hip_ctx* ctx, T** p, size_t count) {
    return ctx->alloc((void**)p, count * sizeof(T));
}

template <typename T>
int upload(feel_hip_ctx* ctx, T** p, const T* host, size_t count) {
    TRY_CALL(devAlloc(ctx, p, count));
    if (count) FEEL_HIP_CHECK(hipMemcpyAsync(*p, host, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return FEEL_HIP_OK;
}

}  // namespace

// ============================================================== context
int feel_hip_ctx::alloc(void** p, size_t n) {
    *p = nullptr;
    if (n == 0) return FEEL_HIP_OK;
    hipError_t e = hipMalloc(p, n);
    if (e != hipSuccess) {
        *p = nullptr;
        return setError(FEEL_HIP_ENOMEM, "hipMalloc(" + std::to_string(n) + " bytes): " + hipGetErrorString(e));
    }
    bytes += (int64_t)n;
    return FEEL_HIP_OK;
}

void feel_hip_ctx::release(void* p, size_t n) {
    if (!p) return;
    (void)hipFree(p);
    bytes -= (int64_t)n;
}

void feel_hip_ctx::tic(const char* name) {
    pending = false;
    if (!timing) return;
    if (period > 1 && timerCalls[name]++ % period != 0) return;
    if (evUsed == events.size()) {
        Ev ev;
        if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) return;
        events.push_back(ev);
    }
    events[evUsed].name = name;
    (void)hipEventRecord(events[evUsed].a, stream);
    pending = true;
}

void feel_hip_ctx::toc() {
    if (!timing || !pending || evUsed >= events.size()) return;
    (void)hipEventRecord(events[evUsed].b, stream);
    ++evUsed;
    pending = false;
}

void feel_hip_ctx::collectTimers() {
    if (evUsed == 0) return;
    (void)hipStreamSynchronize(stream);
    for (size_t i = 0; i < evUsed; ++i) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, events[i].a, events[i].b) == hipSuccess) {
            auto& t = timers[events[i].name];
            t.first += 1;
            t.second += ms;
        }
    }
    evUsed = 0;
}

namespace feelpp_amd {

void retain(feel_hip_ctx* c) { ++c->refs; }
void retain(feel_hip_space* s) { ++s->refs; }
void retain(feel_hip_graph* g) { ++g->refs; }

void unref(feel_hip_ctx* c) {
    if (--c->refs > 0) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->release(c->part, sizeof(double) * kPartDoubles);
    c->release(c->state, sizeof(CgState) + 4 * sizeof(double));
    c->release(c->hist, sizeof(double) * 2 * kHistMax);
    c->release(c->elemmat, c->elemmatBytes);
    if (c->hstate) (void)hipHostFree(c->hstate);
    for (auto& ev : c->events) {
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void unref(feel_hip_space* s) {
    if (--s->refs > 0) return;
    feel_hip_ctx* c = s->ctx;
    (void)hipStreamSynchronize(c->stream);
    c->release(s->coords, sizeof(double) * s->nv * s->dim);
    c->release(s->elems, sizeof(int32_t) * s->ne * (s->dim + 1));
    c->release(s->e2d, sizeof(int32_t) * s->ne * s->nldof);
    c->release(s->rptr, sizeof(uint32_t) * (s->ndof + 1));
    c->release(s->rentry, sizeof(uint32_t) * s->ne * s->nldof);
    c->release(s->elemvec, sizeof(double) * s->ne * s->nldof);
    c->release(s->geo, s->geoBytes);
    for (auto& kv : s->rtab) c->release(kv.second, sizeof(double) * s->hrtab[kv.first].size());
    for (auto& kv : s->qtab)
        c->release(kv.second, sizeof(double) * (size_t)s->qnpts[kv.first] * (1 + s->nloc + s->dim + 1));
    for (auto& kv : s->ntab) c->release(kv.second.first, sizeof(double) * kv.second.second);
    delete s;
    unref(c);
}

void unref(feel_hip_graph* g) {
    if (--g->refs > 0) return;
    feel_hip_ctx* c = g->space->ctx;
    (void)hipStreamSynchronize(c->stream);
    c->release(g->ia, sizeof(int32_t) * (g->n + 1));
    c->release(g->ja, sizeof(int32_t) * g->nnz);
    c->release(g->diagpos, sizeof(int32_t) * g->n);
    if (g->tpos) c->release(g->tpos, sizeof(int32_t) * g->nnz);
    c->release(g->centry, sizeof(uint32_t) * g->ncontrib);
    c->release(g->taskSeg, sizeof(uint32_t) * (g->ntasks + 1));
    c->release(g->taskK, sizeof(uint32_t) * (g->ntasks + 1));
    if (g->segSlot) c->release(g->segSlot, sizeof(int32_t) * g->nseg);
    c->release(g->slicePtr, sizeof(uint32_t) * (g->nslices + 1));
    if (g->sliceRow) c->release(g->sliceRow, sizeof(int32_t) * (g->nslices + 1));
    if (g->sliceShift) c->release(g->sliceShift, sizeof(uint8_t) * g->nslices);
    c->release(g->scol, sizeof(int32_t) * (g->block ? g->nbslots : g->nslots));
    c->release(g->slot2csr, sizeof(int32_t) * g->nslots);
    c->release(g->perm, sizeof(int32_t) * g->n);
    c->release(g->sdiag, sizeof(int32_t) * g->n);
    if (g->compact) {
        c->release(g->colBase, sizeof(int32_t) * g->nslices);
        c->release(g->ccptr, sizeof(int64_t) * g->nslices);
        c->release(g->packed, sizeof(uint16_t) * g->npacked);
    }
    feel_hip_space* sp = g->space;
    delete g;
    unref(sp);
}

}  // namespace feelpp_amd

extern "C" {

const char* feel_hip_last_error(void) { return gLastError.c_str(); }

int feel_hip_ctx_create(int device, feel_hip_ctx** out) {
    if (!out) return setError(FEEL_HIP_EINVAL, "null output handle");
    *out = nullptr;
    int ndev = 0;
    FEEL_HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return setError(FEEL_HIP_EINVAL, "no such HIP device");
    FEEL_HIP_CHECK(hipSetDevice(device));
    if (int rs = noallocSetup(device); rs != FEEL_HIP_OK) return rs;
    auto* c = new feel_hip_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->numCUs = prop.multiProcessorCount;
    if (c->numCUs < 1) c->numCUs = 1;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return setError(FEEL_HIP_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    int rc = c->alloc((void**)&c->part, sizeof(double) * kPartDoubles);
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&c->state, sizeof(CgState) + 4 * sizeof(double));
    if (rc == FEEL_HIP_OK) rc = c->alloc((void**)&c->hist, sizeof(double) * 2 * kHistMax);
    if (rc == FEEL_HIP_OK && hipHostMalloc((void**)&c->hstate, sizeof(CgState), hipHostMallocDefault) != hipSuccess)
        rc = setError(FEEL_HIP_ENOMEM, "hipHostMalloc failed");
    if (rc != FEEL_HIP_OK) {
        feel_hip_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return FEEL_HIP_OK;
}

int feel_hip_ctx_destroy(feel_hip_ctx* c) {
    if (c) unref(c);
    return FEEL_HIP_OK;
}

int feel_hip_ctx_synchronize(feel_hip_ctx* c) {
    if (!c) return setError(FEEL_HIP_EINVAL, "null context");
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    return FEEL_HIP_OK;
}

void* feel_hip_ctx_stream(feel_hip_ctx* c) { return c ? (void*)c->stream : nullptr; }
int64_t feel_hip_ctx_device_bytes(feel_hip_ctx* c) { return c ? c->bytes : -1; }

int feel_hip_ctx_set_timing(feel_hip_ctx* c, int period) {
    if (!c) return setError(FEEL_HIP_EINVAL, "null context");
    if (period < 0) return setError(FEEL_HIP_EINVAL, "timing period must be >= 0");
    c->collectTimers();
    c->timing = period > 0;
    c->period = period;
    c->timerCalls.clear();
    return FEEL_HIP_OK;
}

int feel_hip_ctx_timing(feel_hip_ctx* c, const char* name, int64_t* launches, double* total_ms) {
    if (!c || !name) return setError(FEEL_HIP_EINVAL, "null argument");
    c->collectTimers();
    auto it = c->timers.find(name);
    if (launches) *launches = it == c->timers.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == c->timers.end() ? 0.0 : it->second.second;
    return FEEL_HIP_OK;
}

int feel_hip_ctx_reset_timing(feel_hip_ctx* c) {
    if (!c) return setError(FEEL_HIP_EINVAL, "null context");
    c->collectTimers();
    c->timers.clear();
    return FEEL_HIP_OK;
}

// ============================================================== space
int feel_hip_space_create(feel_hip_ctx* ctx, int dim, int order, int ncomp, int64_t nverts, const double* coords,
                          int64_t nelems, const int32_t* elems, const int32_t* elem2dof, int64_t ndof,
                          feel_hip_space** out) {
    GUARD_BEGIN
    if (!ctx || !out || !coords || !elems || !elem2dof) return setError(FEEL_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (dim < 2 || dim > 3) return setError(FEEL_HIP_EUNSUPPORTED, "dim must be 2 or 3");
    if (order < 1 || order > maxSupportedOrder(dim))
        return setError(FEEL_HIP_EUNSUPPORTED, "order " + std::to_string(order) + " not supported in " +
                                                   std::to_string(dim) + "D");
    if (ncomp != 1 && ncomp != dim) return setError(FEEL_HIP_EINVAL, "ncomp must be 1 or dim");
    if (nverts <= 0 || nelems <= 0 || ndof <= 0) return setError(FEEL_HIP_EINVAL, "empty mesh or space");
    const int nloc = nLocalDof(dim, order), nldof = nloc * ncomp;
    // validate connectivity before it reaches a kernel
    for (int64_t k = 0; k < nelems * (dim + 1); ++k)
        if (elems[k] < 0 || elems[k] >= nverts) return setError(FEEL_HIP_EINVAL, "element vertex index out of range");
    for (int64_t k = 0; k < nelems * nldof; ++k)
        if (elem2dof[k] < 0 || elem2dof[k] >= ndof) return setError(FEEL_HIP_EINVAL, "elem2dof entry out of range");
    if ((uint64_t)nelems * nldof >= (1ULL << 32)) return setError(FEEL_HIP_EUNSUPPORTED, "space too large");
    FEEL_HIP_CHECK(hipSetDevice(ctx->device));
    auto* s = new feel_hip_space();
    s->ctx = ctx;
    retain(ctx);
    s->dim = dim; s->order = order; s->ncomp = ncomp; s->nloc = nloc; s->nldof = nldof;
    s->nv = nverts; s->ne = nelems; s->ndof = ndof;
    s->hElem2dof.assign(elem2dof, elem2dof + nelems * nldof);
    int rc = upload(ctx, &s->coords, coords, (size_t)nverts * dim);
    if (rc == FEEL_HIP_OK) rc = upload(ctx, &s->elems, elems, (size_t)nelems * (dim + 1));
    if (rc == FEEL_HIP_OK) rc = upload(ctx, &s->e2d, elem2dof, (size_t)nelems * nldof);
    if (rc == FEEL_HIP_OK) {
        DofTable dt;
        dt.dim = dim; dt.order = order; dt.ncomp = ncomp; dt.nloc = nloc; dt.nldof = nldof;
        dt.ne = nelems; dt.ndof = ndof;
        dt.elem2dof = s->hElem2dof;
        Incidence inc = buildIncidence(dt);
        std::vector<uint32_t> p32(inc.ptr.begin(), inc.ptr.end());
        rc = upload(ctx, &s->rptr, p32.data(), p32.size());
        if (rc == FEEL_HIP_OK) rc = upload(ctx, &s->rentry, inc.entry.data(), inc.entry.size());
    }
    if (rc == FEEL_HIP_OK) rc = devAlloc(ctx, &s->elemvec, (size_t)nelems * nldof);
    if (rc != FEEL_HIP_OK) {
        feel_hip_space_destroy(s);
        return rc;
    }
    *out = s;
    return FEEL_HIP_OK;
    GUARD_END
}

int feel_hip_space_create_from_mesh(feel_hip_ctx* ctx, int dim, int order, int ncomp, int64_t nverts,
                                    const double* coords, int64_t nelems, const int32_t* elems,
                                    feel_hip_space** out) {
    GUARD_BEGIN
    if (!ctx || !out || !coords || !elems) return setError(FEEL_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (dim < 2 || dim > 3) return setError(FEEL_HIP_EUNSUPPORTED, "dim must be 2 or 3");
    if (order < 1 || order > maxSupportedOrder(dim))
        return setError(FEEL_HIP_EUNSUPPORTED, "order " + std::to_string(order) + " not supported in " +
                                                   std::to_string(dim) + "D");
    if (ncomp != 1 && ncomp != dim) return setError(FEEL_HIP_EINVAL, "ncomp must be 1 or dim");
    if (nverts <= 0 || nelems <= 0) return setError(FEEL_HIP_EINVAL, "empty mesh");
    if (nverts >= (1LL << 31)) return setError(FEEL_HIP_EUNSUPPORTED, "more than 2^31 vertices");
    const int nloc = nLocalDof(dim, order), nldof = nloc * ncomp;
    for (int64_t k = 0; k < nelems * (dim + 1); ++k)
        if (elems[k] < 0 || elems[k] >= nverts) return setError(FEEL_HIP_EINVAL, "element vertex index out of range");
    if ((uint64_t)nelems * nldof >= (1ULL << 31)) return setError(FEEL_HIP_EUNSUPPORTED, "space too large");
    FEEL_HIP_CHECK(hipSetDevice(ctx->device));
    auto* s = new feel_hip_space();
    s->ctx = ctx;
    retain(ctx);
    s->dim = dim; s->order = order; s->ncomp = ncomp; s->nloc = nloc; s->nldof = nldof;
    s->nv = nverts; s->ne = nelems;
    int rc = upload(ctx, &s->coords, coords, (size_t)nverts * dim);
    if (rc == FEEL_HIP_OK) rc = upload(ctx, &s->elems, elems, (size_t)nelems * (dim + 1));
    if (rc == FEEL_HIP_OK) rc = buildDofTableDevice(s);
    if (rc == FEEL_HIP_OK) rc = buildIncidenceDevice(s);
    if (rc == FEEL_HIP_OK) rc = devAlloc(ctx, &s->elemvec, (size_t)nelems * nldof);
    if (rc != FEEL_HIP_OK) {
        feel_hip_space_destroy(s);
        return rc;
    }
    *out = s;
    return FEEL_HIP_OK;
    GUARD_END
}

int feel_hip_space_elem2dof(const feel_hip_space* s, int32_t* elem2dof) {
    if (!s || !elem2dof) return setError(FEEL_HIP_EINVAL, "null argument");
    FEEL_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    FEEL_HIP_CHECK(hipMemcpy(elem2dof, s->e2d, sizeof(int32_t) * s->ne * s->nldof, hipMemcpyDeviceToHost));
    return FEEL_HIP_OK;
}

int feel_hip_space_destroy(feel_hip_space* s) {
    if (s) unref(s);
    return FEEL_HIP_OK;
}

int64_t feel_hip_space_ndof(const feel_hip_space* s) { return s ? s->ndof : -1; }
int64_t feel_hip_space_nelements(const feel_hip_space* s) { return s ? s->ne : -1; }

}  // extern "C"

namespace {

// scatter launch: one wavefront per task at most (4 per 256-thread block)
#define SEG_ARGS g->ntasks, g->taskSeg, g->taskK, g->centry, g->segSlot, s->geo, rt, v64, v32, add
int segGrid(const feel_hip_graph* g) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((g->ntasks + kBlock / 64 - 1) / (kBlock / 64), 8192));
}

// A replacing assembly writes only slots that receive element entries; a
// pattern with slots no element touches gets them cleared first.
int clearEmptySlots(feel_hip_ctx* c, const feel_hip_graph* g, double* v64, float* v32, int add) {
    if (add || !g->segSlot) return FEEL_HIP_OK;
    if (v64) FEEL_HIP_CHECK(hipMemsetAsync(v64, 0, sizeof(double) * g->nnz, c->stream));
    if (v32) FEEL_HIP_CHECK(hipMemsetAsync(v32, 0, sizeof(float) * g->nnz, c->stream));
    return FEEL_HIP_OK;
}

// reference tensors of the scalar forms, entry-major: row ij holds
// [Rsym_0 .. Rsym_{NS-1} | M | 0 padding] (ST = 8 in 3D, 4 in 2D doubles),
// cached per quadrature order
int scalarTensors(feel_hip_space* s, int qorder, const double** out) {
    auto it = s->rtab.find(qorder);
    if (it != s->rtab.end()) { *out = it->second; return FEEL_HIP_OK; }
    ReferenceTensors T = referenceTensors(s->dim, s->order, qorder);
    const int n = T.n, nn = n * n, d = s->dim, ns = numSym(d), st = d == 3 ? 8 : 4;
    std::vector<double> h((size_t)st * nn, 0.0);
    auto R = [&](int a, int b, int ij) { return T.R[(size_t)(a * d + b) * nn + ij]; };
    for (int ij = 0; ij < nn; ++ij) {
        double* row = h.data() + (size_t)ij * st;
        if (d == 3) {
            row[0] = R(0, 0, ij);
            row[1] = R(1, 1, ij);
            row[2] = R(2, 2, ij);
            row[3] = R(0, 1, ij) + R(1, 0, ij);
            row[4] = R(0, 2, ij) + R(2, 0, ij);
            row[5] = R(1, 2, ij) + R(2, 1, ij);
        } else {
            row[0] = R(0, 0, ij);
            row[1] = R(1, 1, ij);
            row[2] = R(0, 1, ij) + R(1, 0, ij);
        }
        row[ns] = T.M[ij];
    }
    double* dptr = nullptr;
    TRY_CALL(upload(s->ctx, &dptr, h.data(), h.size()));
    s->rtab[qorder] = dptr;
    s->hrtab[qorder] = std::move(h);
    *out = dptr;
    return FEEL_HIP_OK;
}

// full reference tensors R_ab (a,b < 3) for the vector forms, cached per quadrature order
// under key -(qorder+1) so they do not collide with the scalar tables
int fullTensors(feel_hip_space* s, int qorder, const double** out) {
    const int key = -(qorder + 1);
    auto it = s->rtab.find(key);
    if (it != s->rtab.end()) { *out = it->second; return FEEL_HIP_OK; }
    ReferenceTensors T = referenceTensors(s->dim, s->order, qorder);
    // entry-major, 10 doubles per local pair (i, j): R_ab at a*3+b, zero pad
    const size_t nn = (size_t)T.n * T.n;
    std::vector<double> h(10 * nn, 0.0);
    for (size_t ij = 0; ij < nn; ++ij)
        for (int t = 0; t < 9; ++t) h[ij * 10 + t] = T.R[t * nn + ij];
    double* dptr = nullptr;
    TRY_CALL(upload(s->ctx, &dptr, h.data(), h.size()));
    s->rtab[key] = dptr;
    s->hrtab[key] = std::move(h);
    *out = dptr;
    return FEEL_HIP_OK;
}

// quadrature table for the linear form: w[Q] | phi[nloc*Q] | lambda[(dim+1)*Q]
int rhsTable(feel_hip_space* s, int qorder, const double** out, int* Q) {
    auto it = s->qtab.find(qorder);
    if (it != s->qtab.end()) { *out = it->second; *Q = s->qnpts[qorder]; return FEEL_HIP_OK; }
    QuadratureRule rule = quadrature(s->dim, qorder);
    Tabulation T = tabulate(s->dim, s->order, rule.pts);
    const int q = rule.npts, d = s->dim;
    std::vector<double> h((size_t)q * (1 + s->nloc + d + 1));
    for (int i = 0; i < q; ++i) h[i] = rule.w[i];
    for (int i = 0; i < s->nloc; ++i)
        for (int k = 0; k < q; ++k) h[q + (size_t)i * q + k] = T.phi[(size_t)i * q + k];
    double* lam = h.data() + q + (size_t)s->nloc * q;
    for (int k = 0; k < q; ++k) {
        double sum = 0;
        for (int c = 0; c < d; ++c) {
            lam[(c + 1) * q + k] = 0.5 * (1.0 + rule.pts[(size_t)k * d + c]);
            sum += lam[(c + 1) * q + k];
        }
        lam[k] = 1.0 - sum;
    }
    double* dptr = nullptr;
    TRY_CALL(upload(s->ctx, &dptr, h.data(), h.size()));
    s->qtab[qorder] = dptr;
    s->qnpts[qorder] = q;
    *out = dptr;
    *Q = q;
    return FEEL_HIP_OK;
}

int ensureGeo(feel_hip_space* s, size_t bytes) {
    if (s->geoBytes >= bytes) return FEEL_HIP_OK;
    s->ctx->release(s->geo, s->geoBytes);
    s->geo = nullptr;
    s->geoBytes = 0;
    TRY_CALL(s->ctx->alloc((void**)&s->geo, bytes));
    s->geoBytes = bytes;
    return FEEL_HIP_OK;
}

}  // namespace

// ============================================================== graph
namespace {

// A vector-space pattern is 3 x 3 blocked when the three rows of every node
// have the same length, hold columns in whole node triples 3j, 3j+1, 3j+2 and
// list the same nodes in the same order.
bool blockedPattern(const std::vector<int32_t>& ia, const std::vector<int32_t>& ja) {
    const int64_t n = (int64_t)ia.size() - 1;
    if (n % 3 != 0) return false;
    for (int64_t i = 0; i < n; i += 3) {
        const int32_t len = ia[i + 1] - ia[i];
        if (len % 3 != 0) return false;
        for (int c = 0; c < 3; ++c) {
            if (ia[i + c + 1] - ia[i + c] != len) return false;
            for (int32_t k = 0; k < len; ++k) {
                const int32_t col = ja[ia[i + c] + k];
                if (col != ja[ia[i] + k] || (k % 3 == 0 ? col % 3 != 0 : col != ja[ia[i] + k - 1] + 1)) return false;
            }
        }
    }
    return true;
}

// diagonal positions (from the host copy of the pattern, rows with their
// diagonal) and the solver's SELL-C-sigma image (row order on the host, the
// image on the device)
int finishGraph(feel_hip_graph* g) {
    feel_hip_ctx* ctx = g->space->ctx;
    const int64_t nrows = g->n;
    const int32_t* ia = g->hia.data();
    const int32_t* ja = g->hja.data();
    std::vector<int32_t> diag(nrows, -1);
    for (int64_t r = 0; r < nrows; ++r) {
        const int32_t* b = ja + ia[r];
        const int32_t* e = ja + ia[r + 1];
        const int32_t* p = std::lower_bound(b, e, (int32_t)r);
        if (p == e || *p != r) return setError(FEEL_HIP_EPATTERN, "graph row without its diagonal entry");
        diag[r] = (int32_t)(p - ja);
    }
    TRY_CALL(upload(ctx, &g->diagpos, diag.data(), diag.size()));
    if (g->space->ncomp == 3 && blockedPattern(g->hia, g->hja)) {
        // vector space: node graph (the first row of each node, every third column)
        const int64_t nb = nrows / 3;
        GraphCSR ng;
        ng.nrows = nb;
        ng.ia.resize(nb + 1);
        ng.ia[0] = 0;
        for (int64_t i = 0; i < nb; ++i) ng.ia[i + 1] = ng.ia[i] + (ia[3 * i + 1] - ia[3 * i]) / 3;
        ng.nnz = ng.ia[nb];
        ng.ja.resize(ng.nnz);
        for (int64_t i = 0; i < nb; ++i)
            for (int32_t p = ia[3 * i], q = ng.ia[i]; p < ia[3 * i + 1]; p += 3, ++q) ng.ja[q] = ja[p] / 3;
        std::vector<int32_t> nodeLen(nb);
        for (int64_t i = 0; i < nb; ++i) nodeLen[i] = ng.ia[i + 1] - ng.ia[i];
        TRY_CALL(buildBlockSellDevice(g, sellPermutation(ng, 1024), nodeLen));
        return FEEL_HIP_OK;
    }
    GraphCSR gc;   // borrows the host pattern for the row ordering
    gc.nrows = nrows; gc.nnz = g->nnz; gc.ia.swap(g->hia); gc.ja.swap(g->hja);
    std::vector<int32_t> perm = sellPermutation(gc, 1024);
    gc.ia.swap(g->hia); gc.ja.swap(g->hja);
    TRY_CALL(buildSellDevice(g, perm));
    return FEEL_HIP_OK;
}

uint64_t nextGraphUid() {
    static std::atomic<uint64_t> next{1};
    return next++;
}

// The rows of g that symmetric elimination by `rows` changes (flagged rows and
// rows holding a flagged column), ascending; listed once per (rows, graph).
int touchedRows(const feel_hip_graph* g, const feel_hip_rows* rows) {
    if (rows->touchedGraph == g->uid) return FEEL_HIP_OK;
    feel_hip_ctx* c = g->space->ctx;
    constexpr int L = 16;
    uint8_t* mark = nullptr;
    TRY_CALL(c->alloc((void**)&mark, (size_t)g->n));
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((g->n + kBlock / L - 1) / (kBlock / L), 16384));
    hipLaunchKernelGGL(kMarkTouched<L>, dim3(grid), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, rows->flag, mark);
    std::vector<uint8_t> h((size_t)g->n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), mark, (size_t)g->n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    c->release(mark, (size_t)g->n);
    if (e != hipSuccess) return setError(FEEL_HIP_EDEVICE, std::string("touched rows: ") + hipGetErrorString(e));
    std::vector<int32_t> list;
    for (int64_t r = 0; r < g->n; ++r)
        if (h[r]) list.push_back((int32_t)r);
    // the list belongs to the row set (its context allocates and frees it)
    if (rows->touched) rows->ctx->release(rows->touched, sizeof(int32_t) * rows->ntouched);
    rows->touched = nullptr;
    rows->ntouched = 0;
    rows->touchedGraph = 0;
    if (!list.empty()) TRY_CALL(upload(rows->ctx, &rows->touched, list.data(), list.size()));
    rows->ntouched = (int64_t)list.size();
    rows->touchedGraph = g->uid;
    return FEEL_HIP_OK;
}

}  // namespace

extern "C" {

int feel_hip_graph_create(feel_hip_space* s, int64_t nrows, const int32_t* ia, const int32_t* ja,
                          feel_hip_graph** out) {
    GUARD_BEGIN
    if (!s || !ia || !ja || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (nrows != s->ndof) return setError(FEEL_HIP_EINVAL, "graph rows must equal the space's dofs");
    if (ia[0] != 0) return setError(FEEL_HIP_EINVAL, "ia[0] must be 0");
    for (int64_t r = 0; r < nrows; ++r) {
        if (ia[r + 1] < ia[r]) return setError(FEEL_HIP_EINVAL, "ia not monotone");
        for (int32_t p = ia[r]; p < ia[r + 1]; ++p) {
            if (ja[p] < 0 || ja[p] >= nrows) return setError(FEEL_HIP_EINVAL, "column index out of range");
            if (p > ia[r] && ja[p] <= ja[p - 1]) return setError(FEEL_HIP_EINVAL, "columns not strictly ascending");
        }
    }
    feel_hip_ctx* ctx = s->ctx;
    FEEL_HIP_CHECK(hipSetDevice(ctx->device));
    auto* g = new feel_hip_graph();
    g->uid = nextGraphUid();
    g->space = s;
    retain(s);
    g->n = nrows;
    g->nnz = ia[nrows];
    g->hia.assign(ia, ia + nrows + 1);
    g->hja.assign(ja, ja + g->nnz);
    int rc = upload(ctx, &g->ia, ia, (size_t)nrows + 1);
    if (rc == FEEL_HIP_OK) rc = upload(ctx, &g->ja, ja, (size_t)g->nnz);
    // gather map of the scatter, built on the device from elem2dof and ia/ja
    if (rc == FEEL_HIP_OK) rc = buildGatherMapDevice(s, g);
    if (rc == FEEL_HIP_OK) rc = finishGraph(g);
    if (rc != FEEL_HIP_OK) {
        feel_hip_graph_destroy(g);
        return rc;
    }
    *out = g;
    return FEEL_HIP_OK;
    GUARD_END
}

int feel_hip_graph_create_from_space(feel_hip_space* s, feel_hip_graph** out) {
    GUARD_BEGIN
    if (!s || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    *out = nullptr;
    feel_hip_ctx* ctx = s->ctx;
    FEEL_HIP_CHECK(hipSetDevice(ctx->device));
    auto* g = new feel_hip_graph();
    g->uid = nextGraphUid();
    g->space = s;
    retain(s);
    int rc = buildGraphDevice(s, g);
    if (rc == FEEL_HIP_OK) {
        g->hia.resize(g->n + 1);
        g->hja.resize(g->nnz);
        if (hipMemcpy(g->hia.data(), g->ia, sizeof(int32_t) * (g->n + 1), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(g->hja.data(), g->ja, sizeof(int32_t) * g->nnz, hipMemcpyDeviceToHost) != hipSuccess)
            rc = setError(FEEL_HIP_EDEVICE, "download of the device-built pattern failed");
    }
    if (rc == FEEL_HIP_OK) rc = finishGraph(g);
    if (rc != FEEL_HIP_OK) {
        feel_hip_graph_destroy(g);
        return rc;
    }
    *out = g;
    return FEEL_HIP_OK;
    GUARD_END
}

int feel_hip_graph_pattern(const feel_hip_graph* g, int32_t* ia, int32_t* ja) {
    if (!g || !ia || !ja) return setError(FEEL_HIP_EINVAL, "null argument");
    std::copy(g->hia.begin(), g->hia.end(), ia);
    std::copy(g->hja.begin(), g->hja.end(), ja);
    return FEEL_HIP_OK;
}

int feel_hip_graph_destroy(feel_hip_graph* g) {
    if (g) unref(g);
    return FEEL_HIP_OK;
}

int64_t feel_hip_graph_nnz(const feel_hip_graph* g) { return g ? g->nnz : -1; }

// ============================================================== matrix
int feel_hip_mat_create(feel_hip_graph* g, int precision, feel_hip_mat** out) {
    if (!g || !out) return setError(FEEL_HIP_EINVAL, "null argument");
    if (precision < 1 || precision > 3) return setError(FEEL_HIP_EINVAL, "bad precision");
    *out = nullptr;
    feel_hip_ctx* c = g->space->ctx;
    auto* A = new feel_hip_mat();
    A->g = g;
    retain(g);
    A->precision = precision;
    int rc = FEEL_HIP_OK;
    // FP32 CSR values are stored only by an FP32-only matrix: the FP32 values
    // of an FP64 + FP32 matrix are its FP64 values rounded, formed where they
    // are needed (the solver's FP32 SELL image, get_values)
    if (precision & FEEL_HIP_FP64) rc = devAlloc(c, &A->v64, (size_t)g->nnz);
    if (rc == FEEL_HIP_OK && precision == FEEL_HIP_FP32) rc = devAlloc(c, &A->v32, (size_t)g->nnz);
    if (rc == FEEL_HIP_OK) rc = devAlloc(c, &A->dinv64, (size_t)g->n);
    if (rc == FEEL_HIP_OK) rc = devAlloc(c, &A->dinv32, (size_t)g->n);
    if (rc != FEEL_HIP_OK) {
        feel_hip_mat_destroy(A);
        return rc;
    }
    *out = A;
    return feel_hip_mat_zero(A);
}

int feel_hip_mat_destroy(feel_hip_mat* A) {
    if (!A) return FEEL_HIP_OK;
    feel_hip_ctx* c = A->g->space->ctx;
    (void)hipStreamSynchronize(c->stream);
    c->release(A->v64, sizeof(double) * A->g->nnz);
    c->release(A->v32, sizeof(float) * A->g->nnz);
    c->release(A->dinv64, sizeof(double) * A->g->n);
    c->release(A->dinv32, sizeof(float) * A->g->n);
    c->release(A->dsq64, sizeof(double) * A->g->n);
    c->release(A->diagFlag, sizeof(int));
    const int64_t n = A->g->n;
    for (double* v : {A->r64, A->z64, A->p64, A->w64}) c->release(v, sizeof(double) * n);
    for (float* v : {A->r32, A->z32, A->p32, A->q32, A->w32, A->x32, A->b32}) c->release(v, sizeof(float) * n);
    for (double* v : {A->bp64, A->xp64}) c->release(v, sizeof(double) * n);
    c->release(A->s64, sizeof(double) * A->g->nslots);
    c->release(A->s32, sizeof(float) * A->g->nslots);
    elemOpFree(A);
    feel_hip_graph* g = A->g;
    delete A;
    unref(g);
    return FEEL_HIP_OK;
}

int feel_hip_mat_zero(feel_hip_mat* A) {
    if (!A) return setError(FEEL_HIP_EINVAL, "null matrix");
    feel_hip_ctx* c = A->g->space->ctx;
    if (A->v64) FEEL_HIP_CHECK(hipMemsetAsync(A->v64, 0, sizeof(double) * A->g->nnz, c->stream));
    if (A->v32) FEEL_HIP_CHECK(hipMemsetAsync(A->v32, 0, sizeof(float) * A->g->nnz, c->stream));
    A->dinvValid = false;
    A->sell64Valid = A->sell32Valid = false;
    elemOpInvalidate(A);
    return FEEL_HIP_OK;
}

int feel_hip_mat_add_matrix(feel_hip_mat* A, int nrows, const int32_t* rows, int ncols, const int32_t* cols,
                            const double* data) {
    if (!A || !rows || !cols || !data || nrows < 0 || ncols < 0) return setError(FEEL_HIP_EINVAL, "bad argument");
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    const int64_t m = (int64_t)nrows * ncols;
    if (m == 0) return FEEL_HIP_OK;
    std::vector<int64_t> pos(m);
    for (int i = 0; i < nrows; ++i) {
        if (rows[i] < 0 || rows[i] >= g->n) return setError(FEEL_HIP_EINVAL, "row out of range");
        const int32_t* b = g->hja.data() + g->hia[rows[i]];
        const int32_t* e = g->hja.data() + g->hia[rows[i] + 1];
        for (int j = 0; j < ncols; ++j) {
            const int32_t* p = std::lower_bound(b, e, cols[j]);
            if (p == e || *p != cols[j]) return setError(FEEL_HIP_EPATTERN, "entry outside the sparsity graph");
            pos[(size_t)i * ncols + j] = p - g->hja.data();
        }
    }
    int64_t* dpos = nullptr;
    double* dval = nullptr;
    TRY_CALL(upload(c, &dpos, pos.data(), pos.size()));
    int rc = upload(c, &dval, data, (size_t)m);
    if (rc == FEEL_HIP_OK) {
        hipLaunchKernelGGL(kAddAt, dim3(1), dim3(1), 0, c->stream, m, dpos, dval, A->v64, A->v32);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = setError(FEEL_HIP_EDEVICE, hipGetErrorString(e));
    }
    c->release(dpos, sizeof(int64_t) * m);
    c->release(dval, sizeof(double) * m);
    A->dinvValid = false;
    A->sell64Valid = A->sell32Valid = false;
    elemOpInvalidate(A);
    return rc;
}

int feel_hip_mat_get_values(feel_hip_mat* A, int precision, void* host) {
    if (!A || !host) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_ctx* c = A->g->space->ctx;
    if (precision == FEEL_HIP_FP64) {
        if (!A->v64) return setError(FEEL_HIP_EINVAL, "matrix holds no FP64 values");
        FEEL_HIP_CHECK(hipMemcpyAsync(host, A->v64, sizeof(double) * A->g->nnz, hipMemcpyDeviceToHost, c->stream));
    } else if (precision == FEEL_HIP_FP32) {
        if (!(A->precision & FEEL_HIP_FP32)) return setError(FEEL_HIP_EINVAL, "matrix holds no FP32 values");
        if (A->v32) {
            FEEL_HIP_CHECK(hipMemcpyAsync(host, A->v32, sizeof(float) * A->g->nnz, hipMemcpyDeviceToHost, c->stream));
        } else {   // the FP64 values rounded
            std::vector<double> v((size_t)A->g->nnz);
            FEEL_HIP_CHECK(hipMemcpyAsync(v.data(), A->v64, sizeof(double) * v.size(), hipMemcpyDeviceToHost, c->stream));
            FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
            float* out = static_cast<float*>(host);
            for (size_t p = 0; p < v.size(); ++p) out[p] = (float)v[p];
        }
    } else {
        return setError(FEEL_HIP_EINVAL, "precision must be FP64 or FP32");
    }
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    return FEEL_HIP_OK;
}

int feel_hip_mat_set_values(feel_hip_mat* A, const double* host) {
    if (!A || !host) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_ctx* c = A->g->space->ctx;
    const int64_t nnz = A->g->nnz;
    if (A->v64) {
        FEEL_HIP_CHECK(hipMemcpyAsync(A->v64, host, sizeof(double) * nnz, hipMemcpyHostToDevice, c->stream));
        if (A->v32) hipLaunchKernelGGL(kCastF32, dim3(gridFor(nnz)), dim3(kBlock), 0, c->stream, nnz, A->v64, A->v32);
    } else {
        std::vector<float> f(host, host + nnz);
        FEEL_HIP_CHECK(hipMemcpyAsync(A->v32, f.data(), sizeof(float) * nnz, hipMemcpyHostToDevice, c->stream));
        FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    FEEL_HIP_CHECK(hipGetLastError());
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    A->dinvValid = false;
    A->sell64Valid = A->sell32Valid = false;
    elemOpInvalidate(A);
    return FEEL_HIP_OK;
}

}  // extern "C"

namespace {

// spmv launch: lanes-per-row chosen from the mean row length
int lanesFor(const feel_hip_graph* g) {
    const double mean = g->n ? (double)g->nnz / (double)g->n : 0.0;
    return mean > 48 ? 16 : (mean > 12 ? 8 : 4);
}

template <typename T>
void launchSpmv(feel_hip_ctx* c, const feel_hip_graph* g, const T* val, const T* x, T* y, double* part,
                const CgState* st, int grid) {
    const int L = lanesFor(g);
    if (L == 16)
        hipLaunchKernelGGL((kSpmv<T, 16>), dim3(grid), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, val, x, y, part, st);
    else if (L == 8)
        hipLaunchKernelGGL((kSpmv<T, 8>), dim3(grid), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, val, x, y, part, st);
    else
        hipLaunchKernelGGL((kSpmv<T, 4>), dim3(grid), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, val, x, y, part, st);
}

int spmvGrid(const feel_hip_graph* g) {
    const int L = lanesFor(g);
    const int64_t groups = (g->n + (kBlock / L) - 1) / (kBlock / L);
    return (int)std::max<int64_t>(1, std::min<int64_t>(groups, 8192));
}

}  // namespace

extern "C" {

int feel_hip_mat_mult(feel_hip_mat* A, const feel_hip_vec* x, feel_hip_vec* y) {
    if (!A || !x || !y) return setError(FEEL_HIP_EINVAL, "null argument");
    if (x->n != A->g->n || y->n != A->g->n) return setError(FEEL_HIP_EINVAL, "size mismatch");
    if (!A->v64) return setError(FEEL_HIP_EINVAL, "matrix holds no FP64 values");
    feel_hip_ctx* c = A->g->space->ctx;
    c->tic("spmv_fp64");
    launchSpmv<double>(c, A->g, A->v64, x->d, y->d, nullptr, nullptr, spmvGrid(A->g));
    c->toc();
    FEEL_HIP_CHECK(hipGetLastError());
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    return FEEL_HIP_OK;
}

// ============================================================== vectors
int feel_hip_vec_create(feel_hip_ctx* c, int64_t n, feel_hip_vec** out) {
    if (!c || !out || n < 0) return setError(FEEL_HIP_EINVAL, "bad argument");
    auto* v = new feel_hip_vec();
    v->ctx = c;
    v->n = n;
    int rc = devAlloc(c, &v->d, (size_t)n);
    if (rc != FEEL_HIP_OK) { delete v; return rc; }
    retain(c);
    if (n) FEEL_HIP_CHECK(hipMemsetAsync(v->d, 0, sizeof(double) * n, c->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    *out = v;
    return FEEL_HIP_OK;
}

int feel_hip_vec_destroy(feel_hip_vec* v) {
    if (!v) return FEEL_HIP_OK;
    (void)hipStreamSynchronize(v->ctx->stream);
    v->ctx->release(v->d, sizeof(double) * v->n);
    feel_hip_ctx* c = v->ctx;
    delete v;
    unref(c);
    return FEEL_HIP_OK;
}

int feel_hip_vec_set(feel_hip_vec* v, const double* host) {
    if (!v || !host) return setError(FEEL_HIP_EINVAL, "null argument");
    FEEL_HIP_CHECK(hipMemcpyAsync(v->d, host, sizeof(double) * v->n, hipMemcpyHostToDevice, v->ctx->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
    return FEEL_HIP_OK;
}

int feel_hip_vec_get(const feel_hip_vec* v, double* host) {
    if (!v || !host) return setError(FEEL_HIP_EINVAL, "null argument");
    FEEL_HIP_CHECK(hipMemcpyAsync(host, v->d, sizeof(double) * v->n, hipMemcpyDeviceToHost, v->ctx->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
    return FEEL_HIP_OK;
}

int feel_hip_vec_zero(feel_hip_vec* v) {
    if (!v) return setError(FEEL_HIP_EINVAL, "null vector");
    if (v->n) FEEL_HIP_CHECK(hipMemsetAsync(v->d, 0, sizeof(double) * v->n, v->ctx->stream));
    return FEEL_HIP_OK;
}

int feel_hip_vec_axpy(feel_hip_vec* y, double a, const feel_hip_vec* x) {
    if (!y || !x || y->n != x->n) return setError(FEEL_HIP_EINVAL, "bad argument");
    feel_hip_ctx* c = y->ctx;
    if (y->n) hipLaunchKernelGGL(kAxpby, dim3(gridFor(y->n)), dim3(kBlock), 0, c->stream, y->n, a, x->d, 1.0, y->d);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

int feel_hip_vec_scale(feel_hip_vec* v, double a) {
    if (!v) return setError(FEEL_HIP_EINVAL, "null vector");
    feel_hip_ctx* c = v->ctx;
    if (v->n) hipLaunchKernelGGL(kAxpby, dim3(gridFor(v->n)), dim3(kBlock), 0, c->stream, v->n, a, v->d, 0.0, v->d);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

int feel_hip_vec_copy(feel_hip_vec* dst, const feel_hip_vec* src) {
    if (!dst || !src || dst->n != src->n) return setError(FEEL_HIP_EINVAL, "bad argument");
    if (dst->n)
        FEEL_HIP_CHECK(hipMemcpyAsync(dst->d, src->d, sizeof(double) * dst->n, hipMemcpyDeviceToDevice, dst->ctx->stream));
    return FEEL_HIP_OK;
}

int feel_hip_vec_set_rows(feel_hip_vec* v, const feel_hip_rows* rows, const feel_hip_vec* values) {
    if (!v || !rows || rows->n != v->n || (values && values->n != v->n)) return setError(FEEL_HIP_EINVAL, "bad argument");
    feel_hip_ctx* c = v->ctx;
    if (rows->nrows)
        hipLaunchKernelGGL(kSetRows, dim3(gridFor(rows->nrows)), dim3(kBlock), 0, c->stream, rows->nrows, rows->rows,
                           values ? values->d : nullptr, v->d);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

int feel_hip_vec_dot(const feel_hip_vec* a, const feel_hip_vec* b, double* out) {
    if (!a || !b || !out || a->n != b->n) return setError(FEEL_HIP_EINVAL, "bad argument");
    feel_hip_ctx* c = a->ctx;
    hipLaunchKernelGGL(kDotPartial, dim3(kReduceBlocks), dim3(kBlock), 0, c->stream, a->n, a->d, b->d, c->part);
    FEEL_HIP_CHECK(hipGetLastError());
    std::vector<double> h(kReduceBlocks);
    FEEL_HIP_CHECK(hipMemcpyAsync(h.data(), c->part, sizeof(double) * kReduceBlocks, hipMemcpyDeviceToHost, c->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    double s = 0;
    for (double x : h) s += x;
    *out = s;
    return FEEL_HIP_OK;
}

int64_t feel_hip_vec_size(const feel_hip_vec* v) { return v ? v->n : -1; }
void* feel_hip_vec_device_ptr(feel_hip_vec* v) { return v ? (void*)v->d : nullptr; }

}  // extern "C"

// ============================================================== assembly
namespace {

// The context's workspace of materialised element matrices, at least bytes.
int ensureElemmat(feel_hip_ctx* c, size_t bytes) {
    if (c->elemmatBytes >= bytes) return FEEL_HIP_OK;
    c->release(c->elemmat, c->elemmatBytes);
    c->elemmat = nullptr;
    c->elemmatBytes = 0;
    TRY_CALL(c->alloc((void**)&c->elemmat, bytes));
    c->elemmatBytes = bytes;
    return FEEL_HIP_OK;
}

// Element matrices on the matrix cores (kElemMatMfma, tile-major into the
// context's workspace), then their segmented sum into the CSR values.
template <int NN, int RPE, int FS, int FR, int RS, int KV, class Entry>
int materialisedPath(feel_hip_mat* A, const double* F, const double* rt, double* v64, float* v32, int add, int grid) {
    feel_hip_graph* g = A->g;
    feel_hip_space* s = g->space;
    feel_hip_ctx* c = s->ctx;
    TRY_CALL(ensureElemmat(c, sizeof(double) * (size_t)MatTiles<NN, RPE>::doubles(s->ne)));
    const int64_t rtiles = (RPE * s->ne + 15) / 16;
    const int mgrid = (int)std::min<int64_t>((rtiles + kBlock / 64 - 1) / (kBlock / 64), 16384);
    hipLaunchKernelGGL((kElemMatMfma<NN, RPE, FS, FR, RS, KV>), dim3(mgrid), dim3(kBlock), 0, c->stream, s->ne, F, rt,
                       c->elemmat);
    hipLaunchKernelGGL(kGatherMaterialised<Entry>, dim3(grid), dim3(kBlock), 0, c->stream, g->ntasks, g->taskSeg,
                       g->taskK, g->centry, g->segSlot, c->elemmat, v64, v32, add);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
}

// Runs the element-matrix path for this assembly (FEEL_HIP_MATRICES_*; 0 =
// the graph's tuned choice).  An untuned graph assembled with add = 0 times
// one run of each path after a warm-up run and keeps the faster; the last run
// leaves valid values.  With add = 1 an untuned graph takes the fused path
// (timing both would accumulate twice).
template <class Fused, class Mfma>
int runElementPath(feel_hip_mat* A, int path, int add, Fused fused, Mfma mfma) {
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    if (path == 0 && add) path = FEEL_HIP_MATRICES_FUSED;
    if (path == FEEL_HIP_MATRICES_MFMA) return mfma();
    if (path == FEEL_HIP_MATRICES_FUSED) return fused();
    hipEvent_t e0, e1;
    FEEL_HIP_CHECK(hipEventCreate(&e0));
    FEEL_HIP_CHECK(hipEventCreate(&e1));
    float ms[2] = {0.0f, 0.0f};
    bool haveMfma = true;
    for (int k = 0; k < 2; ++k) {
        for (int rep = 0; rep < 2; ++rep) {
            if (rep == 1) (void)hipEventRecord(e0, c->stream);
            if (k == 0) {
                TRY_CALL(fused());
            } else if (mfma() != FEEL_HIP_OK) {
                (void)hipGetLastError();
                haveMfma = false;   // workspace does not fit: the fused path stays
                break;
            }
            if (rep == 1) (void)hipEventRecord(e1, c->stream);
        }
        if (k == 1 && !haveMfma) break;
        FEEL_HIP_CHECK(hipEventSynchronize(e1));
        FEEL_HIP_CHECK(hipEventElapsedTime(&ms[k], e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    g->elemPath = haveMfma && ms[1] < ms[0] ? FEEL_HIP_MATRICES_MFMA : FEEL_HIP_MATRICES_FUSED;
    if (!haveMfma) TRY_CALL(fused());
    if (g->elemPath == FEEL_HIP_MATRICES_FUSED) {
        c->release(c->elemmat, c->elemmatBytes);
        c->elemmat = nullptr;
        c->elemmatBytes = 0;
    }
    return FEEL_HIP_OK;
}

}  // namespace

extern "C" {

int feel_hip_assemble_bilinear(feel_hip_mat* A, const feel_hip_bilinear_desc* d, int add) {
    GUARD_BEGIN
    if (!A || !d) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_graph* g = A->g;
    feel_hip_space* s = g->space;
    feel_hip_ctx* c = s->ctx;
    const int wantPrec = d->precision ? d->precision : A->precision;
    // an FP64 + FP32 matrix keeps FP32 values as its FP64 values rounded, so
    // either request assembles its FP64 values
    double* v64 = (wantPrec & FEEL_HIP_FP64) || !A->v32 ? A->v64 : nullptr;
    float* v32 = (wantPrec & FEEL_HIP_FP32) ? A->v32 : nullptr;
    if (!v64 && !v32) return setError(FEEL_HIP_EINVAL, "matrix holds none of the requested precisions");
    elemOpInvalidate(A);
    if (d->kind & FEEL_HIP_EXPR_ELASTICITY) {
        if (d->kind != FEEL_HIP_EXPR_ELASTICITY)
            return setError(FEEL_HIP_EUNSUPPORTED, "elasticity cannot be combined with other terms in one call");
        if (s->dim != 3 || s->ncomp != 3) return setError(FEEL_HIP_EUNSUPPORTED, "elasticity needs Pchv on tetrahedra");
        if (!d->young || !d->poisson || d->young->n != s->ne || d->poisson->n != s->ne)
            return setError(FEEL_HIP_EINVAL, "elasticity needs E and nu per element");
        const double* rt = nullptr;
        TRY_CALL(fullTensors(s, d->quad_order, &rt));
        TRY_CALL(ensureGeo(s, sizeof(double) * kElasticF * s->ne));
        c->tic("assemble_geo");
        hipLaunchKernelGGL(kGeoElastic, dim3(gridFor(s->ne)), dim3(kBlock), 0, c->stream, s->ne, s->coords, s->elems,
                           d->young->d, d->poisson->d, s->geo);
        ++s->geoStamp;
        c->toc();
        FEEL_HIP_CHECK(hipGetLastError());
        const int grid = segGrid(g);
        TRY_CALL(clearEmptySlots(c, g, v64, v32, add));
        if (s->order < 1 || s->order > 4) return setError(FEEL_HIP_EUNSUPPORTED, "no elasticity kernel for this order");
        int path = d->matrices;
        if (path < FEEL_HIP_MATRICES_AUTO || path > FEEL_HIP_MATRICES_MFMA)
            return setError(FEEL_HIP_EINVAL, "unknown element-matrix path");
        if (path == FEEL_HIP_MATRICES_MFMA && s->order < 3)
            return setError(FEEL_HIP_EUNSUPPORTED, "matrix-core element matrices are built for P3 and P4");
        if (path == FEEL_HIP_MATRICES_AUTO) path = s->order < 3 ? FEEL_HIP_MATRICES_FUSED : g->elemPath;
        auto fused = [&]() -> int {
            switch (s->order) {
                case 1: hipLaunchKernelGGL(kGatherElastic<4>, dim3(grid), dim3(kBlock), 0, c->stream, SEG_ARGS); break;
                case 2: hipLaunchKernelGGL(kGatherElastic<10>, dim3(grid), dim3(kBlock), 0, c->stream, SEG_ARGS); break;
                case 3: hipLaunchKernelGGL(kGatherElastic<20>, dim3(grid), dim3(kBlock), 0, c->stream, SEG_ARGS); break;
                default:
                    hipLaunchKernelGGL((kGatherElastic<35, false>), dim3(grid), dim3(kBlock), 0, c->stream, SEG_ARGS);
            }
            FEEL_HIP_CHECK(hipGetLastError());
            return FEEL_HIP_OK;
        };
        // elasticity GEMM: 9 rows per element, factors (c*3+d)*10 + t, tables (i*n+j)*10 + t, t < 9
        auto mfma = [&]() -> int {
            if (s->order == 3)
                return materialisedPath<400, 9, kElasticF, 10, 10, 9, ElasticEntry<20>>(A, s->geo, rt, v64, v32, add,
                                                                                       grid);
            return materialisedPath<1225, 9, kElasticF, 10, 10, 9, ElasticEntry<35>>(A, s->geo, rt, v64, v32, add,
                                                                                    grid);
        };
        c->tic("assemble_gather");
        TRY_CALL(runElementPath(A, path, add, fused, mfma));
        c->toc();
        FEEL_HIP_CHECK(hipGetLastError());
        if (v64 && A->v32 && !v32) hipLaunchKernelGGL(kCastF32, dim3(gridFor(g->nnz)), dim3(kBlock), 0, c->stream, g->nnz, v64, A->v32);
        FEEL_HIP_CHECK(hipGetLastError());
        A->dinvValid = false;
        A->sell64Valid = A->sell32Valid = false;
    elemOpInvalidate(A);
        return FEEL_HIP_OK;
    }
    if ((d->kind & ~(FEEL_HIP_EXPR_MASS | FEEL_HIP_EXPR_DIFFUSION)) || d->kind == 0)
        return setError(FEEL_HIP_EUNSUPPORTED, "expression outside the supported subset");
    if (s->ncomp != 1) return setError(FEEL_HIP_EUNSUPPORTED, "mass/diffusion need a scalar space");
    auto coefPtr = [&](const feel_hip_vec* v, const char* what, const double** out) -> int {
        *out = nullptr;
        if (!v) return FEEL_HIP_OK;
        if (v->n != s->ne) return setError(FEEL_HIP_EINVAL, std::string(what) + " must have one value per element");
        *out = v->d;
        return FEEL_HIP_OK;
    };
    const double *kd = nullptr, *km = nullptr;
    TRY_CALL(coefPtr(d->diff_coef, "diffusion coefficient", &kd));
    TRY_CALL(coefPtr(d->mass_coef, "mass coefficient", &km));
    const double kds = (d->kind & FEEL_HIP_EXPR_DIFFUSION) ? d->diff_scale : 0.0;
    const double kms = (d->kind & FEEL_HIP_EXPR_MASS) ? d->mass_scale : 0.0;
    const double* rt = nullptr;
    TRY_CALL(scalarTensors(s, d->quad_order, &rt));
    const int stride = s->dim == 3 ? 8 : 4;
    TRY_CALL(ensureGeo(s, sizeof(double) * stride * s->ne));
    c->tic("assemble_geo");
    if (s->dim == 3)
        hipLaunchKernelGGL(kGeoScalar<3>, dim3(gridFor(s->ne)), dim3(kBlock), 0, c->stream, s->ne, s->coords, s->elems, kd,
                           kds, km, kms, s->geo);
    else
        hipLaunchKernelGGL(kGeoScalar<2>, dim3(gridFor(s->ne)), dim3(kBlock), 0, c->stream, s->ne, s->coords, s->elems, kd,
                           kds, km, kms, s->geo);
    ++s->geoStamp;
    c->toc();
    FEEL_HIP_CHECK(hipGetLastError());
    const int grid = segGrid(g);
    TRY_CALL(clearEmptySlots(c, g, v64, v32, add));
    const int key = s->dim * 10 + s->order;
    if (key != 21 && key != 22 && (key < 31 || key > 34))
        return setError(FEEL_HIP_EUNSUPPORTED, "no scatter kernel for this space");
    int path = d->matrices;
    if (path < FEEL_HIP_MATRICES_AUTO || path > FEEL_HIP_MATRICES_MFMA)
        return setError(FEEL_HIP_EINVAL, "unknown element-matrix path");
    const bool dense = key == 33 || key == 34;   // P3/P4 tetrahedra
    if (path == FEEL_HIP_MATRICES_MFMA && !dense)
        return setError(FEEL_HIP_EUNSUPPORTED, "matrix-core element matrices are built for P3 and P4 tetrahedra");
    if (path == FEEL_HIP_MATRICES_AUTO) path = dense ? g->elemPath : FEEL_HIP_MATRICES_FUSED;
    auto fused = [&]() -> int {
#define GATHER(NN, ST) hipLaunchKernelGGL((kGatherScalar<NN, ST>), dim3(grid), dim3(kBlock), 0, c->stream, SEG_ARGS)
        switch (key) {
            case 21: GATHER(9, 4); break;
            case 22: GATHER(36, 4); break;
            case 31: GATHER(16, 8); break;
            case 32: GATHER(100, 8); break;
            case 33: GATHER(400, 8); break;
            default:
                hipLaunchKernelGGL((kGatherScalar<1225, 8, false>), dim3(grid), dim3(kBlock), 0, c->stream, SEG_ARGS);
        }
#undef GATHER
        FEEL_HIP_CHECK(hipGetLastError());
        return FEEL_HIP_OK;
    };
    // scalar GEMM: one row per element, factors e*8 + t, tables (i*n+j)*8 + t, t < 7
    auto mfma = [&]() -> int {
        if (key == 33)
            return materialisedPath<400, 1, 8, 0, 8, 7, ScalarEntry<20>>(A, s->geo, rt, v64, v32, add, grid);
        return materialisedPath<1225, 1, 8, 0, 8, 7, ScalarEntry<35>>(A, s->geo, rt, v64, v32, add, grid);
    };
    c->tic("assemble_gather");
    TRY_CALL(runElementPath(A, path, add, fused, mfma));
    c->toc();
    FEEL_HIP_CHECK(hipGetLastError());
    if (v64 && A->v32 && !v32) hipLaunchKernelGGL(kCastF32, dim3(gridFor(g->nnz)), dim3(kBlock), 0, c->stream, g->nnz, v64, A->v32);
    FEEL_HIP_CHECK(hipGetLastError());
    A->dinvValid = false;
    A->sell64Valid = A->sell32Valid = false;
    // diffusion alone, operator=, a rule exact for the degree-2 integrand of P2
    TRY_CALL(elemOpAssembled(A, d->kind == FEEL_HIP_EXPR_DIFFUSION && !add && d->quad_order >= 2 && v64));
    return FEEL_HIP_OK;
    GUARD_END
}

int feel_hip_assemble_linear(feel_hip_space* s, const feel_hip_linear_desc* d, feel_hip_vec* b, int add) {
    GUARD_BEGIN
    if (!s || !d || !b) return setError(FEEL_HIP_EINVAL, "null argument");
    if (b->n != s->ndof) return setError(FEEL_HIP_EINVAL, "rhs size must equal the space's dofs");
    if (d->kind < 0 || d->kind > 2) return setError(FEEL_HIP_EUNSUPPORTED, "unknown source kind");
    if (d->kind == FEEL_HIP_SRC_VECTOR && s->ncomp != 3) return setError(FEEL_HIP_EINVAL, "vector source needs Pchv");
    if (d->kind != FEEL_HIP_SRC_VECTOR && s->ncomp != 1) return setError(FEEL_HIP_EINVAL, "scalar source needs Pch");
    feel_hip_ctx* c = s->ctx;
    const double* qt = nullptr;
    int Q = 0;
    TRY_CALL(rhsTable(s, d->quad_order, &qt, &Q));
    if (s->dim == 3)
        hipLaunchKernelGGL(kElemRhs<3>, dim3(gridFor(s->ne)), dim3(kBlock), 0, c->stream, s->ne, s->nloc, s->ncomp, Q,
                           s->coords, s->elems, qt, d->kind, d->params[0], d->params[1], d->params[2], s->elemvec);
    else
        hipLaunchKernelGGL(kElemRhs<2>, dim3(gridFor(s->ne)), dim3(kBlock), 0, c->stream, s->ne, s->nloc, s->ncomp, Q,
                           s->coords, s->elems, qt, d->kind, d->params[0], d->params[1], d->params[2], s->elemvec);
    FEEL_HIP_CHECK(hipGetLastError());
    constexpr int L = 8;   // lanes per row
    hipLaunchKernelGGL(kGatherVec<L>, dim3((int)std::min<int64_t>((s->ndof + kBlock / L - 1) / (kBlock / L), 16384)),
                       dim3(kBlock), 0, c->stream, s->ndof, s->rptr, s->rentry, s->elemvec, b->d, add);
    FEEL_HIP_CHECK(hipGetLastError());
    return FEEL_HIP_OK;
    GUARD_END
}

// ============================================================== Dirichlet
int feel_hip_rows_create(feel_hip_ctx* c, int64_t n, int64_t nrows, const int32_t* rows, feel_hip_rows** out) {
    if (!c || !out || (nrows && !rows) || n <= 0 || nrows < 0) return setError(FEEL_HIP_EINVAL, "bad argument");
    std::vector<uint8_t> flag(n, 0);
    for (int64_t i = 0; i < nrows; ++i) {
        if (rows[i] < 0 || rows[i] >= n) return setError(FEEL_HIP_EINVAL, "row index out of range");
        flag[rows[i]] = 1;
    }
    auto* r = new feel_hip_rows();
    r->ctx = c;
    retain(c);
    r->n = n;
    r->nrows = nrows;
    std::vector<int32_t> flagged;
    for (int64_t i = 0; i < n; ++i)
        if (flag[i]) flagged.push_back((int32_t)i);
    r->nflagged = (int64_t)flagged.size();
    int rc = upload(c, &r->rows, rows, (size_t)nrows);
    if (rc == FEEL_HIP_OK) rc = upload(c, &r->flag, flag.data(), flag.size());
    if (rc == FEEL_HIP_OK && r->nflagged) rc = upload(c, &r->flagged, flagged.data(), flagged.size());
    if (rc != FEEL_HIP_OK) { feel_hip_rows_destroy(r); return rc; }
    *out = r;
    return FEEL_HIP_OK;
}

int feel_hip_rows_destroy(feel_hip_rows* r) {
    if (!r) return FEEL_HIP_OK;
    (void)hipStreamSynchronize(r->ctx->stream);
    r->ctx->release(r->rows, sizeof(int32_t) * r->nrows);
    r->ctx->release(r->flag, r->n);
    if (r->flagged) r->ctx->release(r->flagged, sizeof(int32_t) * r->nflagged);
    if (r->touched) r->ctx->release(r->touched, sizeof(int32_t) * r->ntouched);
    feel_hip_ctx* c = r->ctx;
    delete r;
    unref(c);
    return FEEL_HIP_OK;
}

int feel_hip_mat_zero_rows(feel_hip_mat* A, const feel_hip_rows* rows, const feel_hip_vec* values, feel_hip_vec* rhs,
                           int mode, double diag) {
    if (!A || !rows) return setError(FEEL_HIP_EINVAL, "null argument");
    feel_hip_graph* g = A->g;
    if (rows->n != g->n) return setError(FEEL_HIP_EINVAL, "row set built for another size");
    if (values && values->n != g->n) return setError(FEEL_HIP_EINVAL, "values size mismatch");
    if (rhs && rhs->n != g->n) return setError(FEEL_HIP_EINVAL, "rhs size mismatch");
    if (mode != FEEL_HIP_ON_ELIMINATION && mode != FEEL_HIP_ON_ELIMINATION_SYMMETRIC)
        return setError(FEEL_HIP_EINVAL, "unknown elimination mode");
    feel_hip_ctx* c = g->space->ctx;
    constexpr int L = 16;   // lanes per row
    const double* gv = values ? values->d : nullptr;
    double* b = rhs ? rhs->d : nullptr;
    const int sym = mode == FEEL_HIP_ON_ELIMINATION_SYMMETRIC;
    // only the rows the call changes: the flagged ones, and with symmetric
    // elimination the rows holding a flagged column (listed once per graph)
    if (sym) TRY_CALL(touchedRows(g, rows));
    const int32_t* list = sym ? rows->touched : rows->flagged;
    const int64_t count = sym ? rows->ntouched : rows->nflagged;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((count + kBlock / L - 1) / (kBlock / L), 16384));
    if ((A->v64 || A->v32) && count > 0)
        hipLaunchKernelGGL(kDirichletRows<L>, dim3(grid), dim3(kBlock), 0, c->stream, count, list, g->ia, g->ja,
                           A->v64, A->v32, b, rows->flag, gv, sym, diag);
    FEEL_HIP_CHECK(hipGetLastError());
    A->dinvValid = false;
    A->sell64Valid = A->sell32Valid = false;
    TRY_CALL(elemOpZeroRows(A, rows, mode, diag));
    return FEEL_HIP_OK;
}

}  // extern "C"
