// Internal definitions shared by the device-layer translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "feelpp_hip.h"

namespace feelpp_amd {

int setError(int code, const std::string& msg);

#define FEEL_HIP_CHECK(expr)                                                                   \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::feelpp_amd::setError(FEEL_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Solver state shared by the CG kernels; written only by block 0 of a kernel,
// read by the following kernels (kernel boundaries order the accesses).
struct CgState {
    double rtol, atol, dtol;
    double ttol, rnorm0, dp, beta, dpi_prev;
    int maxit, its, done, reason;
    int iter;       // kernel-side iteration counter (number of p-updates launched and not skipped)
    int fresh;      // iteration at which the search direction (re)starts as p = z
    int pad[2];
};

constexpr int kReduceBlocks = 1024;   // grid of the reduction-carrying vector kernels
constexpr int kBlock = 256;
constexpr int kSpmvGrid = 2048;       // default grid of the solver SpMV (one p.Ap partial per block)
constexpr int kSpmvPartCap = 4096;    // largest SpMV grid (partials buffer capacity)
// the context's partials buffer: z.r and z.z (set A), FP64 residual, SpMV
// p.Ap, SpMV w.D^{-1}w, z.r and z.z (set B: the pipelined FP32 iteration
// reads one set and writes the other)
constexpr int kPartDoubles = 5 * kReduceBlocks + 2 * kSpmvPartCap;
constexpr int kSpmvBlock = 1024;
// the fused CG step kernel's launch: 512 blocks of 512 threads read the
// prologue's partials half as often as 1024 x 256 (C2 mixed +0.75 %; 256 x 1024
// measured no better than 1024 x 256)
constexpr int kStepBlock = 512;
constexpr int kStepBlocks = 512;
constexpr int kHistMax = 8192;        // CG coefficient history kept for the condition estimate
constexpr int kElasticF = 90;         // elasticity factors per element: 9 blocks of 9 + pad
constexpr int kSegTask = 256;         // scatter entries per task (whole slots)
// 3 x 3 block image: lanes per node row, chosen per slice: 2 (32 rows per
// slice) unless the slice's longest row would then take more than
// kBlockStepCap steps, then 4, 8 or 16 (16, 8 or 4 rows per slice).
// Measured on the P3/P4 beams (cap, largest L): (48, 8), (32, 16), (24, 16),
// (32, 8).
constexpr int kBlockMinShift = 1, kBlockMaxShift = 4;
constexpr int kBlockStepCap = 32;

}  // namespace feelpp_amd

namespace feelpp_amd {
// gather map of the scatter for (space, graph), built on the device (gathermap.hip)
int buildGatherMapDevice(feel_hip_space* s, feel_hip_graph* g);
// the space's COUPLED pattern (ia/ja) and the gather map from one sort (gathermap.hip)
int buildGraphDevice(feel_hip_space* s, feel_hip_graph* g);
// first-touch elem2dof and ndof of the space's mesh; dof incidence (doftable.hip)
int buildDofTableDevice(feel_hip_space* s);
int buildIncidenceDevice(feel_hip_space* s);
// the solver's SELL-C-sigma image of g for the row order perm (sell.hip)
int buildSellDevice(feel_hip_graph* g, const std::vector<int32_t>& perm);
// the 3 x 3 block image of a vector-space graph for the node order nperm (sell.hip)
int buildBlockSellDevice(feel_hip_graph* g, const std::vector<int32_t>& nperm, const std::vector<int32_t>& nodeLen);
// device address of the static staging array of feel_hip_mat_add_matrix_noalloc,
// looked up once per device with that device current (noalloc.hip)
int noallocSetup(int device);
// element-by-element FP32 operator of P2 diffusion matrices (elemop.hip)
struct ElemOp;
int elemOpAssembled(feel_hip_mat* A, bool qualifies);
int elemOpZeroRows(feel_hip_mat* A, const feel_hip_rows* rows, int mode, double diag);
void elemOpInvalidate(feel_hip_mat* A);
bool elemOpReady(const feel_hip_mat* A);
int elemOpPrepare(feel_hip_mat* A);
void elemOpFree(feel_hip_mat* A);
int elemOpApply(feel_hip_mat* A, const float* p, float* w, double* part, const CgState* st, int* npart);
void retain(feel_hip_ctx* c);
void retain(feel_hip_space* s);
void retain(feel_hip_graph* g);
void unref(feel_hip_ctx* c);
void unref(feel_hip_space* s);
void unref(feel_hip_graph* g);
}  // namespace feelpp_amd

// Lifetime: a context, space or graph is reference counted.  Its creator holds
// one reference and every object built on it (space -> ctx, graph -> space,
// matrix -> graph, vector/rows/batch -> ctx, batch family -> graph) holds
// another, so the *_destroy calls may come in any order; the memory goes when
// the last reference does.  Counts are not atomic: one host thread per context.
struct feel_hip_ctx {
    int refs = 1;
    int device = 0;
    int numCUs = 0;
    hipStream_t stream = nullptr;
    int64_t bytes = 0;
    // timing
    bool timing = false;
    int period = 1;                           // time every period-th launch of each timer
    bool pending = false;                     // the last tic recorded its start event
    std::map<std::string, int64_t> timerCalls;
    struct Ev { hipEvent_t a, b; std::string name; };
    std::vector<Ev> events;
    size_t evUsed = 0;
    std::map<std::string, std::pair<int64_t, double>> timers;
    // CG scratch: partial sums (3 arrays of kReduceBlocks doubles) and state
    double* part = nullptr;   // kPartDoubles doubles
    feelpp_amd::CgState* state = nullptr;
    feelpp_amd::CgState* hstate = nullptr;   // pinned host mirror
    double* hist = nullptr;   // 2 x kHistMax CG coefficients (beta, p.Ap) of the running solve
    // element matrices materialised by the matrix-core assembly path; shared by
    // every assembly on this context (they are ordered on its stream), grown
    // to the largest request and kept until the context goes
    double* elemmat = nullptr;
    size_t elemmatBytes = 0;

    int alloc(void** p, size_t bytes);
    void release(void* p, size_t bytes);
    void tic(const char* name);   // records the start event of a timed launch
    void toc();                   // records the stop event
    void collectTimers();
};

struct feel_hip_space {
    int refs = 1;
    feel_hip_ctx* ctx = nullptr;
    int dim = 0, order = 0, ncomp = 1, nloc = 0, nldof = 0;
    int64_t nv = 0, ne = 0, ndof = 0;
    double* coords = nullptr;
    int32_t* elems = nullptr;
    int32_t* e2d = nullptr;
    std::vector<int32_t> hElem2dof;   // kept to build scatter maps
    // dof -> (element, local) incidence, element order, for the linear-form gather
    uint32_t* rptr = nullptr;
    uint32_t* rentry = nullptr;
    // per-element scratch (geometric factors / element vectors)
    double* geo = nullptr;
    size_t geoBytes = 0;
    uint64_t geoStamp = 0;        // bumped by every assembly that rewrites geo
    double* elemvec = nullptr;
    // reference tensors per quadrature order: [Rsym(NS) | M | m] device copies
    std::map<int, double*> rtab;
    std::map<int, std::vector<double>> hrtab;
    // per quadrature order: weights, basis values, P1 geometry basis values
    std::map<int, double*> qtab;
    std::map<int, int> qnpts;
    // per quadrature order, for the norms: weights, values, gradients, P1 geometry values
    std::map<int, std::pair<double*, size_t>> ntab;   // device table, doubles
    std::map<int, int> nnpts;
};

struct feel_hip_graph {
    int refs = 1;
    uint64_t uid = 0;   // process-unique, never reused (keys per-graph caches)
    feel_hip_space* space = nullptr;
    int64_t n = 0, nnz = 0;
    int32_t* ia = nullptr;
    int32_t* ja = nullptr;
    int32_t* diagpos = nullptr;
    // gather map: per CSR slot (in slot order) the element-matrix entry ids
    // e*nldof^2 + i*nldof + j landing in it, ascending element order; bit 31
    // marks a slot's last entry.  Tasks: runs of whole slots for the scatter.
    uint32_t* centry = nullptr;   // ncontrib
    uint32_t* taskSeg = nullptr;  // ntasks+1: first segment (non-empty slot) of each task
    uint32_t* taskK = nullptr;    // ntasks+1: first entry of each task
    int32_t* segSlot = nullptr;   // nseg: CSR slot of each segment, only when some slot has no entry
    int64_t ntasks = 0, nseg = 0;
    // SpMV launch shape picked by the autotuner, per precision (0 FP64, 1 FP32):
    // index into the candidate list, -1 until the first solve measures them
    int spmvShape[2] = {-1, -1};
    // elasticity element-matrix path picked by timing both (FEEL_HIP_MATRICES_*),
    // 0 until the first assembly with add = 0 measures them
    int elemPath = 0;
    int64_t ncontrib = 0;
    std::vector<int32_t> hia, hja;
    // SELL-C-sigma image used by the solver (permuted numbering)
    int64_t nslices = 0, nslots = 0;
    uint32_t* slicePtr = nullptr;
    int32_t* scol = nullptr;
    int32_t* slot2csr = nullptr;
    int32_t* perm = nullptr;      // new -> old row
    int32_t* sdiag = nullptr;     // diagonal slot per new row
    bool compact = false;         // packed 16/32-bit column stream available
    bool block = false;           // 3 x 3 block image (vector spaces): slicePtr/scol in block slots
    int64_t nbslots = 0;          // block slots of the block image (nslots = 9 nbslots value slots)
    int32_t* sliceRow = nullptr;  // block image: first node row of each slice (nslices + 1)
    uint8_t* sliceShift = nullptr;  // block image: log2 of the lanes per node row of each slice
    int32_t* colBase = nullptr;   // nslices, -1 = slice stores int32 columns
    int64_t* ccptr = nullptr;     // nslices, slice start in packed (uint16 units)
    uint16_t* packed = nullptr;   // npacked
    int64_t npacked = 0;
    // transpose map of a structurally symmetric pattern (built on first use):
    // tpos[p] = CSR position of (j, i) for the entry p = (i, j)
    int32_t* tpos = nullptr;
    int symmetric = 0;            // 0 unknown, 1 structurally symmetric, -1 not
};

struct feel_hip_mat {
    feel_hip_graph* g = nullptr;
    int precision = FEEL_HIP_FP64;
    double* v64 = nullptr;
    float* v32 = nullptr;
    double* dinv64 = nullptr;
    float* dinv32 = nullptr;
    bool dinvValid = false;
    // SELL values for the solver, refreshed from the CSR values when stale
    double* s64 = nullptr;
    float* s32 = nullptr;
    bool sell64Valid = false, sell32Valid = false;
    // scalar image: s32 holds D^{-1/2} A D^{-1/2} (rounded from the FP64
    // values; every diagonal entry positive), dsq64 = D^{-1/2}
    bool s32Scaled = false;
    double* dsq64 = nullptr;
    int* diagFlag = nullptr;   // set by kDiagSqrt when a diagonal entry is not positive
    // solver work vectors, allocated on first use and kept with the matrix
    double *r64 = nullptr, *z64 = nullptr, *p64 = nullptr, *w64 = nullptr;
    float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *w32 = nullptr, *x32 = nullptr, *b32 = nullptr;
    float* q32 = nullptr;   // second search-direction buffer of the pipelined FP32 iteration
    double *bp64 = nullptr, *xp64 = nullptr;   // b and x in the solver numbering
    // element-by-element FP32 operator: 1 while the values are a P2 diffusion
    // assembly followed only by symmetric unit-diagonal eliminations
    int mfState = 0;
    uint64_t geoStamp = 0;        // the space's geoStamp after this matrix's last assembly
    uint8_t* mfRows = nullptr;    // Dirichlet rows of its symmetric eliminations (Feel++ numbering)
    feelpp_amd::ElemOp* eop = nullptr;
};

struct feel_hip_vec {
    feel_hip_ctx* ctx = nullptr;
    int64_t n = 0;
    double* d = nullptr;
};

struct feel_hip_rows {
    feel_hip_ctx* ctx = nullptr;
    int64_t n = 0, nrows = 0;
    int32_t* rows = nullptr;
    uint8_t* flag = nullptr;
    int32_t* flagged = nullptr;   // the distinct flagged rows, ascending
    int64_t nflagged = 0;
    // rows that symmetric elimination changes on one graph (flagged rows and
    // rows holding a flagged column), built on the first such call there
    mutable uint64_t touchedGraph = 0;
    mutable int32_t* touched = nullptr;
    mutable int64_t ntouched = 0;
};
