// MatrixSparse::addMatrix(rows, nrows, cols, ncols, data) (ADD semantics,
// feelalg/matrixsparse.hpp:496-498) for callers that cannot allocate device
// memory: the staging of rows, columns, values and CSR positions, and the
// error flag, live in one static device array per GPU instead of buffers
// allocated per call.  The array is held until the process ends; calls that
// overlap on one GPU share it and race, so they must be ordered (one context
// per GPU, or external ordering between contexts of one GPU).
#include <mutex>
#include <string>

#include "common.hpp"
#include "kernels.hpp"

using namespace feelpp_amd;

// capacity of one call: nrows * ncols entries, and nrows, ncols each
constexpr int64_t kNoAllocMaxEntries = 65536;
constexpr int64_t kNoAllocMaxIndices = 65536;
// layout (bytes): positions int64[E] | values f64[E] | rows i32[I] | cols i32[I] | flag i32
constexpr size_t kOffPos = 0;
constexpr size_t kOffVal = kOffPos + sizeof(int64_t) * kNoAllocMaxEntries;
constexpr size_t kOffRows = kOffVal + sizeof(double) * kNoAllocMaxEntries;
constexpr size_t kOffCols = kOffRows + sizeof(int32_t) * kNoAllocMaxIndices;
constexpr size_t kOffFlag = kOffCols + sizeof(int32_t) * kNoAllocMaxIndices;
constexpr size_t kScratchBytes = kOffFlag + 256;   // 1.5 MiB + 256 B

// The staging array.  Host code never uses this name as an address: it holds
// the per-device address from hipGetSymbolAddress (gScratch).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// device address of `scratch` on each GPU (each holds its own copy)
constexpr int kMaxDevices = 64;
void* gScratch[kMaxDevices] = {};
std::mutex gScratchMutex;

enum : int32_t { kBadRow = 1, kOutsidePattern = 2 };

__global__ void kClearFlag(int32_t* __restrict__ flag) {
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 0;
}

// CSR position of entry (rows[i], cols[j]) by binary search in the row's
// ascending columns; flags a row out of range or an entry outside the graph.
__global__ __launch_bounds__(kBlock) void kLocateEntries(int64_t n, const int32_t* __restrict__ ia,
                                                         const int32_t* __restrict__ ja, int nrows,
                                                         const int32_t* __restrict__ rows, int ncols,
                                                         const int32_t* __restrict__ cols, int64_t* __restrict__ pos,
                                                         int32_t* __restrict__ flag) {
    const int64_t m = (int64_t)nrows * ncols;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = k / ncols;
        const int32_t r = rows[i], c = cols[k - i * ncols];
        if (r < 0 || r >= n) {
            pos[k] = -1;
            atomicOr(flag, kBadRow);
            continue;
        }
        int32_t lo = ia[r], hi = ia[r + 1];
        while (lo < hi) {
            const int32_t mid = lo + ((hi - lo) >> 1);
            if (ja[mid] < c) lo = mid + 1;
            else hi = mid;
        }
        if (lo == ia[r + 1] || ja[lo] != c) {
            pos[k] = -1;
            atomicOr(flag, kOutsidePattern);
        } else {
            pos[k] = lo;
        }
    }
}

}  // namespace

namespace feelpp_amd {

// Once per device, with that device current (feel_hip_ctx_create): the
// device address of the static staging array.
int noallocSetup(int device) {
    if (device < 0 || device >= kMaxDevices) return setError(FEEL_HIP_EINVAL, "device id beyond the staging table");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (gScratch[device]) return FEEL_HIP_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return setError(FEEL_HIP_EDEVICE, std::string("hipGetSymbolAddress(staging array): ") + hipGetErrorString(e));
    gScratch[device] = ptr;
    return FEEL_HIP_OK;
}

}  // namespace feelpp_amd

extern "C" int feel_hip_mat_add_matrix_noalloc(feel_hip_mat* A, int nrows, const int32_t* rows, int ncols,
                                               const int32_t* cols, const double* data) {
    if (!A || !rows || !cols || !data || nrows < 0 || ncols < 0) return setError(FEEL_HIP_EINVAL, "bad argument");
    const int64_t m = (int64_t)nrows * ncols;
    if (m == 0) return FEEL_HIP_OK;
    if (m > kNoAllocMaxEntries || nrows > kNoAllocMaxIndices || ncols > kNoAllocMaxIndices)
        return setError(FEEL_HIP_EINVAL, "add_matrix_noalloc: " + std::to_string(nrows) + " x " + std::to_string(ncols) +
                                             " entries exceed the staging capacity of " +
                                             std::to_string(kNoAllocMaxEntries) + " entries");
    feel_hip_graph* g = A->g;
    feel_hip_ctx* c = g->space->ctx;
    if (c->device < 0 || c->device >= kMaxDevices || !gScratch[c->device])
        return setError(FEEL_HIP_EDEVICE, "add_matrix_noalloc: staging array not set up on this device");
    unsigned char* base = static_cast<unsigned char*>(gScratch[c->device]);
    int64_t* dpos = reinterpret_cast<int64_t*>(base + kOffPos);
    double* dval = reinterpret_cast<double*>(base + kOffVal);
    int32_t* drows = reinterpret_cast<int32_t*>(base + kOffRows);
    int32_t* dcols = reinterpret_cast<int32_t*>(base + kOffCols);
    int32_t* dflag = reinterpret_cast<int32_t*>(base + kOffFlag);
    FEEL_HIP_CHECK(hipMemcpyAsync(drows, rows, sizeof(int32_t) * nrows, hipMemcpyHostToDevice, c->stream));
    FEEL_HIP_CHECK(hipMemcpyAsync(dcols, cols, sizeof(int32_t) * ncols, hipMemcpyHostToDevice, c->stream));
    FEEL_HIP_CHECK(hipMemcpyAsync(dval, data, sizeof(double) * m, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(kClearFlag, dim3(1), dim3(64), 0, c->stream, dflag);
    const int grid = (int)std::min<int64_t>((m + kBlock - 1) / kBlock, 1024);
    hipLaunchKernelGGL(kLocateEntries, dim3(grid), dim3(kBlock), 0, c->stream, g->n, g->ia, g->ja, nrows, drows, ncols,
                       dcols, dpos, dflag);
    int32_t flag = 0;
    FEEL_HIP_CHECK(hipMemcpyAsync(&flag, dflag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (flag & kBadRow) return setError(FEEL_HIP_EINVAL, "row out of range");
    if (flag & kOutsidePattern) return setError(FEEL_HIP_EPATTERN, "entry outside the sparsity graph");
    // the same ordered ADD as feel_hip_mat_add_matrix (entries in row-major order)
    hipLaunchKernelGGL(kAddAt, dim3(1), dim3(1), 0, c->stream, m, dpos, dval, A->v64, A->v32);
    FEEL_HIP_CHECK(hipGetLastError());
    FEEL_HIP_CHECK(hipStreamSynchronize(c->stream));
    A->dinvValid = false;
    A->sell64Valid = A->sell32Valid = false;
    elemOpInvalidate(A);
    return FEEL_HIP_OK;
}
