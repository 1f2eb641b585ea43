// HIP kernels of the device layer (gfx950, 64-wide wavefronts).
//
// Reductions never use floating-point atomics or grid-wide fences: a kernel
// writes one partial per block (fixed grid of kReduceBlocks blocks), and every
// block of the NEXT kernel re-reduces those partials in the same fixed order,
// so all blocks see bit-identical scalars and results are run-to-run
// deterministic.  Scalars that must persist are written by block 0 into
// double-buffered slots indexed by the iteration number the host passes in.
#pragma once
#include "common.hpp"

namespace feelpp_amd {

// ---------------------------------------------------------------- reductions
__device__ inline double waveReduceSum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// sum over the block (blockDim.x == kBlock); result valid in every thread
__device__ inline double blockReduceSum(double v) {
    __shared__ double sh[kBlock / 64];
    __shared__ double total;
    v = waveReduceSum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
#pragma unroll
        for (int i = 0; i < kBlock / 64; ++i) s += sh[i];
        total = s;
    }
    __syncthreads();
    const double r = total;
    __syncthreads();
    return r;
}

// Reads the solver's done flag once per block and broadcasts it, so that every
// thread of a block takes the same early exit (the flag may be set by block 0
// of the running kernel while other blocks start).
__device__ inline bool blockDone(const CgState* st) {
    __shared__ int flag;
    if (threadIdx.x == 0) flag = st->done;
    __syncthreads();
    return flag != 0;
}

// N block sums at once (one pair of barriers instead of N); each is combined
// in blockReduceSum's order, so the results are the same bits; valid in every
// thread
template <int BS, int N>
__device__ inline void blockReduceSums(double (&v)[N]) {
    __shared__ double sh[N][BS / 64];
    __shared__ double total[N];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const double t = waveReduceSum(v[k]);
        if (lane == 0) sh[k][w] = t;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        double t = 0;
#pragma unroll
        for (int i = 0; i < BS / 64; ++i) t += sh[threadIdx.x][i];
        total[threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = total[k];
    __syncthreads();
}

// This thread's share of the sum of count partials (BS threads): four
// independent accumulators so that its loads are in flight together, then a
// fixed combination order.
template <int BS = kBlock>
__device__ inline double partialShare(const double* __restrict__ part, int count) {
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    int i = threadIdx.x;
    for (; i + 3 * BS < count; i += 4 * BS) {
        v0 += part[i];
        v1 += part[i + BS];
        v2 += part[i + 2 * BS];
        v3 += part[i + 3 * BS];
    }
    for (; i < count; i += BS) v0 += part[i];
    return (v0 + v1) + (v2 + v3);
}

// Sum of count partials, identical in every block.
__device__ inline double reducePartials(const double* __restrict__ part, int count = kReduceBlocks) {
    return blockReduceSum(partialShare(part, count));
}

// 16-byte vector of T (four floats or two doubles) for the streaming CG
// vector kernels: one dwordx4 load or store per lane instead of four or two.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};

template <typename... P>
__device__ inline bool aligned16(const P*... ptr) {
    return ((reinterpret_cast<uintptr_t>(ptr) | ...) & 15) == 0;
}

// ---------------------------------------------------------------- geometry
// Affine map of a P1 simplex on the [-1,1] reference element
// (feelpoly/geomap.hpp:2258-2355): K[:,a] = (x_{a+1} - x_0)/2, J = |det K|
// (:1972-1976).  Returns Kinv (K^{-1}) and J.
template <int DIM>
__device__ inline double affineMap(const double* __restrict__ coords, const int32_t* __restrict__ el,
                                   double Kinv[DIM][DIM]) {
    double X[DIM + 1][DIM];
#pragma unroll
    for (int v = 0; v <= DIM; ++v)
#pragma unroll
        for (int c = 0; c < DIM; ++c) X[v][c] = coords[(size_t)el[v] * DIM + c];
    double K[DIM][DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int a = 0; a < DIM; ++a) K[r][a] = 0.5 * (X[a + 1][r] - X[0][r]);
    if constexpr (DIM == 2) {
        const double det = K[0][0] * K[1][1] - K[0][1] * K[1][0];
        const double id = 1.0 / det;
        Kinv[0][0] = K[1][1] * id; Kinv[0][1] = -K[0][1] * id;
        Kinv[1][0] = -K[1][0] * id; Kinv[1][1] = K[0][0] * id;
        return fabs(det);
    } else {
        const double c00 = K[1][1] * K[2][2] - K[1][2] * K[2][1];
        const double c01 = K[1][2] * K[2][0] - K[1][0] * K[2][2];
        const double c02 = K[1][0] * K[2][1] - K[1][1] * K[2][0];
        const double det = K[0][0] * c00 + K[0][1] * c01 + K[0][2] * c02;
        const double id = 1.0 / det;
        Kinv[0][0] = c00 * id;
        Kinv[1][0] = c01 * id;
        Kinv[2][0] = c02 * id;
        Kinv[0][1] = (K[0][2] * K[2][1] - K[0][1] * K[2][2]) * id;
        Kinv[1][1] = (K[0][0] * K[2][2] - K[0][2] * K[2][0]) * id;
        Kinv[2][1] = (K[0][1] * K[2][0] - K[0][0] * K[2][1]) * id;
        Kinv[0][2] = (K[0][1] * K[1][2] - K[0][2] * K[1][1]) * id;
        Kinv[1][2] = (K[0][2] * K[1][0] - K[0][0] * K[1][2]) * id;
        Kinv[2][2] = (K[0][0] * K[1][1] - K[0][1] * K[1][0]) * id;
        return fabs(det);
    }
}

// Per-element factors of the scalar forms.  With grad phi = B grad_ref phi,
// B = K^{-T} (polynomialset_impl.hpp:274-302), the element matrix
//   A_ij = sum_q w_q J (kd grad phi_i . grad phi_j + km phi_i phi_j)
//        = sum_{a<=b} F_ab Rsym_ab(i,j) + F_m M(i,j)
// with C = B^T B = K^{-1} K^{-T}, F_ab = kd J C_ab, F_m = km J and the
// reference tensors Rsym, M computed once per (order, quadrature).
// Layout: 3D [C00 C11 C22 C01 C02 C12 | m | pad] (stride 8), 2D [C00 C11 C01 | m] (stride 4).
template <int DIM>
__global__ __launch_bounds__(kBlock) void kGeoScalar(int64_t ne, const double* __restrict__ coords,
                                                     const int32_t* __restrict__ elems,
                                                     const double* __restrict__ kd, double kds,
                                                     const double* __restrict__ km, double kms,
                                                     double* __restrict__ F) {
    constexpr int STRIDE = DIM == 3 ? 8 : 4;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        double Ki[DIM][DIM];
        const double J = affineMap<DIM>(coords, elems + e * (DIM + 1), Ki);
        const double a = kds * (kd ? kd[e] : 1.0) * J;
        const double m = kms * (km ? km[e] : 1.0) * J;
        double C[DIM][DIM];
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) {
                double s = 0;
#pragma unroll
                for (int r = 0; r < DIM; ++r) s += Ki[i][r] * Ki[j][r];
                C[i][j] = s;
            }
        double* f = F + e * STRIDE;
        if constexpr (DIM == 3) {
            f[0] = a * C[0][0]; f[1] = a * C[1][1]; f[2] = a * C[2][2];
            f[3] = a * C[0][1]; f[4] = a * C[0][2]; f[5] = a * C[1][2];
            f[6] = m; f[7] = 0.0;
        } else {
            f[0] = a * C[0][0]; f[1] = a * C[1][1]; f[2] = a * C[0][1]; f[3] = m;
        }
    }
}

// ---------------------------------------------------------------- scatter
// Wavefront-level segmented reduction of element contributions into CSR
// slots.  The gather map lists, slot by slot, the element-matrix entries that
// land in each CSR slot in ascending element order (the order MatSetValues
// ADD applies them); the top bit of an entry marks the last contribution of
// its slot.  A task is a run of whole slots (about kSegTask contributions);
// a wavefront walks its tasks 64 contributions at a time, one lane per
// contribution, so the entry ids are read coalesced and the element-factor
// gathers of a chunk are independent.  A shuffle-based inclusive segmented
// scan (fixed tree, so results are run-to-run identical) sums each slot; the
// lane holding a slot's last contribution writes it.  Slot index = task's
// first segment + ends before the lane; segSlot maps segments to CSR slots
// when the pattern has slots without contributions (else identity).
// DPP move of v from another lane of the wavefront (dpp_ctrl CTRL, rows
// enabled by ROWS); lanes without a source and lanes of disabled rows get 0.
// row_shr:n (0x110 + n) reads lane - n of the same 16-lane row;
// row_bcast:15 (0x142) gives row r+1 lane 15 of row r; row_bcast:31 (0x143)
// gives rows 2 and 3 lane 31.
template <int CTRL, int ROWS>
__device__ inline int dppMove(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ inline double dppMove(double v) {
    const int lo = dppMove<CTRL, ROWS>(__double2loint(v));
    const int hi = dppMove<CTRL, ROWS>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// One step of a segmented inclusive scan: x takes the value moved in unless a
// segment starts at or after its source; head records a start.
template <int CTRL, int ROWS>
__device__ inline void segScanStep(double& x, int& head) {
    const double y = dppMove<CTRL, ROWS>(x);
    const int hy = dppMove<CTRL, ROWS>(head);
    if (!head) x = y + x;
    head |= hy;
}

template <class Value>
__device__ inline void segScatter(int64_t ntasks, const uint32_t* __restrict__ taskSeg,
                                  const uint32_t* __restrict__ taskK, const uint32_t* __restrict__ centry,
                                  const int32_t* __restrict__ segSlot, double* __restrict__ v64,
                                  float* __restrict__ v32, int add, const Value& value) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int64_t t = wave; t < ntasks; t += nwaves) {
        uint32_t seg = taskSeg[t];
        const uint32_t kend = taskK[t + 1];
        double carry = 0.0;
        bool open = false;
        for (uint32_t k = taskK[t]; k < kend; k += 64) {
            const uint32_t kk = k + lane;
            const bool valid = kk < kend;
            const uint32_t c = valid ? centry[kk] : 0u;
            const bool isEnd = valid && (c >> 31);
            double x = valid ? value(c & 0x7fffffffu) : 0.0;
            const uint64_t ends = __ballot(isEnd);
            int head = lane == 0 ? !open : (int)((ends >> (lane - 1)) & 1ull);
            if (lane == 0 && open) x = carry + x;
            // segmented inclusive scan in registers: Hillis-Steele inside each
            // 16-lane row, then the row carries (lane 15 into row 1 and row 3,
            // lane 31 into rows 2 and 3); a fixed order, so run-to-run identical
            segScanStep<0x111, 0xf>(x, head);
            segScanStep<0x112, 0xf>(x, head);
            segScanStep<0x114, 0xf>(x, head);
            segScanStep<0x118, 0xf>(x, head);
            segScanStep<0x142, 0xa>(x, head);
            segScanStep<0x143, 0xc>(x, head);
            if (isEnd) {
                const uint32_t sidx = seg + (uint32_t)__popcll(ends & below);
                const int64_t p = segSlot ? (int64_t)segSlot[sidx] : (int64_t)sidx;
                if (v64) {
                    const double v = add ? v64[p] + x : x;
                    v64[p] = v;
                    if (v32) v32[p] = (float)v;
                } else if (v32) {
                    v32[p] = add ? (float)((double)v32[p] + x) : (float)x;
                }
            }
            seg += (uint32_t)__popcll(ends);
            const int last = (int)((kend - 1 - k) < 63u ? (kend - 1 - k) : 63u);
            carry = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), last),
                                     __builtin_amdgcn_readlane(__double2loint(x), last));
            open = ((ends >> last) & 1ull) == 0;
        }
    }
}

// Scalar forms: entry (i,j) of element e = sum_t F[e][t] R[ij][t], with the
// element factors F (ST doubles per element, zero padded) and the reference
// tensors entry-major (zero padded), both read with 16-byte loads and summed
// in the order t = 0 .. ST-1.  Tables larger than the LDS budget (P4: 1225 x 8
// doubles) are read through L1/L2.
template <int NN, int ST, bool USE_LDS = true>
__global__ __launch_bounds__(kBlock) void kGatherScalar(int64_t ntasks, const uint32_t* __restrict__ taskSeg,
                                                        const uint32_t* __restrict__ taskK,
                                                        const uint32_t* __restrict__ centry,
                                                        const int32_t* __restrict__ segSlot,
                                                        const double* __restrict__ F,
                                                        const double* __restrict__ rtab,
                                                        double* __restrict__ v64, float* __restrict__ v32,
                                                        int add) {
    constexpr int H = ST / 2;
    __shared__ double2 Rs[USE_LDS ? NN * H : 1];
    if constexpr (USE_LDS) {
        for (int i = threadIdx.x; i < NN * H; i += kBlock) Rs[i] = reinterpret_cast<const double2*>(rtab)[i];
        __syncthreads();
    }
    const double2* R = USE_LDS ? Rs : reinterpret_cast<const double2*>(rtab);
    const double2* F2 = reinterpret_cast<const double2*>(F);
    auto value = [&](uint32_t c) {
        const uint32_t el = c / NN;
        const uint32_t ij = c - el * NN;
        const double2* f = F2 + (size_t)el * H;
        const double2* r = R + (size_t)ij * H;
        double a = 0.0;
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const double2 fv = f[h], rv = r[h];
            a += fv.x * rv.x;
            a += fv.y * rv.y;
        }
        return a;
    };
    segScatter(ntasks, taskSeg, taskK, centry, segSlot, v64, v32, add, value);
}

// Per-element factors of isotropic linear elasticity inner(sigma(u), grad v),
// sigma = lambda tr(eps) I + 2 mu eps (quickstart/qs_elasticity.cpp:41-63).
// With B = K^{-T}, C = B^T B and R_ab(i,j) = sum_q w_q d_a phi_i d_b phi_j on the
// reference element, the (c,i),(d,j) entry is sum_ab Ke[c][d][a][b] R_ab(i,j) with
//   Ke[c][d][a][b] = J (lambda B_ca B_db + mu B_da B_cb + mu delta_cd C_ab).
// Layout: kElasticF = 90 doubles per element, block (c,d) at (c*3+d)*10 holding
// index a*3+b and a zero pad (16-byte aligned blocks).
static __global__ __launch_bounds__(kBlock) void kGeoElastic(int64_t ne, const double* __restrict__ coords,
                                                             const int32_t* __restrict__ elems,
                                                             const double* __restrict__ young,
                                                             const double* __restrict__ poisson,
                                                             double* __restrict__ F) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        double Ki[3][3];
        const double J = affineMap<3>(coords, elems + e * 4, Ki);
        const double E = young[e], nu = poisson[e];
        const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
        const double mu = E / (2.0 * (1.0 + nu));
        double B[3][3], C[3][3];
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int a = 0; a < 3; ++a) B[x][a] = Ki[a][x];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) C[a][b] = B[0][a] * B[0][b] + B[1][a] * B[1][b] + B[2][a] * B[2][b];
        double* f = F + e * kElasticF;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b)
                        f[(c * 3 + d) * 10 + a * 3 + b] =
                            J * (lam * B[c][a] * B[d][b] + mu * B[d][a] * B[c][b] + (c == d ? mu * C[a][b] : 0.0));
                f[(c * 3 + d) * 10 + 9] = 0.0;
            }
    }
}

// Vector elasticity form: local entry ij = (c*n + i) * 3n + (d*n + j) of
// element e = sum_ab F_e[c,d,a,b] R_ab(i,j); reference tensors entry-major,
// R[(i*n + j)*10 + a*3 + b] with a zero pad (16-byte loads, summed in the
// order t = a*3 + b), staged in LDS when they fit.
template <int N, bool USE_LDS = true>
__global__ __launch_bounds__(kBlock) void kGatherElastic(int64_t ntasks, const uint32_t* __restrict__ taskSeg,
                                                         const uint32_t* __restrict__ taskK,
                                                         const uint32_t* __restrict__ centry,
                                                         const int32_t* __restrict__ segSlot,
                                                         const double* __restrict__ F,
                                                         const double* __restrict__ rtab,
                                                         double* __restrict__ v64, float* __restrict__ v32, int add) {
    constexpr int NN = N * N, N3 = 3 * N, NN3 = N3 * N3;
    __shared__ double2 Rs[USE_LDS ? 5 * NN : 1];
    if constexpr (USE_LDS) {
        for (int i = threadIdx.x; i < 5 * NN; i += kBlock) Rs[i] = reinterpret_cast<const double2*>(rtab)[i];
        __syncthreads();
    }
    const double2* R = USE_LDS ? Rs : reinterpret_cast<const double2*>(rtab);
    const double2* F2 = reinterpret_cast<const double2*>(F);
    auto value = [&](uint32_t c) {
        const uint32_t el = c / NN3;
        const uint32_t loc = c - el * NN3;
        const uint32_t rl = loc / N3, cl = loc - rl * N3;
        const uint32_t cc = rl / N, i = rl - cc * N, dd = cl / N, j = cl - dd * N;
        const double2* f = F2 + (size_t)el * (kElasticF / 2) + (cc * 3 + dd) * 5;
        const double2* r = R + (size_t)(i * N + j) * 5;
        double a = 0.0;
#pragma unroll
        for (int h = 0; h < 5; ++h) {
            const double2 fv = f[h], rv = r[h];
            a += fv.x * rv.x;
            a += fv.y * rv.y;
        }
        return a;
    };
    segScatter(ntasks, taskSeg, taskK, centry, segSlot, v64, v32, add, value);
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Element matrices materialised on the FP64 matrix cores.  For all elements
// at once the reference-tensor contraction is one GEMM: RPE rows per element
// (elasticity: m = 9e + 3c + d; scalar forms: m = e) times the reference
// tensors, columns i*nloc + j, over the factor index t < KV (padded to a
// multiple of 4 with zeros):
//   elasticity  Ke[(c,i),(d,j)] = sum_{t<9} F[e][(c,d)][t] R[(i,j)][t]
//   scalar      A_e(i,j)        = sum_{t<7} F[e][t] R[(i,j)][t]
// Factor of row m, index t: F[(m / RPE) * FS + (m % RPE) * FR + t]; table
// entry: R[(i*nloc + j) * RS + t].  A wave owns a 16-row tile and sweeps every
// 16-column tile with v_mfma_f64_16x16x4f64.  Operands: lane l holds
// A[l%16][l/16] and B[l/16][l%16]; the 16x16 result holds row l/16 + 4r,
// column l%16 in register r.  Em is tile-major (tile (rt, ct) at
// (rt*NCT + ct)*256, row-major inside), so each result register is stored as
// 512 contiguous bytes; the entry maps below give a gather-map entry's place.
template <int NN, int RPE>
struct MatTiles {
    static constexpr int NCT = (NN + 15) / 16;
    static int64_t doubles(int64_t ne) { return (ne * RPE + 15) / 16 * NCT * 256; }
    __device__ static int64_t at(int64_t m, uint32_t col) {
        return ((m >> 4) * NCT + (col >> 4)) * 256 + (m & 15) * 16 + (col & 15);
    }
};

// entry id e*(3n)^2 + (c*n+i)*3n + d*n + j of an elasticity element matrix
template <int N>
struct ElasticEntry {
    using Tiles = MatTiles<N * N, 9>;
    __device__ static int64_t at(uint32_t c) {
        constexpr uint32_t N3 = 3 * N, NN3 = N3 * N3;
        const uint32_t el = c / NN3;
        const uint32_t loc = c - el * NN3;
        const uint32_t rl = loc / N3, cl = loc - rl * N3;
        const uint32_t cc = rl / N, i = rl - cc * N, dd = cl / N, j = cl - dd * N;
        return Tiles::at((int64_t)el * 9 + cc * 3 + dd, i * N + j);
    }
};

// entry id e*n^2 + i*n + j of a scalar element matrix
template <int N>
struct ScalarEntry {
    using Tiles = MatTiles<N * N, 1>;
    __device__ static int64_t at(uint32_t c) {
        const uint32_t el = c / (uint32_t)(N * N);
        return Tiles::at(el, c - el * (uint32_t)(N * N));
    }
};

template <int NN, int RPE, int FS, int FR, int RS, int KV>
__global__ __launch_bounds__(kBlock) void kElemMatMfma(int64_t ne, const double* __restrict__ F,
                                                       const double* __restrict__ R, double* __restrict__ Em) {
    using T = MatTiles<NN, RPE>;
    constexpr int KS = (KV + 3) / 4;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int64_t rows = ne * RPE, rtiles = (rows + 15) / 16;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t rt = wave; rt < rtiles; rt += nwaves) {
        const int64_t m = rt * 16 + li;
        double a[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = 4 * s + lk;
            a[s] = (m < rows && k < KV) ? F[(m / RPE) * FS + (m % RPE) * FR + k] : 0.0;
        }
        double* tile = Em + rt * T::NCT * 256 + lk * 16 + li;
        for (int ct = 0; ct < T::NCT; ++ct, tile += 256) {
            const int ij = ct * 16 + li;
            f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + lk;
                const double b = (ij < NN && k < KV) ? R[ij * RS + k] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b, acc, 0, 0, 0);
            }
            // rows past RPE ne and columns past nloc^2 are zero padding, never read
#pragma unroll
            for (int r = 0; r < 4; ++r) __builtin_nontemporal_store(acc[r], tile + 64 * r);
        }
    }
}

// Scatter of materialised element matrices: entry c of the map is read from
// its tile-major place in Em.
template <class Entry>
__global__ __launch_bounds__(kBlock) void kGatherMaterialised(int64_t ntasks, const uint32_t* __restrict__ taskSeg,
                                                              const uint32_t* __restrict__ taskK,
                                                              const uint32_t* __restrict__ centry,
                                                              const int32_t* __restrict__ segSlot,
                                                              const double* __restrict__ Em,
                                                              double* __restrict__ v64, float* __restrict__ v32,
                                                              int add) {
    auto value = [&](uint32_t c) { return Em[Entry::at(c)]; };
    segScatter(ntasks, taskSeg, taskK, centry, segSlot, v64, v32, add, value);
}

// Element vectors of f*id(v) (linearformcontext.hpp:294-305):
// b_e(c*nl+i) = sum_q w_q J f_c(x_q) phi_i(q), x_q = sum_v X_v lambda_v(q).
// qtab layout: w[Q] | phi[nl*Q] | lambda[(DIM+1)*Q].
template <int DIM>
__global__ __launch_bounds__(kBlock) void kElemRhs(int64_t ne, int nl, int ncomp, int Q,
                                                   const double* __restrict__ coords,
                                                   const int32_t* __restrict__ elems,
                                                   const double* __restrict__ qtab, int kind, double p0,
                                                   double p1, double p2, double* __restrict__ be) {
    const double* w = qtab;
    const double* phi = qtab + Q;
    const double* lam = phi + (size_t)nl * Q;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t* el = elems + e * (DIM + 1);
        double Ki[DIM][DIM];
        const double J = affineMap<DIM>(coords, el, Ki);
        double* out = be + e * (size_t)nl * ncomp;
        double X[DIM + 1][DIM];
        if (kind == FEEL_HIP_SRC_SINSIN)
#pragma unroll
            for (int v = 0; v <= DIM; ++v)
#pragma unroll
                for (int c = 0; c < DIM; ++c) X[v][c] = coords[(size_t)el[v] * DIM + c];
        // each entry is accumulated over q in a register and written once
        for (int c = 0; c < ncomp; ++c) {
            const double fconst = kind == FEEL_HIP_SRC_VECTOR ? (c == 0 ? p0 : c == 1 ? p1 : p2) : p0;
            for (int i = 0; i < nl; ++i) {
                double acc = 0.0;
                for (int q = 0; q < Q; ++q) {
                    double fc = fconst;
                    if (kind == FEEL_HIP_SRC_SINSIN) {
                        double x[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; ++d) x[d] = 0.0;
#pragma unroll
                        for (int v = 0; v <= DIM; ++v)
#pragma unroll
                            for (int d = 0; d < DIM; ++d) x[d] += X[v][d] * lam[v * Q + q];
                        fc = p0 * sin(M_PI * x[0]) * sin(M_PI * x[1]);
                    }
                    const double wJ = w[q] * J;
                    acc += wJ * fc * phi[i * Q + q];
                }
                out[c * nl + i] = acc;
            }
        }
    }
}

// ---------------------------------------------------------------- norms
// Reference functions of the error norms, one per component.
struct NormFn {
    int kind[3];
    double c[3][4];
};

template <int DIM>
__device__ inline double normFnValue(const NormFn& f, int comp, const double* x, double* grad) {
    const double* c = f.c[comp];
    if (f.kind[comp] == FEEL_HIP_FN_LINEAR) {
        double v = c[0];
#pragma unroll
        for (int d = 0; d < DIM; ++d) { v += c[d + 1] * x[d]; grad[d] = c[d + 1]; }
        return v;
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) grad[d] = 0.0;
    if (f.kind[comp] == FEEL_HIP_FN_SINSIN) {
        const double sx = sin(M_PI * x[0]), sy = sin(M_PI * x[1]);
        grad[0] = c[0] * M_PI * cos(M_PI * x[0]) * sy;
        grad[1] = c[0] * M_PI * sx * cos(M_PI * x[1]);
        return c[0] * sx * sy;
    }
    return 0.0;
}

// Block partials of integrate(elements, inner(u - f)) (semi = 0) or of
// integrate(elements, inner(grad(u) - grad(f))) (semi = 1) on affine simplices
// (feelvf/norml2.hpp:37-57, normsemih1.hpp): u_h(x_q) = sum_l phi_l(q) u_l,
// grad u_h = sum_l u_l B grad phi_l(q) with B = K^{-T}, x_q = sum_v lambda_v(q) X_v.
// tab: w[Q] | phi[nloc*Q] | dphi[(l*Q+q)*DIM+a] | lambda[(DIM+1)*Q].
template <int DIM>
__global__ __launch_bounds__(kBlock) void kElemNorm(int64_t ne, int nloc, int ncomp, int Q,
                                                    const double* __restrict__ coords,
                                                    const int32_t* __restrict__ elems,
                                                    const int32_t* __restrict__ e2d, const double* __restrict__ u,
                                                    const double* __restrict__ tab, NormFn fn, int semi,
                                                    double* __restrict__ part) {
    const double* w = tab;
    const double* phi = tab + Q;
    const double* dphi = phi + (size_t)nloc * Q;
    const double* lam = dphi + (size_t)nloc * Q * DIM;
    const int nldof = nloc * ncomp;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t* el = elems + e * (DIM + 1);
        double Ki[DIM][DIM];
        const double J = affineMap<DIM>(coords, el, Ki);
        double X[DIM + 1][DIM];
#pragma unroll
        for (int v = 0; v <= DIM; ++v)
#pragma unroll
            for (int d = 0; d < DIM; ++d) X[v][d] = coords[(size_t)el[v] * DIM + d];
        for (int q = 0; q < Q; ++q) {
            double x[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                x[d] = 0.0;
#pragma unroll
                for (int v = 0; v <= DIM; ++v) x[d] += X[v][d] * lam[v * Q + q];
            }
            const double wJ = w[q] * J;
            for (int c = 0; c < ncomp; ++c) {
                const int32_t* l2g = e2d + e * nldof + c * nloc;
                double gf[DIM];
                const double fv = normFnValue<DIM>(fn, c, x, gf);
                if (!semi) {
                    double uh = 0.0;
                    for (int l = 0; l < nloc; ++l) uh += phi[(size_t)l * Q + q] * u[l2g[l]];
                    const double d = uh - fv;
                    acc += wJ * d * d;
                } else {
                    double g[DIM];
#pragma unroll
                    for (int r = 0; r < DIM; ++r) g[r] = 0.0;
                    for (int l = 0; l < nloc; ++l) {
                        const double ul = u[l2g[l]];
                        const double* dp = dphi + ((size_t)l * Q + q) * DIM;
#pragma unroll
                        for (int r = 0; r < DIM; ++r) {
                            double s = 0.0;
#pragma unroll
                            for (int a = 0; a < DIM; ++a) s += Ki[a][r] * dp[a];
                            g[r] += ul * s;
                        }
                    }
                    double n2 = 0.0;
#pragma unroll
                    for (int r = 0; r < DIM; ++r) n2 += (g[r] - gf[r]) * (g[r] - gf[r]);
                    acc += wJ * n2;
                }
            }
        }
    }
    acc = blockReduceSum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Linear-form scatter: b_r = sum of the element-vector entries landing in row
// r (rentry, ascending element order), L lanes per row so that the row's entry
// list is read coalesced; lane partials are combined in a fixed order.
template <int L>
__global__ __launch_bounds__(kBlock) void kGatherVec(int64_t n, const uint32_t* __restrict__ rptr,
                                                     const uint32_t* __restrict__ rentry,
                                                     const double* __restrict__ be, double* __restrict__ b, int add) {
    const int lane = threadIdx.x & (L - 1);
    constexpr int RPB = kBlock / L;
    for (int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / L; r < n; r += (int64_t)gridDim.x * RPB) {
        double s = 0.0;
        for (uint32_t k = rptr[r] + lane; k < rptr[r + 1]; k += L) s += be[rentry[k]];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        if (lane == 0) b[r] = add ? b[r] + s : s;
    }
}

// ---------------------------------------------------------------- Dirichlet
// MatZeroRows / MatZeroRowsColumns semantics (matrixpetsc.cpp:1764-1850).
// Both precisions of on() in one pass, L lanes per row so that a row's
// column indices and values are read coalesced: flagged rows become
// diag * e_r (rhs_r = diag g_r); with symmetric elimination the flagged
// columns of the other rows are zeroed and rhs_r -= sum A_rc g_c, summed from
// the FP64 values when present (else the FP32 ones) over the row's lanes.
// Rows are list[0 .. count) (every row when list is null): rows outside the
// list are rows the call leaves unchanged.
template <int L>
__global__ __launch_bounds__(kBlock) void kDirichletRows(int64_t count, const int32_t* __restrict__ list,
                                                         const int32_t* __restrict__ ia,
                                                         const int32_t* __restrict__ ja, double* __restrict__ v64,
                                                         float* __restrict__ v32, double* __restrict__ rhs,
                                                         const uint8_t* __restrict__ flag,
                                                         const double* __restrict__ g, int symmetric, double diag) {
    const int lane = threadIdx.x & (L - 1);
    constexpr int RPB = kBlock / L;
    for (int64_t i = (int64_t)blockIdx.x * RPB + threadIdx.x / L; i < count; i += (int64_t)gridDim.x * RPB) {
        const int64_t r = list ? (int64_t)list[i] : i;
        const int b = ia[r], e = ia[r + 1];
        if (flag[r]) {
            for (int p = b + lane; p < e; p += L) {
                const bool d = ja[p] == r;
                if (v64) v64[p] = d ? diag : 0.0;
                if (v32) v32[p] = d ? (float)diag : 0.0f;
            }
            if (lane == 0 && rhs) rhs[r] = diag * (g ? g[r] : 0.0);
        } else if (symmetric) {
            double corr = 0.0;
            for (int p = b + lane; p < e; p += L) {
                const int c = ja[p];
                if (flag[c]) {
                    if (g) corr += (v64 ? v64[p] : (double)v32[p]) * g[c];
                    if (v64) v64[p] = 0.0;
                    if (v32) v32[p] = 0.0f;
                }
            }
            if (rhs && g) {
#pragma unroll
                for (int off = L / 2; off > 0; off >>= 1) corr += __shfl_xor(corr, off, L);
                if (lane == 0) rhs[r] -= corr;
            }
        }
    }
}

// mark[r] = 1 for the rows symmetric elimination changes: flagged rows and
// rows holding a flagged column (L lanes per row)
template <int L>
__global__ __launch_bounds__(kBlock) void kMarkTouched(int64_t n, const int32_t* __restrict__ ia,
                                                       const int32_t* __restrict__ ja, const uint8_t* __restrict__ flag,
                                                       uint8_t* __restrict__ mark) {
    const int lane = threadIdx.x & (L - 1);
    constexpr int RPB = kBlock / L;
    for (int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / L; r < n; r += (int64_t)gridDim.x * RPB) {
        int any = 0;
        for (int p = ia[r] + lane; p < ia[r + 1]; p += L) any |= flag[ja[p]];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) any |= __shfl_xor(any, off, L);
        if (lane == 0) mark[r] = (uint8_t)(any | flag[r]);
    }
}

// ---------------------------------------------------------------- CG kernels
// y = A x with L lanes per row (grid-stride over row groups).  When part is
// given, also writes the block partial of sum_r x_r y_r (= p.Ap).
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void kSpmv(int64_t n, const int32_t* __restrict__ ia,
                                                const int32_t* __restrict__ ja, const T* __restrict__ val,
                                                const T* __restrict__ x, T* __restrict__ y,
                                                double* __restrict__ part, const CgState* __restrict__ st) {
    if (st && blockDone(st)) return;
    const int lane = threadIdx.x & (L - 1);
    constexpr int RPB = kBlock / L;
    double dot = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / L; r < n; r += (int64_t)gridDim.x * RPB) {
        const int b = ia[r], e = ia[r + 1];
        T s = 0;
        for (int k = b + lane; k < e; k += L) s += val[k] * x[ja[k]];
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
        if (lane == 0) {
            y[r] = s;
            if (part) dot += (double)x[r] * (double)s;
        }
    }
    if (part) {
        const double t = blockReduceSum(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// p = z + (beta/beta_prev) q, preceded by the convergence test on the
// preconditioned residual norm (KSPConvergedDefault with UIRNorm); q is the
// previous direction (p itself, or the other buffer of the pipelined FP32
// iteration).  With hist, block 0 records beta = (r, z) of iteration it at
// hist[2 it] (the update kernel records p.Ap at hist[2 it + 1]): the Lanczos
// coefficients of the run, from which the host estimates the condition number.
template <typename T>
__global__ __launch_bounds__(kBlock) void kCgDirection(int64_t n, int it, const T* __restrict__ z, const T* q, T* p,
                                                       const double* __restrict__ zrPart,
                                                       const double* __restrict__ zzPart, CgState* st,
                                                       double* betaSlots, double* hist) {
    if (blockDone(st)) return;
    double sums[2] = {partialShare(zrPart, kReduceBlocks), partialShare(zzPart, kReduceBlocks)};
    blockReduceSums<kBlock, 2>(sums);
    const double beta = sums[0];
    const double dp = sqrt(sums[1]);
    const double ttol = it == 0 ? fmax(st->rtol * dp, st->atol) : st->ttol;
    const double rnorm0 = it == 0 ? dp : st->rnorm0;
    int reason = 0;
    if (dp != dp) reason = -9;
    else if (dp <= ttol) reason = dp < st->atol ? 3 : 2;
    else if (it > 0 && dp >= st->dtol * rnorm0) reason = -4;
    else if (it >= st->maxit) reason = -3;
    else if (beta == 0.0) reason = 3;
    else if (it > 0 && beta * betaSlots[(it - 1) & 1] < 0.0) reason = -8;   // indefinite preconditioner
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->dp = dp;
        if (it == 0) { st->ttol = ttol; st->rnorm0 = rnorm0; }
        if (reason) { st->done = 1; st->reason = reason; st->its = reason == -8 ? it + 1 : it; }
        betaSlots[it & 1] = beta;
        if (hist && it < kHistMax) hist[2 * it] = beta;
    }
    if (reason) return;
    // the direction (re)starts as p = z at it == st->fresh (0 unless a
    // mixed-precision solve restarted it)
    const bool first = it == st->fresh;
    const double bb = first ? 0.0 : beta / betaSlots[(it - 1) & 1];
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    int64_t tail = 0;
    if (aligned16(z, q, p)) {
        typedef typename Vec16<T>::type VT;
        constexpr int V = 16 / sizeof(T);
        const int64_t nv = n / V;
        const VT* zv = reinterpret_cast<const VT*>(z);
        const VT* qv = reinterpret_cast<const VT*>(q);
        VT* pv = reinterpret_cast<VT*>(p);
        for (int64_t i = gid; i < nv; i += stride) {
            const VT zi = zv[i];
            if (first) {
                pv[i] = zi;
            } else {
                VT pi = qv[i];
#pragma unroll
                for (int l = 0; l < V; ++l) pi[l] = (T)((double)zi[l] + bb * (double)pi[l]);
                pv[i] = pi;
            }
        }
        tail = nv * V;
    }
    for (int64_t i = tail + gid; i < n; i += stride) p[i] = first ? z[i] : (T)((double)z[i] + bb * (double)q[i]);
}

// x += a p, r -= a w, z = D^{-1} r; partials of z.r and z.z.  dpiPart holds
// the npart block partials of p.w left by the SpMV.
template <typename T>
__global__ __launch_bounds__(kBlock) void kCgUpdate(int64_t n, int it, T* __restrict__ x, T* __restrict__ r,
                                                    T* __restrict__ z, const T* __restrict__ p,
                                                    const T* __restrict__ w, const T* __restrict__ dinv,
                                                    const double* __restrict__ dpiPart, int npart,
                                                    double* __restrict__ zrPart, double* __restrict__ zzPart,
                                                    CgState* st, const double* betaSlots, double* dpiSlots,
                                                    double* hist) {
    if (blockDone(st)) return;
    const double dpi = reducePartials(dpiPart, npart);
    const bool indefinite = dpi == 0.0 || (it > 0 && dpi * dpiSlots[(it - 1) & 1] <= 0.0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        dpiSlots[it & 1] = dpi;
        if (hist && it < kHistMax) hist[2 * it + 1] = dpi;
        if (indefinite) { st->done = 1; st->reason = -10; st->its = it + 1; }
    }
    if (indefinite) return;
    const double a = betaSlots[it & 1] / dpi;
    double zr = 0.0, zz = 0.0;
    // one entry: x += a p, r -= a w, z = D^{-1} r, accumulate z.r and z.z
    auto step = [&](T& xi, T& ri, T& zi, T pi, T wi, T di) {
        xi = (T)((double)xi + a * (double)pi);
        ri = (T)((double)ri - a * (double)wi);
        zi = di * ri;
        zr += (double)zi * (double)ri;
        zz += (double)zi * (double)zi;
    };
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    int64_t tail = 0;
    if (aligned16(x, r, z, p, w, dinv)) {
        typedef typename Vec16<T>::type VT;
        constexpr int V = 16 / sizeof(T);
        const int64_t nv = n / V;
        for (int64_t i = gid; i < nv; i += stride) {
            VT xv = reinterpret_cast<const VT*>(x)[i], rv = reinterpret_cast<const VT*>(r)[i], zv;
            const VT pv = reinterpret_cast<const VT*>(p)[i], wv = reinterpret_cast<const VT*>(w)[i],
                     dv = reinterpret_cast<const VT*>(dinv)[i];
#pragma unroll
            for (int l = 0; l < V; ++l) {
                T xi = xv[l], ri = rv[l], zi;
                step(xi, ri, zi, pv[l], wv[l], dv[l]);
                xv[l] = xi;
                rv[l] = ri;
                zv[l] = zi;
            }
            reinterpret_cast<VT*>(x)[i] = xv;
            reinterpret_cast<VT*>(r)[i] = rv;
            reinterpret_cast<VT*>(z)[i] = zv;
        }
        tail = nv * V;
    }
    for (int64_t i = tail + gid; i < n; i += stride) step(x[i], r[i], z[i], p[i], w[i], dinv[i]);
    double sums[2] = {zr, zz};
    blockReduceSums<kBlock, 2>(sums);
    if (threadIdx.x == 0) { zrPart[blockIdx.x] = sums[0]; zzPart[blockIdx.x] = sums[1]; }
}

// One pipelined Jacobi-PCG iteration (the FP32 inner solve): the update of
// iteration it (x += a p, r -= a w, z = D^{-1} r, partials of z.r and z.z into
// zrOut/zzOut) fused with the next search direction q = z + b p.  With
// w = Ap and a = (r, z) / (p, w):
// (r', z') = (r, z) - 2 a (w, z) + a^2 (w, D^{-1} w), and (w, z) = (w, p) since
// p = z + b_prev p_prev is A-conjugate to p_prev, so (r', z') =
// a^2 (w, D^{-1} w) - (r, z) (the recurrence behind single-reduction CG
// variants, Chronopoulos and Gear 1989).  b = (r', z') / (r, z) then needs only
// the SpMV's partials, and the direction kernel of the classic iteration is
// not launched.  The true (r', z')
// and (z', z') partials go to the next launch, whose first step is the stopping
// test on them (kCgDirection's test; when that kernel ran in this iteration,
// at the start of a chain or after a reliable update, it ran the test and
// stored (r, z)).  The partial sets alternate with it, since every block reads
// the previous launch's set while blocks that finished write theirs.
// SCALED: the iteration runs on D^{-1/2} A D^{-1/2} (unit diagonal), the same
// Jacobi-PCG in exact arithmetic (x~ = D^{1/2} x, r~ = D^{-1/2} r), where z = r
// and dinv is not read.
template <typename T, bool SCALED>
__global__ __launch_bounds__(kStepBlock) void kCgStep(int64_t n, int it, T* __restrict__ x, T* __restrict__ r,
                                                      const T* __restrict__ p, T* __restrict__ q,
                                                      const T* __restrict__ w, const T* __restrict__ dinv,
                                                      const double* __restrict__ dpiPart,
                                                      const double* __restrict__ wdwPart, int npart,
                                                      const double* __restrict__ zrIn, const double* __restrict__ zzIn,
                                                      double* __restrict__ zrOut, double* __restrict__ zzOut,
                                                      CgState* st, double* betaSlots, double* dpiSlots, double* hist,
                                                      int tested) {
    if (blockDone(st)) return;
    // the four scalars of the step in one block reduction; unless the
    // direction kernel ran this iteration's test (tested), zrIn / zzIn hold the
    // kStepBlocks partials of the previous step
    const bool test = !tested;
    double sums[4] = {test ? partialShare<kStepBlock>(zrIn, kStepBlocks) : 0.0,
                      test ? partialShare<kStepBlock>(zzIn, kStepBlocks) : 0.0,
                      partialShare<kStepBlock>(dpiPart, npart), partialShare<kStepBlock>(wdwPart, npart)};
    blockReduceSums<kStepBlock, 4>(sums);
    double rz;
    if (test) {
        rz = sums[0];
        const double dp = sqrt(sums[1]);
        int reason = 0;
        if (dp != dp) reason = -9;
        else if (dp <= st->ttol) reason = dp < st->atol ? 3 : 2;
        else if (dp >= st->dtol * st->rnorm0) reason = -4;
        else if (it >= st->maxit) reason = -3;
        else if (rz == 0.0) reason = 3;
        else if (rz * betaSlots[(it - 1) & 1] < 0.0) reason = -8;   // indefinite preconditioner
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->dp = dp;
            if (reason) { st->done = 1; st->reason = reason; st->its = reason == -8 ? it + 1 : it; }
            betaSlots[it & 1] = rz;
            if (hist && it < kHistMax) hist[2 * it] = rz;
        }
        if (reason) return;
    } else {
        rz = betaSlots[it & 1];
    }
    const double dpi = sums[2];
    const bool indefinite = dpi == 0.0 || (it > 0 && dpi * dpiSlots[(it - 1) & 1] <= 0.0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        dpiSlots[it & 1] = dpi;
        if (hist && it < kHistMax) hist[2 * it + 1] = dpi;
        if (indefinite) { st->done = 1; st->reason = -10; st->its = it + 1; }
    }
    if (indefinite) return;
    const double a = rz / dpi;
    double b = (a * a * sums[3] - rz) / rz;
    if (!(b > 0.0) || !(b < 1e300)) b = 0.0;   // the recurrence lost its sign: restart the direction
    double zr = 0.0, zz = 0.0;
    auto step = [&](T& xi, T& ri, T& qi, T pi, T wi, T di) {
        xi = (T)((double)xi + a * (double)pi);
        ri = (T)((double)ri - a * (double)wi);
        const T zi = SCALED ? ri : di * ri;
        qi = (T)((double)zi + b * (double)pi);
        zr += (double)zi * (double)ri;
        zz += (double)zi * (double)zi;
    };
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    int64_t tail = 0;
    if (aligned16(x, r, p, q, w, SCALED ? x : dinv)) {
        typedef typename Vec16<T>::type VT;
        constexpr int V = 16 / sizeof(T);
        const int64_t nv = n / V;
        for (int64_t i = gid; i < nv; i += stride) {
            VT xv = reinterpret_cast<const VT*>(x)[i], rv = reinterpret_cast<const VT*>(r)[i], qv, dv;
            const VT pv = reinterpret_cast<const VT*>(p)[i], wv = reinterpret_cast<const VT*>(w)[i];
            if constexpr (!SCALED) dv = reinterpret_cast<const VT*>(dinv)[i];
#pragma unroll
            for (int l = 0; l < V; ++l) {
                T xi = xv[l], ri = rv[l], qi;
                step(xi, ri, qi, pv[l], wv[l], dv[l]);
                xv[l] = xi;
                rv[l] = ri;
                qv[l] = qi;
            }
            reinterpret_cast<VT*>(x)[i] = xv;
            reinterpret_cast<VT*>(r)[i] = rv;
            reinterpret_cast<VT*>(q)[i] = qv;
        }
        tail = nv * V;
    }
    for (int64_t i = tail + gid; i < n; i += stride) step(x[i], r[i], q[i], p[i], w[i], SCALED ? (T)1 : dinv[i]);
    double out[2] = {zr, zz};
    blockReduceSums<kStepBlock, 2>(out);
    if (threadIdx.x == 0) { zrOut[blockIdx.x] = out[0]; zzOut[blockIdx.x] = out[1]; }
}

// r = b - w (w = A x0), z = D^{-1} r; partials of z.r and z.z.
template <typename T>
__global__ __launch_bounds__(kBlock) void kCgInit(int64_t n, const T* __restrict__ b, const T* __restrict__ w,
                                                  T* __restrict__ r, T* __restrict__ z, const T* __restrict__ dinv,
                                                  double* __restrict__ zrPart, double* __restrict__ zzPart) {
    double zr = 0.0, zz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T ri = b[i] - w[i];
        r[i] = ri;
        const T zi = dinv[i] * ri;
        z[i] = zi;
        zr += (double)zi * (double)ri;
        zz += (double)zi * (double)zi;
    }
    zr = blockReduceSum(zr);
    zz = blockReduceSum(zz);
    if (threadIdx.x == 0) { zrPart[blockIdx.x] = zr; zzPart[blockIdx.x] = zz; }
}

// r64 = b - w; r32 = s * r64 (for the FP32 inner solve); partial of ||D^{-1} r64||^2.
static __global__ __launch_bounds__(kBlock) void kResidual64(int64_t n, const double* __restrict__ b,
                                                      const double* __restrict__ w, const double* __restrict__ dinv,
                                                      double* __restrict__ r, double* __restrict__ zzPart) {
    double zz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ri = b[i] - w[i];
        r[i] = ri;
        const double zi = dinv[i] * ri;
        zz += zi * zi;
    }
    zz = blockReduceSum(zz);
    if (threadIdx.x == 0) zzPart[blockIdx.x] = zz;
}

static __global__ __launch_bounds__(kBlock) void kScaleToF32(int64_t n, const double* __restrict__ a, double s,
                                                      float* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)(s * a[i]);
}

// x += a (s o d) (s = D^{-1/2} when the FP32 iteration ran on the scaled
// image, else 1)
static __global__ __launch_bounds__(kBlock) void kAxpyF32(int64_t n, double a, const float* __restrict__ d,
                                                   const double* __restrict__ s, double* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] += a * (s ? s[i] * (double)d[i] : (double)d[i]);
}

// D^{-1/2} from D^{-1}; flag = 1 when some entry is not positive (the scaled
// image then does not exist)
static __global__ __launch_bounds__(kBlock) void kDiagSqrt(int64_t n, const double* __restrict__ dinv,
                                                    double* __restrict__ dsq, int* __restrict__ flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double d = dinv[i];
        if (d > 0.0) {
            dsq[i] = sqrt(d);
        } else {
            dsq[i] = 1.0;
            *flag = 1;
        }
    }
}

// The FP32 scalar image of D^{-1/2} A D^{-1/2} from the FP64 one: one
// wavefront per slice, 64 consecutive slots per step (slot base + o of the
// slice belongs to lane ((o >> 2) & 63), i.e. row 64 s + that lane; padding
// slots hold 0 and stay 0)
static __global__ __launch_bounds__(kBlock) void kSellScale32(int64_t n, int64_t nslices,
                                                       const uint32_t* __restrict__ slicePtr,
                                                       const int32_t* __restrict__ scol,
                                                       const double* __restrict__ s64,
                                                       const double* __restrict__ dsq, float* __restrict__ s32) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < nslices; s += nw) {
        const uint32_t base = slicePtr[s], len = slicePtr[s + 1] - base;
        for (uint32_t o = lane; o < len; o += 64) {
            const int64_t row = s * 64 + ((o >> 2) & 63);
            const size_t q = (size_t)base + o;
            const double ds = row < n ? dsq[row] : 0.0;
            s32[q] = (float)(s64[q] * ds * dsq[scol[q]]);
        }
    }
}

// reliable-update restart of the scaled FP32 iteration: r~ = D^{-1/2} r64 / sigma,
// x~ = 0, partials of r~.r~ into both z.r and z.z (z = r on the scaled image)
static __global__ __launch_bounds__(kBlock) void kCgScaledReset(int64_t n, const double* __restrict__ r64,
                                                         const double* __restrict__ dsq, double invSigma,
                                                         float* __restrict__ r, float* __restrict__ xs,
                                                         double* __restrict__ zrPart, double* __restrict__ zzPart) {
    double rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float ri = (float)(r64[i] * dsq[i] * invSigma);
        r[i] = ri;
        xs[i] = 0.0f;
        rr += (double)ri * (double)ri;
    }
    rr = blockReduceSum(rr);
    if (threadIdx.x == 0) { zrPart[blockIdx.x] = rr; zzPart[blockIdx.x] = rr; }
}

// Jacobi: D^{-1} from the stored diagonal; PCJACOBI treats a zero diagonal as 1.
static __global__ __launch_bounds__(kBlock) void kDiagInv(int64_t n, const int32_t* __restrict__ diagpos,
                                                   const double* __restrict__ v64, const float* __restrict__ v32,
                                                   double* __restrict__ d64, float* __restrict__ d32) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (d64) {
            const double d = v64 ? v64[diagpos[i]] : (double)v32[diagpos[i]];
            d64[i] = d != 0.0 ? 1.0 / d : 1.0;
        }
        if (d32) {
            const double d = v32 ? (double)v32[diagpos[i]] : v64[diagpos[i]];
            d32[i] = (float)(d != 0.0 ? 1.0 / d : 1.0);
        }
    }
}

static __global__ __launch_bounds__(kBlock) void kCastF32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// y = a x + b y (b = 0: y = a x)
static __global__ __launch_bounds__(kBlock) void kAxpby(int64_t n, double a, const double* __restrict__ x, double b,
                                                 double* __restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];
}

// v[r] = values ? values[r] : 0 for the listed rows
static __global__ __launch_bounds__(kBlock) void kSetRows(int64_t nrows, const int32_t* __restrict__ rows,
                                                   const double* __restrict__ values, double* __restrict__ v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * blockDim.x)
        v[rows[i]] = values ? values[rows[i]] : 0.0;
}

static __global__ __launch_bounds__(kBlock) void kDotPartial(int64_t n, const double* __restrict__ a,
                                                      const double* __restrict__ b, double* __restrict__ part) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += a[i] * b[i];
    s = blockReduceSum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// sequential ADD of host-supplied entries at known CSR slots (MatrixSparse::addMatrix slow path)
static __global__ void kAddAt(int64_t m, const int64_t* __restrict__ pos, const double* __restrict__ vals,
                       double* __restrict__ v64, float* __restrict__ v32) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int64_t k = 0; k < m; ++k) {
        if (v64) {
            v64[pos[k]] += vals[k];
            if (v32) v32[pos[k]] = (float)v64[pos[k]];
        } else if (v32) {
            v32[pos[k]] = (float)((double)v32[pos[k]] + vals[k]);
        }
    }
}

// ---------------------------------------------------------------- SELL solver format
template <int BS>
__device__ inline double blockReduceSumN(double v) {
    __shared__ double sh[BS / 64];
    __shared__ double total;
    v = waveReduceSum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
#pragma unroll
        for (int i = 0; i < BS / 64; ++i) s += sh[i];
        total = s;
    }
    __syncthreads();
    const double r = total;
    __syncthreads();
    return r;
}

// Column stream of the SELL-C-sigma image (shared by the SpMV kernels).
// COMPACT: columns come from the packed stream, where a slice spanning fewer
// than 2^16 columns stores 16-bit offsets from colBase[s] (2 bytes per slot)
// and any other slice (colBase[s] = -1) stores int32 columns.
struct SellCols {
    const uint32_t* __restrict__ slicePtr;
    const int32_t* __restrict__ col;
    const int32_t* __restrict__ colBase;
    const int64_t* __restrict__ cptr;
    const uint16_t* __restrict__ packed;
};

// Slot order inside a slice of the SELL image: the slice's width (padded to a
// multiple of 4) is cut into quads of 4 consecutive entries of each row, and
// quad q holds lane l's entries 4q..4q+3 at 256q + 4l + (0..3).  A lane then
// reads each quad of values and columns with one 16-byte (8 for 16-bit
// columns) load, and a wavefront's load covers 1 KiB of values.
constexpr int kSellQuad = 4;
__device__ inline size_t sellSlot(uint32_t base, int k, int lane) {
    return (size_t)base + ((size_t)(k >> 2) << 8) + ((size_t)lane << 2) + (size_t)(k & 3);
}

template <typename T>
struct Quad;
template <>
struct Quad<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Quad<double> {
    typedef double type __attribute__((ext_vector_type(4)));
};
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef uint8_t u8x4 __attribute__((ext_vector_type(4)));

// Row (s*64 + lane) of A x for slice s, executed by one wavefront: the
// lane's quads of values and columns are streamed, UQ quads per step with
// 4 x UQ independent partial sums for memory-level parallelism.  FP32 values
// (one 16-byte load per quad) are read non-temporally, keeping L2 for the x
// gathers; FP64 values take two 16-byte loads per quad, which need the line
// to stay in L2 between them, and columns (8 or 16 bytes) likewise.
template <typename T, int UQ, bool COMPACT>
__device__ inline T sellRow(const SellCols& S, int64_t s, int lane, const T* __restrict__ val,
                            const T* __restrict__ x) {
    typedef typename Quad<T>::type V;
    const uint32_t base = S.slicePtr[s];
    const int nq = (int)((S.slicePtr[s + 1] - base) >> 8);
    const V* v = reinterpret_cast<const V*>(val + base) + lane;
    T acc[UQ][4];
#pragma unroll
    for (int u = 0; u < UQ; ++u) acc[u][0] = acc[u][1] = acc[u][2] = acc[u][3] = 0;
    // one wave per slice: the narrow/wide choice is uniform across the wave
    const int32_t cb = COMPACT ? S.colBase[s] : -1;
    if (cb >= 0) {
        const u16x4* c = reinterpret_cast<const u16x4*>(S.packed + S.cptr[s]) + lane;
        const T* xs = x + cb;
        int k = 0;
        for (; k + UQ <= nq; k += UQ) {
            V vv[UQ];
            u16x4 cc[UQ];
#pragma unroll
            for (int u = 0; u < UQ; ++u) {
                vv[u] = sizeof(T) == 4 ? __builtin_nontemporal_load(v + (size_t)(k + u) * 64) : v[(size_t)(k + u) * 64];
                cc[u] = c[(size_t)(k + u) * 64];
            }
#pragma unroll
            for (int u = 0; u < UQ; ++u) {
                acc[u][0] += vv[u].x * xs[cc[u].x];
                acc[u][1] += vv[u].y * xs[cc[u].y];
                acc[u][2] += vv[u].z * xs[cc[u].z];
                acc[u][3] += vv[u].w * xs[cc[u].w];
            }
        }
        for (; k < nq; ++k) {
            const V vv = v[(size_t)k * 64];
            const u16x4 cc = c[(size_t)k * 64];
            acc[0][0] += vv.x * xs[cc.x];
            acc[0][1] += vv.y * xs[cc.y];
            acc[0][2] += vv.z * xs[cc.z];
            acc[0][3] += vv.w * xs[cc.w];
        }
    } else {
        const i32x4* c = (COMPACT ? reinterpret_cast<const i32x4*>(S.packed + S.cptr[s])
                                  : reinterpret_cast<const i32x4*>(S.col + base)) + lane;
        int k = 0;
        for (; k + UQ <= nq; k += UQ) {
            V vv[UQ];
            i32x4 cc[UQ];
#pragma unroll
            for (int u = 0; u < UQ; ++u) {
                vv[u] = sizeof(T) == 4 ? __builtin_nontemporal_load(v + (size_t)(k + u) * 64) : v[(size_t)(k + u) * 64];
                cc[u] = c[(size_t)(k + u) * 64];
            }
#pragma unroll
            for (int u = 0; u < UQ; ++u) {
                acc[u][0] += vv[u].x * x[cc[u].x];
                acc[u][1] += vv[u].y * x[cc[u].y];
                acc[u][2] += vv[u].z * x[cc[u].z];
                acc[u][3] += vv[u].w * x[cc[u].w];
            }
        }
        for (; k < nq; ++k) {
            const V vv = v[(size_t)k * 64];
            const i32x4 cc = c[(size_t)k * 64];
            acc[0][0] += vv.x * x[cc.x];
            acc[0][1] += vv.y * x[cc.y];
            acc[0][2] += vv.z * x[cc.z];
            acc[0][3] += vv.w * x[cc.w];
        }
    }
    T a = 0;
#pragma unroll
    for (int u = 0; u < UQ; ++u) a += (acc[u][0] + acc[u][1]) + (acc[u][2] + acc[u][3]);
    return a;
}

// y = A x on the SELL-C-sigma image: one wavefront per 64-row slice, one lane
// per row, grid-stride over slices with a fixed grid.  With part, also the
// block partial of x.y (p.Ap); with part2, that of y.D^{-1}y (dinv = D^{-1}),
// which the pipelined CG iteration needs (y.y when dinv is null: the
// symmetrically scaled image).
template <typename T, int BS, int U, bool COMPACT>
__global__ __launch_bounds__(BS) void kSellSpmv(int64_t nrows, int64_t nslices, SellCols S, const T* __restrict__ val,
                                                const T* __restrict__ x, T* __restrict__ y, double* __restrict__ part,
                                                const CgState* __restrict__ st, const T* __restrict__ dinv,
                                                double* __restrict__ part2) {
    if (st && blockDone(st)) return;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * BS + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * BS) >> 6;
    double dot = 0.0, dot2 = 0.0;
    for (int64_t s = wave; s < nslices; s += nwaves) {
        const T a = sellRow<T, U, COMPACT>(S, s, lane, val, x);
        const int64_t r = s * 64 + lane;
        if (r < nrows) {
            y[r] = a;
            if (part) dot += (double)x[r] * (double)a;
            if (part2) dot2 += (double)a * (double)a * (dinv ? (double)dinv[r] : 1.0);
        }
    }
    if (part2) {
        double t[2] = {dot, dot2};
        blockReduceSums<BS, 2>(t);
        if (threadIdx.x == 0) {
            if (part) part[blockIdx.x] = t[0];
            part2[blockIdx.x] = t[1];
        }
    } else if (part) {
        const double t = blockReduceSumN<BS>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// y = A x on the 3 x 3 block image of a vector space: one wavefront per slice,
// L = 2^sliceShift[s] lanes per node row (64 / L rows from sliceRow[s]), each
// lane taking every L-th block of its row, the L lanes' sums added at the end;
// per block slot one column index, nine values (planes of 64, read
// non-temporally) and three consecutive x entries.  UB block slots per step (4
// for FP32, 2 for FP64: measured) keep independent partial sums.  With part,
// also the block partial of x.y (p.Ap).
template <typename T, int BS>
__global__ __launch_bounds__(BS) void kBlockSellSpmv(int64_t nslices, const uint32_t* __restrict__ slicePtr,
                                                     const int32_t* __restrict__ sliceRow,
                                                     const uint8_t* __restrict__ sliceShift,
                                                     const int32_t* __restrict__ bcol, const T* __restrict__ val,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     double* __restrict__ part, const CgState* __restrict__ st) {
    constexpr int UB = sizeof(T) == 4 ? 4 : 2;
    if (st && blockDone(st)) return;
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * BS + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * BS) >> 6;
    double dot = 0.0;
    for (int64_t s = wave; s < nslices; s += nwaves) {
        const uint32_t base = slicePtr[s];
        const int width = (int)((slicePtr[s + 1] - base) >> 6);
        const int sh = __builtin_amdgcn_readfirstlane((int)sliceShift[s]);
        const int32_t* c = bcol + base + lane;
        const T* v = val + 9 * (size_t)base + lane;
        T acc[UB][3];
#pragma unroll
        for (int u = 0; u < UB; ++u) acc[u][0] = acc[u][1] = acc[u][2] = 0;
        int k = 0;
        for (; k + UB <= width; k += UB) {
            int32_t j[UB];
            T vv[UB][9];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                j[u] = c[(size_t)(k + u) * 64];
#pragma unroll
                for (int q = 0; q < 9; ++q) vv[u][q] = __builtin_nontemporal_load(v + ((size_t)(k + u) * 9 + q) * 64);
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const T x0 = x[3 * (int64_t)j[u]], x1 = x[3 * (int64_t)j[u] + 1], x2 = x[3 * (int64_t)j[u] + 2];
                acc[u][0] += vv[u][0] * x0 + vv[u][1] * x1 + vv[u][2] * x2;
                acc[u][1] += vv[u][3] * x0 + vv[u][4] * x1 + vv[u][5] * x2;
                acc[u][2] += vv[u][6] * x0 + vv[u][7] * x1 + vv[u][8] * x2;
            }
        }
        for (; k < width; ++k) {
            const int32_t jj = c[(size_t)k * 64];
            const T x0 = x[3 * (int64_t)jj], x1 = x[3 * (int64_t)jj + 1], x2 = x[3 * (int64_t)jj + 2];
            const T* w = v + (size_t)k * 9 * 64;
            acc[0][0] += w[0] * x0 + w[64] * x1 + w[128] * x2;
            acc[0][1] += w[192] * x0 + w[256] * x1 + w[320] * x2;
            acc[0][2] += w[384] * x0 + w[448] * x1 + w[512] * x2;
        }
        T a[3];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            a[cc] = 0;
#pragma unroll
            for (int u = 0; u < UB; ++u) a[cc] += acc[u][cc];
            for (int o = 1; o < (1 << sh); o <<= 1) a[cc] += __shfl_xor(a[cc], o);   // the L lanes of the row
        }
        const int64_t r = (int64_t)sliceRow[s] + (lane >> sh);
        if (r < sliceRow[s + 1] && (lane & ((1 << sh) - 1)) == 0) {
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                y[3 * r + cc] = a[cc];
                if (part) dot += (double)x[3 * r + cc] * (double)a[cc];
            }
        }
    }
    if (part) {
        const double t = blockReduceSumN<BS>(dot);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// Reliable update, device half: r32 = r64 / sigma, z32 = D^{-1} r32, xs = 0,
// partials of r.z and z.z for the next direction kernel.
static __global__ __launch_bounds__(kBlock) void kCgStepReset(int64_t n, const double* __restrict__ r64, double invSigma,
                                                       const float* __restrict__ dinv, float* __restrict__ r,
                                                       float* __restrict__ z, float* __restrict__ xs,
                                                       double* __restrict__ gzPart, double* __restrict__ zzPart) {
    double zr = 0.0, zz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float ri = (float)(r64[i] * invSigma);
        r[i] = ri;
        const float zi = dinv[i] * ri;
        z[i] = zi;
        xs[i] = 0.0f;
        zr += (double)zi * (double)ri;
        zz += (double)zi * (double)zi;
    }
    zr = blockReduceSum(zr);
    zz = blockReduceSum(zz);
    if (threadIdx.x == 0) { gzPart[blockIdx.x] = zr; zzPart[blockIdx.x] = zz; }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kCsrToSell(int64_t nslots, const int32_t* __restrict__ slot2csr,
                                                     const double* __restrict__ c64, const float* __restrict__ c32,
                                                     T* __restrict__ out) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = slot2csr[s];
        out[s] = p < 0 ? (T)0 : (c64 ? (T)c64[p] : (T)c32[p]);
    }
}

// Both SELL value images in one pass over the slot map: FP64 from the FP64
// CSR values, FP32 from the FP32 ones (the FP64 ones rounded when the matrix
// stores none).
static __global__ __launch_bounds__(kBlock) void kCsrToSellBoth(int64_t nslots, const int32_t* __restrict__ slot2csr,
                                                                const double* __restrict__ c64,
                                                                const float* __restrict__ c32,
                                                                double* __restrict__ s64, float* __restrict__ s32) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = slot2csr[s];
        const double v = p < 0 ? 0.0 : c64[p];
        s64[s] = v;
        s32[s] = c32 ? (p < 0 ? 0.0f : c32[p]) : (float)v;
    }
}

// D^{-1} in the solver numbering from the SELL diagonal slots
static __global__ __launch_bounds__(kBlock) void kDiagInvSell(int64_t n, const int32_t* __restrict__ sdiag,
                                                              const double* __restrict__ s64,
                                                              const float* __restrict__ s32, double* __restrict__ d64,
                                                              float* __restrict__ d32) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (d64 && s64) {
            const double d = s64[sdiag[i]];
            d64[i] = d != 0.0 ? 1.0 / d : 1.0;
        }
        if (d32) {   // from the FP64 diagonal when there is one, so D^{-1} does not depend on the FP32 image
            const double d = s64 ? s64[sdiag[i]] : (double)s32[sdiag[i]];
            d32[i] = (float)(d != 0.0 ? 1.0 / d : 1.0);
        }
    }
}

// out[i] = in[perm[i]] (old -> solver numbering) and back
static __global__ __launch_bounds__(kBlock) void kPermuteIn(int64_t n, const int32_t* __restrict__ perm,
                                                            const double* __restrict__ in, double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[perm[i]];
}

static __global__ __launch_bounds__(kBlock) void kPermuteOut(int64_t n, const int32_t* __restrict__ perm,
                                                             const double* __restrict__ in, double* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[perm[i]] = in[i];
}

}  // namespace feelpp_amd
