// bf_knn.hip -- exact k-nearest neighbours by brute force, the ground truth of recall@k.
//
// One wavefront owns 16 queries and streams every base row: distances come from
// v_mfma_i32_16x16x64_i8 (two K-steps cover a 128-byte row), d = |q|^2 + |p|^2 - 2 q.p
// in exact int32.  Selection is fused: each query keeps its best k keys (d << 32 | id)
// in LDS and a threshold (the k-th key); rows below it are appended to a per-query
// buffer that is merged into the best-k list with an in-register bitonic sort whenever
// it could overflow.  Keys are unique, so the result is the exact top-k in (d, id)
// order, with no n x nq distance matrix written to HBM.
#include "pann_internal.h"

namespace pann {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kTopCap = 128;  // best-k list capacity per query (k <= 128)
constexpr int kBufCap = 128;  // candidate buffer per query
constexpr int kTileCols = 32; // base rows per iteration (two 16-column MFMA tiles)
constexpr int kQ = 16;        // queries per wavefront

__device__ __forceinline__ int laneId() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0));
}

__device__ __forceinline__ uint64_t shflXor64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

template <int C>
__device__ __forceinline__ void bitonicSort(uint64_t (&key)[C], int lane) {
    constexpr int N = 64 * C;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int c2 = c ^ (j >> 6);
                    if (c < c2) {
                        const bool asc = ((c * 64 + lane) & k) == 0;
                        const uint64_t a = key[c], b = key[c2];
                        const bool sw = asc ? (a > b) : (a < b);
                        key[c] = sw ? b : a;
                        key[c2] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const bool asc = ((c * 64 + lane) & k) == 0;
                    const bool lower = (lane & j) == 0;
                    const uint64_t p = shflXor64(key[c], j);
                    const uint64_t mn = key[c] < p ? key[c] : p;
                    const uint64_t mx = key[c] < p ? p : key[c];
                    key[c] = (lower == asc) ? mn : mx;
                }
            }
        }
    }
}

struct BfLds {
    uint64_t top[kQ][kTopCap];
    uint64_t buf[kQ][kBufCap];
    uint32_t ntop[kQ];
    uint32_t nbuf[kQ];
    uint64_t thr[kQ];
};

// merge query row r's buffer into its best-k list (whole wavefront)
__device__ void mergeRow(BfLds& S, int r, uint32_t k, int lane) {
    uint64_t key[4];
    const uint32_t nt = S.ntop[r], nb = S.nbuf[r];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const uint32_t i = c * 64 + lane;
        key[c] = i < nt ? S.top[r][i] : ~0ull;
        key[c + 2] = i < nb ? S.buf[r][i] : ~0ull;
    }
    bitonicSort<4>(key, lane);
    const uint32_t total = min(k, nt + nb);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const uint32_t i = c * 64 + lane;
        if (i < total) S.top[r][i] = key[c];
    }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        S.ntop[r] = total;
        S.nbuf[r] = 0;
        S.thr[r] = (total >= k && k > 0) ? S.top[r][k - 1] : ~0ull;
    }
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(64) void bfKnnKernel(const int8_t* base, const int32_t* bnorm,
                                                 uint32_t n, const int8_t* queries,
                                                 const int32_t* qnorm, uint32_t nq, uint32_t k,
                                                 bool normsInRow, uint32_t* outIds,
                                                 int32_t* outDists) {
    __shared__ BfLds S;
    const int lane = laneId();
    const uint32_t q0 = blockIdx.x * kQ;
    const int col = lane & 15, grp = lane >> 4;

    if (lane < kQ) {
        S.ntop[lane] = 0;
        S.nbuf[lane] = 0;
        S.thr[lane] = ~0ull;
    }
    // A operand: query row (q0 + col), bytes [16*grp, +16) and [64 + 16*grp, +16)
    v4i a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
    if (q0 + col < nq) {
        const int8_t* qr = queries + (size_t)(q0 + col) * kVecBytes;
        a0 = *reinterpret_cast<const v4i*>(qr + 16 * grp);
        a1 = *reinterpret_cast<const v4i*>(qr + 64 + 16 * grp);
        // bytes 124..127 hold norms: zero the query's so q.p ignores them
        if (normsInRow && grp == 3) a1.w = 0;
    }
    // output rows held by this lane: 4*grp + r
    int qn[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t q = q0 + 4 * grp + r;
        qn[r] = q < nq ? qnorm[q] : 0;
    }
    __builtin_amdgcn_wave_barrier();
    uint64_t thr[4] = {~0ull, ~0ull, ~0ull, ~0ull};

    for (uint32_t p0 = 0; p0 < n; p0 += kTileCols) {
#pragma unroll
        for (int t = 0; t < kTileCols / 16; t++) {
            const uint32_t p = p0 + 16 * t + col;
            v4i b0 = {0, 0, 0, 0}, b1 = {0, 0, 0, 0};
            int pn = 0;
            if (p < n) {
                const int8_t* pr = base + (size_t)p * kVecBytes;
                b0 = *reinterpret_cast<const v4i*>(pr + 16 * grp);
                b1 = *reinterpret_cast<const v4i*>(pr + 64 + 16 * grp);
                pn = bnorm[p];
            }
            v4i acc = {0, 0, 0, 0};
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a0, b0, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a1, b1, acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int d = qn[r] + pn - 2 * acc[r];
                const uint64_t key = ((uint64_t)(uint32_t)d << 32) | p;
                const int row = 4 * grp + r;
                if (p < n && q0 + row < nq && key < thr[r]) {
                    const uint32_t slot = atomicAdd(&S.nbuf[row], 1u);
                    S.buf[row][slot] = key;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        // a tile adds at most kTileCols keys per query: merge rows that could overflow
        const uint32_t nb = lane < kQ ? S.nbuf[lane] : 0u;
        uint64_t need = __ballot(lane < kQ && nb + kTileCols > (uint32_t)kBufCap);
        if (need) {
            while (need) {
                const int r = __ffsll((unsigned long long)need) - 1;
                mergeRow(S, r, k, lane);
                need &= need - 1;
            }
#pragma unroll
            for (int r = 0; r < 4; r++) thr[r] = S.thr[4 * grp + r];
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (int r = 0; r < kQ; r++)
        if (q0 + r < nq) mergeRow(S, r, k, lane);
    for (int r = 0; r < kQ; r++) {
        const uint32_t q = q0 + r;
        if (q >= nq) break;
        for (uint32_t j = lane; j < k; j += 64) {
            const bool have = j < S.ntop[r];
            const uint64_t key = have ? S.top[r][j] : ~0ull;
            outIds[(size_t)q * k + j] = have ? (uint32_t)key : kEmpty;
            outDists[(size_t)q * k + j] = have ? (int32_t)(key >> 32) : 0x7FFFFFFF;
        }
    }
}

}  // namespace

hipError_t launchBruteForceKnn(const int8_t* base, const int32_t* bnorm, uint32_t n,
                               const int8_t* queries, const int32_t* qnorm, uint32_t nq,
                               uint32_t k, bool normsInRow, uint32_t* outIds, int32_t* outDists,
                               hipStream_t s) {
    if (nq == 0) return hipSuccess;
    if (k == 0 || k > (uint32_t)kTopCap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bfKnnKernel, dim3((nq + kQ - 1) / kQ), dim3(64), 0, s, base, bnorm, n,
                       queries, qnorm, nq, k, normsInRow, outIds, outDists);
    return hipGetLastError();
}

}  // namespace pann
