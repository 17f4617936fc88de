// bf_knn.hip -- exact k-nearest neighbours by brute force, the ground truth of recall@k.
//
// Exact top-k (k <= 128) of every query over all base rows, in (distance, id) order, as an
// int8 GEMM on v_mfma_i32_16x16x64_i8 with the selection split so that the GEMM itself
// does one compare per distance:
//
//   acc(q, p) = q.p - floor(|p|^2 / 2)     (the MFMA accumulator starts at -floor(|p|^2/2))
//   d(q, p)   = |q|^2 + (|p|^2 & 1) - 2 acc(q, p)
//
// so for one query a larger acc means a smaller d, and every p with acc >= A has a
// smaller d than every p with acc < A (the odd bit moves d by at most 1, a step of acc
// by 2).  Hence {p : acc(q, p) >= A} is always a prefix of the (d, id) order.
//
//   1. threshold pass: over a sample of base blocks, each lane keeps the largest acc it
//      sees per query in each of L disjoint subsets (its column of the 16-row blocks, and
//      the block number mod L); the query's threshold A is the smallest of those 16 L
//      maxima, so at least 16 L distinct points reach it.
//   2. filter pass: the full GEMM over all base rows; a distance is a candidate iff
//      acc >= A, one v_cmp per distance; the rare candidates go to per-(query, slice)
//      lists as (d << 32 | id) keys.
//   3. select pass: one workgroup per query sorts its candidates and writes the first k.
//   A query whose lists overflowed is filtered again with lists sized from its exact
//   per-slice counts; one with fewer than k candidates (or too many to sort) is searched
//   by the one-pass kernel (bfKnnExactKernel), which keeps its best k in LDS while it
//   streams every base row.
//
// The GEMM passes: 64 queries per wavefront (the A operand, 32 VGPRs, stays in
// registers), four wavefronts per workgroup over the same base rows (the later ones find
// them in L1/L2), each lane keeping the loads of the next four 16-row blocks in flight.
// The base is split into slices so the grid fills the chip, and the workgroups of one XCD
// (blockIdx mod 8) take consecutive jobs of one slice so that its L2 holds the rows they
// share.  No nq x n distance matrix is written.
#include "pann_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace pann {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kGemmWaves = 4;                            // wavefronts per workgroup
constexpr int kRowsPerWave = 64;                         // queries per wavefront
constexpr int kQueriesPerBlock = kGemmWaves * kRowsPerWave;
constexpr int32_t kPadAcc = -(1 << 28);                  // accumulator start past the last row
constexpr int kSelectThreads = 256;
constexpr uint32_t kCandPerQuery = 2048;                 // candidate slots per query (all slices)
constexpr uint32_t kSelectCap = 8192;                    // keys one select workgroup sorts
constexpr uint32_t kSmallBase = 262144;                  // up to this n the sample is the whole base
constexpr int kTopCap = 128;                             // k <= 128
constexpr uint32_t kMaxSlices = 128;                     // slices of one pass

__device__ __forceinline__ int laneId() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0));
}

__device__ __forceinline__ uint64_t shflXor64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

struct GemmArgs {
    const int8_t* base;
    const int32_t* bnorm;
    uint32_t n;
    const int8_t* queries;
    const int32_t* qnorm;
    const uint32_t* qlist; // nullable: row i of the pass is query qlist[i] (else i)
    uint32_t nsel;         // rows of the pass
    int32_t normsInRow;
    uint32_t qTiles;       // workgroups of kQueriesPerBlock rows
    uint32_t slices;       // parts of the scanned blocks
    uint32_t jobs;         // qTiles * slices
    uint32_t nBlocks;      // 16-row blocks of the base
    uint32_t scanBlocks;   // blocks scanned by this pass (the sample, or all)
    uint32_t stride;       // scan block s is base block s * stride
    int32_t* accmax;       // threshold pass: per row 16 L maxima (start at INT32_MIN)
    const int32_t* thr;    // filter pass: candidate iff acc > thr[query]
    uint64_t* cand;        // filter pass: keys, row-major (row, slice) lists
    uint32_t* counts;      // filter pass: nsel x slices candidate counts
    uint32_t capS;         // list length when offs is null
    const uint64_t* offs;  // nullable: list (i, s) is cand[offs[i S + s] .. offs[i S + s + 1])
};

__device__ __forceinline__ uint32_t queryOf(const GemmArgs& a, uint32_t i) {
    return a.qlist ? a.qlist[i] : i;
}

// A operand: rows q0 + 16 rb + col, bytes [16 grp, +16) and [64 + 16 grp, +16); a query's
// norm bytes 124..127 are zeroed so that q.p ignores them.
__device__ __forceinline__ void loadQueries(const GemmArgs& a, uint32_t q0, int col, int grp,
                                            v4i (&A0)[4], v4i (&A1)[4]) {
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        const uint32_t i = q0 + 16 * rb + col;
        A0[rb] = v4i{0, 0, 0, 0};
        A1[rb] = v4i{0, 0, 0, 0};
        if (i < a.nsel) {
            const int8_t* r = a.queries + (size_t)queryOf(a, i) * kVecBytes;
            A0[rb] = *reinterpret_cast<const v4i*>(r + 16 * grp);
            A1[rb] = *reinterpret_cast<const v4i*>(r + 64 + 16 * grp);
            if (a.normsInRow && grp == 3) A1[rb].w = 0;
        }
    }
}

// Scan block s of a pass is base block s * stride.
__device__ __forceinline__ uint32_t baseBlockOf(const GemmArgs& a, uint32_t s) {
    return s * a.stride;
}

// Scan block s for lane (col, grp): its B operand (base row 16 baseBlockOf(s) + col) and
// |p|^2, or -1 past the slice end or the base (nothing loaded).  The accumulator start and
// the norm's low bit are derived where the block is used, so that nothing waits for a
// load at the time it is issued.
struct BBlock {
    v4i b0, b1;
    int32_t pn;
    __device__ bool valid() const { return pn >= 0; }
    __device__ int32_t start() const { return pn >= 0 ? -(pn >> 1) : kPadAcc; }
    __device__ int32_t odd() const { return pn & 1; }
};

__device__ __forceinline__ BBlock loadBlock(const GemmArgs& a, uint32_t s, uint32_t s1, int col,
                                            int grp) {
    BBlock B;
    const uint32_t p = s < s1 ? baseBlockOf(a, s) * 16 + col : 0xFFFFFFFFu;
    B.b0 = v4i{0, 0, 0, 0};
    B.b1 = v4i{0, 0, 0, 0};
    B.pn = -1;
    if (p < a.n) {
        const int8_t* r = a.base + (size_t)p * kVecBytes;
        B.b0 = *reinterpret_cast<const v4i*>(r + 16 * grp);
        B.b1 = *reinterpret_cast<const v4i*>(r + 64 + 16 * grp);
        B.pn = a.bnorm[p];
    }
    return B;
}

__device__ __forceinline__ void blockProducts(const v4i (&A0)[4], const v4i (&A1)[4],
                                              const BBlock& B, v4i (&acc)[4]) {
    const int32_t st = B.start();
    const v4i c = {st, st, st, st};
    // the four independent first halves, then the second halves that depend on them
#pragma unroll
    for (int rb = 0; rb < 4; rb++) acc[rb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A0[rb], B.b0, c, 0, 0, 0);
#pragma unroll
    for (int rb = 0; rb < 4; rb++) acc[rb] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A1[rb], B.b1, acc[rb], 0, 0, 0);
}

// The job of a workgroup: workgroup b runs on XCD b mod 8; each XCD takes a contiguous
// range of the jobs, which are ordered slice-major, so the workgroups resident on one XCD
// scan the same slice.
__device__ __forceinline__ bool jobOf(const GemmArgs& a, uint32_t* slice, uint32_t* qtile) {
    const uint32_t per = (a.jobs + 7) / 8;
    const uint32_t j = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (j >= a.jobs) return false;
    *slice = j / a.qTiles;
    *qtile = j % a.qTiles;
    return true;
}

// Base rows reach the four wavefronts of a workgroup through LDS: the workgroup copies a
// chunk of kChunk scan blocks (16 KB) into one of two buffers while it computes on the
// other, so each row is fetched from L2 once per workgroup instead of once per
// wavefront.  Piece p (16 bytes) of chunk row r sits at rows[8 r + (p ^ (r & 7))]: the
// eight lanes that read one piece of eight consecutive rows hit eight bank groups.
constexpr int kChunk = 8;
constexpr int kChunkRows = 16 * kChunk;
constexpr int kPieces = kVecBytes / 16;

struct Chunk {
    v4i rows[kChunkRows * kPieces];
    int32_t pn[kChunkRows];  // |p|^2, or -1 past the slice or the base
};

// What one thread moves into a chunk: four pieces, and one norm for the first 128 threads.
struct Staging {
    v4i piece[4];
    int32_t pn;
};

__device__ __forceinline__ void fetchChunk(const GemmArgs& a, uint32_t c0, uint32_t s1,
                                           Staging& st) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const uint32_t i = t + 256u * u;
        const uint32_t r = i / kPieces;
        const uint32_t s = c0 + r / 16;
        const uint32_t row = s < s1 ? baseBlockOf(a, s) * 16 + r % 16 : 0xFFFFFFFFu;
        st.piece[u] = v4i{0, 0, 0, 0};
        if (row < a.n)
            st.piece[u] = *reinterpret_cast<const v4i*>(a.base + (size_t)row * kVecBytes + 16 * (i % kPieces));
    }
    st.pn = -1;
    if (t < (uint32_t)kChunkRows) {
        const uint32_t s = c0 + t / 16;
        const uint32_t row = s < s1 ? baseBlockOf(a, s) * 16 + t % 16 : 0xFFFFFFFFu;
        if (row < a.n) st.pn = a.bnorm[row];
    }
}

__device__ __forceinline__ void storeChunk(Chunk& C, const Staging& st) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const uint32_t i = t + 256u * u;
        const uint32_t r = i / kPieces, p = i % kPieces;
        C.rows[r * kPieces + (p ^ (r & 7))] = st.piece[u];
    }
    if (t < (uint32_t)kChunkRows) C.pn[t] = st.pn;
}

__device__ __forceinline__ BBlock readBlock(const Chunk& C, int b, int col, int grp) {
    BBlock B;
    const uint32_t r = 16 * b + col;
    B.b0 = C.rows[r * kPieces + (grp ^ (r & 7))];
    B.b1 = C.rows[r * kPieces + ((4 + grp) ^ (r & 7))];
    B.pn = C.pn[r];
    return B;
}

// Runs body(B, s) for every scan block s of [s0, s1), chunk by chunk, and after each
// chunk between(last) (no chunk load is in flight then).  Uniform over the workgroup.
template <class Body, class Between>
__device__ __forceinline__ void scanChunks(const GemmArgs& a, Chunk (&chunk)[2], uint32_t s0,
                                           uint32_t s1, Body body, Between between) {
    Staging st;
    fetchChunk(a, s0, s1, st);
    storeChunk(chunk[0], st);
    __syncthreads();
    int cur = 0;
    for (uint32_t c0 = s0; c0 < s1; c0 += kChunk) {
        const bool more = c0 + kChunk < s1;
        if (more) fetchChunk(a, c0 + kChunk, s1, st);
        const uint32_t nb = min((uint32_t)kChunk, s1 - c0);
#pragma unroll 1
        for (uint32_t b = 0; b < nb; b++) body(readBlock(chunk[cur], b, laneId() & 15, laneId() >> 4), c0 + b);
        if (more) storeChunk(chunk[cur ^ 1], st);
        between(!more);
        __syncthreads();
        cur ^= 1;
    }
}

// ---- pass 1: per query, lane column and subset, the largest acc over a sample ----
template <int L>
__global__ __launch_bounds__(256) void bfThresholdKernel(GemmArgs a) {
    __shared__ Chunk chunk[2];
    const int lane = laneId(), wave = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4;
    uint32_t slice, qtile;
    if (!jobOf(a, &slice, &qtile)) return;
    const uint32_t q0 = qtile * kQueriesPerBlock + wave * kRowsPerWave;
    v4i A0[4], A1[4];
    loadQueries(a, q0, col, grp, A0, A1);
    int32_t best[L][16];
#pragma unroll
    for (int u = 0; u < L; u++)
#pragma unroll
        for (int e = 0; e < 16; e++) best[u][e] = INT32_MIN;
    const uint32_t s0 = (uint32_t)((uint64_t)a.scanBlocks * slice / a.slices);
    const uint32_t s1 = (uint32_t)((uint64_t)a.scanBlocks * (slice + 1) / a.slices);
    scanChunks(a, chunk, s0, s1, [&](const BBlock& B, uint32_t sb) {
        v4i acc[4];
        blockProducts(A0, A1, B, acc);
        const int sub = (int)(sb % L);  // the subset of this scan block
#pragma unroll
        for (int v = 0; v < L; v++) {
            if (v == sub) {
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        best[v][rb * 4 + j] = max(best[v][rb * 4 + j], acc[rb][j]);
            }
        }
    }, [](bool) {});
    // subset v * 16 + col of the query; the slices of the sample combine by atomicMax
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t i = q0 + 16 * rb + 4 * grp + j;
            if (i < a.nsel) {
#pragma unroll
                for (int v = 0; v < L; v++)
                    atomicMax(&a.accmax[(size_t)i * 16 * L + v * 16 + col], best[v][rb * 4 + j]);
            }
        }
}

// thr[query of row i] = (smallest of the row's 16 L maxima) - 1: a candidate is acc > thr.
__global__ __launch_bounds__(256) void bfReduceThresholdKernel(const int32_t* accmax, int subsets,
                                                              const uint32_t* qlist, uint32_t nsel,
                                                              int32_t* thr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsel) return;
    int32_t m = INT32_MAX;
    for (int t = 0; t < subsets; t++) m = min(m, accmax[(size_t)i * subsets + t]);
    thr[qlist ? qlist[i] : i] = m == INT32_MIN ? INT32_MIN : m - 1;
}

// ---- pass 2: the full GEMM; candidates acc > thr[query] to per-(row, slice) lists ----
// Candidates wait in a per-wavefront LDS pool and go to HBM between chunks, once half the
// pool is used: global stores among the block computations would make the waits for the
// next chunk's loads drain every store as well.  A row whose candidates find the pool
// full is marked kDropped in its count (the count itself stays exact), which sends it to
// the refilter.
constexpr uint32_t kPool = 256;
constexpr uint32_t kDropped = 0x80000000u;

__global__ __launch_bounds__(256) void bfFilterKernel(GemmArgs a) {
    __shared__ Chunk chunk[2];
    __shared__ int32_t qnLds[kGemmWaves][kRowsPerWave];
    __shared__ uint32_t cntLds[kGemmWaves][kRowsPerWave];   // exact, | kDropped
    __shared__ uint64_t firstLds[kGemmWaves][kRowsPerWave];
    __shared__ uint32_t capLds[kGemmWaves][kRowsPerWave];
    __shared__ int32_t thrLds[kGemmWaves][kRowsPerWave];
    __shared__ uint32_t poolAt[kGemmWaves][kPool];          // row << 16 | slot
    __shared__ uint64_t poolKey[kGemmWaves][kPool];
    __shared__ uint32_t poolN[kGemmWaves];
    const int lane = laneId(), wave = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4;
    uint32_t slice, qtile;
    if (!jobOf(a, &slice, &qtile)) return;
    const uint32_t q0 = qtile * kQueriesPerBlock + wave * kRowsPerWave;
    v4i A0[4], A1[4];
    loadQueries(a, q0, col, grp, A0, A1);
    {
        const bool live = q0 + lane < a.nsel;
        const size_t list = (size_t)(q0 + lane) * a.slices + slice;
        qnLds[wave][lane] = live ? a.qnorm[queryOf(a, q0 + lane)] : 0;
        cntLds[wave][lane] = 0;
        const uint64_t first = !live ? 0 : a.offs ? a.offs[list] : (uint64_t)list * a.capS;
        firstLds[wave][lane] = first;
        // slots are 16-bit in the pool; longer lists are used up to 65,535
        const uint64_t cap = !live ? 0 : a.offs ? a.offs[list + 1] - first : a.capS;
        capLds[wave][lane] = (uint32_t)min(cap, (uint64_t)0xFFFF);
        thrLds[wave][lane] = live ? a.thr[queryOf(a, q0 + lane)] : INT32_MAX;
        if (lane == 0) poolN[wave] = 0;
    }
    __builtin_amdgcn_wave_barrier();
    // this lane's rows: e = 4 rb + j holds row 16 rb + 4 grp + j of the wavefront
    int32_t thr[16];
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int j = 0; j < 4; j++) thr[rb * 4 + j] = thrLds[wave][16 * rb + 4 * grp + j];
    const uint32_t s0 = (uint32_t)((uint64_t)a.scanBlocks * slice / a.slices);
    const uint32_t s1 = (uint32_t)((uint64_t)a.scanBlocks * (slice + 1) / a.slices);
    scanChunks(a, chunk, s0, s1, [&](const BBlock& B, uint32_t sb) {
        v4i acc[4];
        blockProducts(A0, A1, B, acc);
        uint64_t pm[16];  // per element, the lanes with a candidate
        uint64_t any = 0;
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                pm[rb * 4 + j] = __ballot(acc[rb][j] > thr[rb * 4 + j]);
                any |= pm[rb * 4 + j];
            }
        if (!any || !B.valid()) return;
        // rare: each lane pools its candidates; list slots and pool places by LDS atomics
        const uint32_t p = baseBlockOf(a, sb) * 16 + col;
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (!((pm[rb * 4 + j] >> lane) & 1)) continue;
                const uint32_t row = 16 * rb + 4 * grp + j;
                const uint32_t slot = atomicAdd(&cntLds[wave][row], 1u) & ~kDropped;
                if (slot >= capLds[wave][row]) continue;
                const uint32_t at = atomicAdd(&poolN[wave], 1u);
                if (at < kPool) {
                    poolAt[wave][at] = row << 16 | slot;
                    poolKey[wave][at] = ((uint64_t)(uint32_t)(qnLds[wave][row] + B.odd() - 2 * acc[rb][j]) << 32) | p;
                } else {
                    atomicOr(&cntLds[wave][row], kDropped);
                }
            }
    }, [&](bool last) {
        // between chunks: the pooled candidates to their lists when half the pool is used
        __builtin_amdgcn_wave_barrier();
        const uint32_t pooled = min(kPool, poolN[wave]);
        if (!last && pooled < kPool / 2) return;
        for (uint32_t x = lane; x < pooled; x += 64) {
            const uint32_t w = poolAt[wave][x];
            a.cand[firstLds[wave][w >> 16] + (w & 0xFFFF)] = poolKey[wave][x];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) poolN[wave] = 0;
        __builtin_amdgcn_wave_barrier();
    });
    __builtin_amdgcn_wave_barrier();
    if (q0 + lane < a.nsel) a.counts[(size_t)(q0 + lane) * a.slices + slice] = cntLds[wave][lane];
}

// ---- pass 3: per row of the pass, sort its candidates and keep the first k ----
__global__ __launch_bounds__(kSelectThreads) void bfSelectKernel(const uint64_t* cand,
                                                                const uint32_t* counts,
                                                                const uint64_t* offs,
                                                                const uint32_t* qlist,
                                                                uint32_t slices, uint32_t capS,
                                                                uint32_t k, uint32_t* outIds,
                                                                int32_t* outDists,
                                                                uint32_t* retry) {
    __shared__ uint64_t keys[kSelectCap];
    __shared__ uint32_t segStart[kMaxSlices + 1];
    __shared__ uint32_t total, bad;
    const uint32_t i = blockIdx.x;
    const uint32_t q = qlist ? qlist[i] : i;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        uint32_t s = 0, over = 0;
        for (uint32_t j = 0; j < slices; j++) {
            const size_t list = (size_t)i * slices + j;
            const uint32_t cap = offs ? (uint32_t)(offs[list + 1] - offs[list]) : capS;
            const uint32_t c = counts[list] & 0x7FFFFFFFu;
            segStart[j] = s;
            over |= c > cap || (counts[list] >> 31);
            s += min(c, cap);
        }
        segStart[slices] = s;
        total = s;
        bad = over || s < k || s > kSelectCap;
        retry[i] = bad;
    }
    __syncthreads();
    if (bad) return;
    uint32_t P = kSelectThreads;
    while (P < total) P <<= 1;
    for (uint32_t x = t; x < P; x += kSelectThreads) {
        uint64_t key = ~0ull;
        if (x < total) {
            uint32_t s = 0;
            while (segStart[s + 1] <= x) s++;
            const size_t list = (size_t)i * slices + s;
            const uint64_t first = offs ? offs[list] : (uint64_t)list * capS;
            key = cand[first + (x - segStart[s])];
        }
        keys[x] = key;
    }
    __syncthreads();
    // bitonic sort of P keys, ascending
    for (uint32_t size = 2; size <= P; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t x = t; x < P / 2; x += kSelectThreads) {
                const uint32_t lo = 2 * x - (x & (stride - 1));
                const uint32_t hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const uint64_t u = keys[lo], v = keys[hi];
                if ((u > v) == asc) {
                    keys[lo] = v;
                    keys[hi] = u;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t j = t; j < k; j += kSelectThreads) {
        const uint64_t key = keys[j];
        outIds[(size_t)q * k + j] = (uint32_t)key;
        outDists[(size_t)q * k + j] = (int32_t)(key >> 32);
    }
}

// ---- one-pass kernel: 16 queries per wavefront, their best k kept in LDS ----
constexpr int kBufCap = 128;  // candidate buffer per query
constexpr int kTileCols = 32; // base rows per iteration (two 16-column MFMA tiles)
constexpr int kQ = 16;        // queries per wavefront

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

// Queries qlist[16 blockIdx.x + r] for r < 16 (below nsel), or 16 blockIdx.x + r when qlist
// is null: every base row streams past the wavefront; rows below a query's k-th key go to
// its buffer, merged into its best-k list whenever the next tile could overflow it.
__global__ __launch_bounds__(64) void bfKnnExactKernel(const int8_t* base, const int32_t* bnorm,
                                                      uint32_t n, const int8_t* queries,
                                                      const int32_t* qnorm, const uint32_t* qlist,
                                                      uint32_t nsel, uint32_t k, bool normsInRow,
                                                      uint32_t* outIds, int32_t* outDists) {
    __shared__ BfLds S;
    const int lane = laneId();
    const int col = lane & 15, grp = lane >> 4;
    const uint32_t sel0 = blockIdx.x * kQ;
    auto queryOf = [&](uint32_t r) -> uint32_t {
        const uint32_t i = sel0 + r;
        return i >= nsel ? 0xFFFFFFFFu : qlist ? qlist[i] : i;
    };

    if (lane < kQ) {
        S.ntop[lane] = 0;
        S.nbuf[lane] = 0;
        S.thr[lane] = ~0ull;
    }
    // A operand: the wavefront's query `col`, bytes [16*grp, +16) and [64 + 16*grp, +16)
    v4i a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
    const uint32_t qc = queryOf(col);
    if (qc != 0xFFFFFFFFu) {
        const int8_t* qr = queries + (size_t)qc * kVecBytes;
        a0 = *reinterpret_cast<const v4i*>(qr + 16 * grp);
        a1 = *reinterpret_cast<const v4i*>(qr + 64 + 16 * grp);
        // bytes 124..127 hold norms: zero the query's so q.p ignores them
        if (normsInRow && grp == 3) a1.w = 0;
    }
    // output rows held by this lane: 4*grp + r
    int qn[4];
    bool live[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t q = queryOf(4 * grp + r);
        live[r] = q != 0xFFFFFFFFu;
        qn[r] = live[r] ? qnorm[q] : 0;
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
                if (p < n && live[r] && key < thr[r]) {
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
        if (queryOf(r) != 0xFFFFFFFFu) mergeRow(S, r, k, lane);
    for (int r = 0; r < kQ; r++) {
        const uint32_t q = queryOf(r);
        if (q == 0xFFFFFFFFu) break;
        for (uint32_t j = lane; j < k; j += 64) {
            const bool have = j < S.ntop[r];
            const uint64_t key = have ? S.top[r][j] : ~0ull;
            outIds[(size_t)q * k + j] = have ? (uint32_t)key : kEmpty;
            outDists[(size_t)q * k + j] = have ? (int32_t)(key >> 32) : 0x7FFFFFFF;
        }
    }
}


template <int L>
hipError_t launchThreshold(const GemmArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((bfThresholdKernel<L>), dim3((a.jobs + 7) / 8 * 8), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Workgroups of `kernel` (256 threads) that all CUs keep resident at once.
template <class Kernel>
uint32_t residentBlocks(Kernel kernel) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) != hipSuccess || per < 1 || cus < 1)
        return 768;
    return (uint32_t)(per * cus);
}

// Slices of `blocks` for qTiles query tiles (at most maxSlices, at least minBlocks blocks
// per slice): the qTiles * slices workgroups should fill whole rounds of the `resident`
// ones, since a round that is only partly filled takes as long as a full one (805
// workgroups over 768 resident ran the 1M filter pass in two rounds).  The slice count with
// the largest busy fraction wins, the fewest slices on ties.
uint32_t pickSlices(uint32_t qTiles, uint32_t blocks, uint32_t resident, uint32_t maxSlices,
                    uint32_t minBlocks) {
    const uint32_t sMax = std::max(1u, std::min(maxSlices, blocks / std::max(1u, minBlocks)));
    uint32_t best = 1;
    double bestFill = -1;
    for (uint32_t sl = 1; sl <= sMax; sl++) {
        const uint64_t jobs = (uint64_t)qTiles * sl;
        const uint64_t rounds = (jobs + resident - 1) / resident;
        const double fill = (double)jobs / (double)(rounds * resident);
        if (fill > bestFill + 1e-9) {
            best = sl;
            bestFill = fill;
        }
    }
    return best;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
};

// The device memory of one call: one allocation cut into 256-byte aligned pieces, taken
// in the order they were added.  A call allocates once and frees once (about a millisecond
// each for the half-gigabyte candidate lists of 29,316 queries over 1M rows), instead of
// once per pass and buffer.
class Arena {
public:
    void add(size_t bytes) { total_ += pad(bytes); }
    hipError_t allocate() { return hipMalloc(&mem_.p, std::max<size_t>(total_, 256)); }
    template <class T>
    T* take(size_t count) {
        const size_t b = pad(count * sizeof(T));
        if (!mem_.p || used_ + b > total_) return nullptr;
        T* p = reinterpret_cast<T*>(static_cast<char*>(mem_.p) + used_);
        used_ += b;
        return p;
    }

private:
    static size_t pad(size_t b) { return (std::max<size_t>(b, 8) + 255) & ~size_t(255); }
    DevBuf mem_;
    size_t total_ = 0, used_ = 0;
};

// Pass 1 for the rows of `a` (all queries, or a.qlist): thr[query] from L subsets per lane
// column over every `stride`-th block of the base; accmax holds a.nsel * 16 * L ints.
hipError_t thresholds(GemmArgs a, int L, uint32_t stride, int32_t* accmax, int32_t* thr,
                      hipStream_t s) {
    const uint32_t sampleBlocks = std::max(1u, a.nBlocks / stride);
    const uint32_t subsets = 16u * (uint32_t)L;
    hipError_t e = hipMemsetD32Async((hipDeviceptr_t)accmax, (int)0x80000000, (size_t)a.nsel * subsets, s);
    if (e != hipSuccess) return e;
    a.scanBlocks = sampleBlocks;
    a.stride = stride;
    const uint32_t res = L == 1 ? residentBlocks(bfThresholdKernel<1>) : L == 2 ? residentBlocks(bfThresholdKernel<2>)
                       : L == 4 ? residentBlocks(bfThresholdKernel<4>) : residentBlocks(bfThresholdKernel<8>);
    a.slices = pickSlices(a.qTiles, sampleBlocks, res, 64, 512);
    a.jobs = a.qTiles * a.slices;
    a.accmax = accmax;
    switch (L) {
        case 1: e = launchThreshold<1>(a, s); break;
        case 2: e = launchThreshold<2>(a, s); break;
        case 4: e = launchThreshold<4>(a, s); break;
        default: e = launchThreshold<8>(a, s); break;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bfReduceThresholdKernel, dim3((a.nsel + 255) / 256), dim3(256), 0, s,
                       static_cast<const int32_t*>(accmax), (int)subsets, a.qlist, a.nsel, thr);
    return hipGetLastError();
}

// Passes 2 and 3 for the rows of `a`, with candidate slots `cand`, list counters `cnt`
// (a.nsel * a.slices) and flags (a.nsel); retry[i] is set for the rows the select pass
// could not finish.  With offs, list (i, s) holds exactly offs[i S + s + 1] - offs[i S + s]
// keys.
hipError_t filterSelect(GemmArgs a, const uint64_t* offs, uint64_t* cand, uint32_t* cnt,
                        uint32_t* flags, uint32_t k, uint32_t* outIds, int32_t* outDists,
                        std::vector<uint32_t>& counts, std::vector<uint32_t>& retry, hipStream_t s) {
    a.cand = cand;
    a.counts = cnt;
    a.offs = offs;
    a.jobs = a.qTiles * a.slices;
    hipLaunchKernelGGL(bfFilterKernel, dim3((a.jobs + 7) / 8 * 8), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bfSelectKernel, dim3(a.nsel), dim3(kSelectThreads), 0, s,
                       static_cast<const uint64_t*>(cand), static_cast<const uint32_t*>(cnt),
                       offs, a.qlist, a.slices, a.capS, k, outIds, outDists, flags);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    counts.resize((size_t)a.nsel * a.slices);
    retry.resize(a.nsel);
    e = hipMemcpyAsync(counts.data(), cnt, counts.size() * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(retry.data(), flags, retry.size() * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

}  // namespace

hipError_t launchBruteForceKnn(const int8_t* base, const int32_t* bnorm, uint32_t n,
                               const int8_t* queries, const int32_t* qnorm, uint32_t nq,
                               uint32_t k, bool normsInRow, uint32_t* outIds, int32_t* outDists,
                               hipStream_t s) {
    if (nq == 0) return hipSuccess;
    if (k == 0 || k > (uint32_t)kTopCap) return hipErrorInvalidValue;
    std::vector<uint32_t> exact;  // queries for the one-pass kernel
    uint32_t* exactDev = nullptr;  // room for nq ids
    auto runExact = [&]() -> hipError_t {
        if (exact.empty()) return hipSuccess;
        hipError_t e = hipMemcpyAsync(exactDev, exact.data(), exact.size() * 4, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        const uint32_t ns = (uint32_t)exact.size();
        hipLaunchKernelGGL(bfKnnExactKernel, dim3((ns + kQ - 1) / kQ), dim3(64), 0, s, base, bnorm,
                           n, queries, qnorm, static_cast<const uint32_t*>(exactDev), ns, k,
                           normsInRow, outIds, outDists);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        return hipStreamSynchronize(s);
    };
    if (n == 0) {
        Arena ar;
        ar.add((size_t)nq * 4);
        hipError_t e = ar.allocate();
        if (e != hipSuccess) return e;
        exactDev = ar.take<uint32_t>(nq);
        for (uint32_t q = 0; q < nq; q++) exact.push_back(q);
        return runExact();
    }
    // candidate slots per query; PANN_BFKNN_CANDIDATES lowers them (tests of the retry paths)
    uint32_t perQuery = kCandPerQuery;
    if (const char* env = getenv("PANN_BFKNN_CANDIDATES")) {
        const long v = strtol(env, nullptr, 10);
        if (v >= 16 && v < (long)kCandPerQuery) perQuery = (uint32_t)v;
    }
    GemmArgs a = {};
    a.base = base;
    a.bnorm = bnorm;
    a.n = n;
    a.queries = queries;
    a.qnorm = qnorm;
    a.nsel = nq;
    a.normsInRow = normsInRow ? 1 : 0;
    a.qTiles = (nq + kQueriesPerBlock - 1) / kQueriesPerBlock;
    a.nBlocks = (n + 15) / 16;
    // the sample: up to kSmallBase rows the whole base with 16 L >= k subsets (at least k
    // points pass); above it every fourth block with one subset per lane column
    const int Lk = k <= 16 ? 1 : k <= 32 ? 2 : k <= 64 ? 4 : 8;
    const bool small = n <= kSmallBase;
    const uint32_t sampleStride = small ? 1 : 4;
    const uint32_t resFilter = residentBlocks(bfFilterKernel);
    const uint32_t mainSlices = pickSlices(a.qTiles, a.nBlocks, resFilter, 32, 512);
    const uint32_t mainCap = std::max(16u, perQuery / mainSlices);
    const uint64_t mainSlots = (uint64_t)nq * mainSlices * mainCap;

    // every piece sized for all nq queries: thresholds, the maxima of up to 16 * 8 subsets,
    // candidate slots, list counters and exact offsets for up to kMaxSlices slices, flags
    // and three query lists (looser threshold, refiltered, one-pass)
    Arena ar;
    ar.add((size_t)nq * 4);
    ar.add((size_t)nq * 128 * 4);
    ar.add(mainSlots * 8);
    ar.add((size_t)nq * kMaxSlices * 4);
    ar.add(((size_t)nq * kMaxSlices + 1) * 8);
    for (int i = 0; i < 4; i++) ar.add((size_t)nq * 4);
    hipError_t e = ar.allocate();
    if (e != hipSuccess) return e;
    int32_t* thr = ar.take<int32_t>(nq);
    int32_t* accmax = ar.take<int32_t>((size_t)nq * 128);
    uint64_t* cand = ar.take<uint64_t>(mainSlots);
    uint32_t* cnt = ar.take<uint32_t>((size_t)nq * kMaxSlices);
    uint64_t* offsDev = ar.take<uint64_t>((size_t)nq * kMaxSlices + 1);
    uint32_t* flags = ar.take<uint32_t>(nq);
    uint32_t* looseDev = ar.take<uint32_t>(nq);
    uint32_t* againDev = ar.take<uint32_t>(nq);
    exactDev = ar.take<uint32_t>(nq);
    if (!exactDev) return hipErrorOutOfMemory;

    e = thresholds(a, small ? Lk : 1, sampleStride, accmax, thr, s);
    if (e != hipSuccess) return e;
    a.thr = thr;
    a.scanBlocks = a.nBlocks;
    a.stride = 1;
    a.slices = mainSlices;
    a.capS = mainCap;
    std::vector<uint32_t> counts, retry;
    e = filterSelect(a, nullptr, cand, cnt, flags, k, outIds, outDists, counts, retry, s);
    if (e != hipSuccess) return e;

    // Queries the select pass could not finish.  Those with fewer than k candidates get a
    // threshold from 16 * 8 subsets of the sample (at least 128 points pass); then they and
    // those whose lists overflowed are filtered again over many slices, first counting
    // each (query, slice) list exactly, then writing them at exact offsets.  A query with
    // fewer than k or more than kSelectCap candidates then goes to the one-pass kernel.
    std::vector<uint32_t> again, loose;
    for (uint32_t q = 0; q < nq; q++) {
        if (!retry[q]) continue;
        uint64_t tot = 0;
        for (uint32_t j = 0; j < a.slices; j++) tot += counts[(size_t)q * a.slices + j] & 0x7FFFFFFFu;
        again.push_back(q);
        if (tot < k) loose.push_back(q);
    }
    if (!loose.empty()) {
        e = hipMemcpyAsync(looseDev, loose.data(), loose.size() * 4, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        GemmArgs r = a;
        r.qlist = looseDev;
        r.nsel = (uint32_t)loose.size();
        r.qTiles = (r.nsel + kQueriesPerBlock - 1) / kQueriesPerBlock;
        e = thresholds(r, 8, sampleStride, accmax, thr, s);
        if (e != hipSuccess) return e;
    }
    if (!again.empty()) {
        e = hipMemcpyAsync(againDev, again.data(), again.size() * 4, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        GemmArgs r = a;
        r.qlist = againDev;
        r.nsel = (uint32_t)again.size();
        r.qTiles = (r.nsel + kQueriesPerBlock - 1) / kQueriesPerBlock;
        r.scanBlocks = r.nBlocks;
        r.stride = 1;
        r.slices = pickSlices(r.qTiles, r.nBlocks, resFilter, kMaxSlices, 64);
        r.capS = 0;  // count only
        std::vector<uint32_t> c1, r1;
        e = filterSelect(r, nullptr, cand, cnt, flags, k, outIds, outDists, c1, r1, s);
        if (e != hipSuccess) return e;
        std::vector<uint64_t> offs(1, 0);
        for (uint32_t x = 0; x < r.nsel; x++) {
            uint64_t tot = 0;
            for (uint32_t j = 0; j < r.slices; j++) tot += c1[(size_t)x * r.slices + j] & 0x7FFFFFFFu;
            const bool ok = tot >= k && tot <= kSelectCap;
            if (!ok) exact.push_back(again[x]);
            for (uint32_t j = 0; j < r.slices; j++)
                offs.push_back(offs.back() + (ok ? (c1[(size_t)x * r.slices + j] & 0x7FFFFFFFu) : 0));
        }
        e = hipMemcpyAsync(offsDev, offs.data(), offs.size() * 8, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        // the main pass's lists are done with; more keys than they held get their own slots
        DevBuf more;
        uint64_t* cand2 = cand;
        if (offs.back() > mainSlots) {
            if ((e = hipMalloc(&more.p, offs.back() * 8)) != hipSuccess) return e;
            cand2 = static_cast<uint64_t*>(more.p);
        }
        std::vector<uint32_t> c2, r2;
        e = filterSelect(r, offsDev, cand2, cnt, flags, k, outIds, outDists, c2, r2, s);
        if (e != hipSuccess) return e;
        for (uint32_t x = 0; x < r.nsel; x++) {
            const bool wasOk = offs[(size_t)(x + 1) * r.slices] > offs[(size_t)x * r.slices];
            if (r2[x] && wasOk) exact.push_back(again[x]);
        }
    }
    if (getenv("PANN_BFKNN_TRACE"))
        fprintf(stderr, "[bf_knn] %u queries, %u slices: %zu refiltered, %zu of them with a looser "
                "threshold, %zu one-pass\n", nq, a.slices, again.size(), loose.size(), exact.size());
    return runExact();
}

}  // namespace pann
