// build_gpu_trees.hip -- HCNNG and PyNNDescent graph construction on gfx950.
//
// The algorithm is the host builder's (build_hcnng.cpp over the cluster trees of
// build_trees.h), which follows ParlayANN's HCNNG and the reference's plan for its GPU
// build (README.md:351-353: clustering trees with leaves of at most 1,000 points, dense
// nearest-neighbour work per leaf, a degree-bounded spanning tree per leaf, thousands of
// leaves in parallel, edges united at the end).  Per tree:
//   1. treeSplitKernel, one workgroup per segment still larger than a leaf, level by
//      level: two pivots drawn from the segment with the host's seeded generator, every
//      point sent to the nearer (ties to the first), a stable partition (near side, then
//      far side, each in input order) by block-wide ballot scans; a degenerate split
//      halves the segment, as on the host;
//   2. hcnngLeafKernel, one workgroup per leaf with the leaf's rows in LDS (up to 1,024
//      rows of 128 bytes): each point's 16 nearest leaf mates by (distance, leaf index)
//      from v_dot4 products, the resulting edges bitonic-sorted by (distance, endpoints)
//      in LDS, then Kruskal with the degree bound on one wavefront -- edges already
//      inside one component or at a full endpoint are rejected 64 at a time against the
//      state at the start of their batch (both only become more true), the rest are
//      taken in order by one lane.
// A node's list is the concatenation over trees of the edges its leaf accepted, in
// acceptance order, which is the host builder's list, entry for entry.
#include "build_prune.h"
#include "devmem.h"
#include "device_common.h"
#include "pann_internal.h"
#include "tools_common.h"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <numeric>
#include <vector>

namespace pann {
namespace {

using namespace dev;

constexpr int kSplitThreads = 256;
constexpr int kLeafThreads = 1024;
constexpr uint32_t kLeafMax = 1024;   // points of a leaf held in LDS
constexpr uint32_t kLeafNbrs = 16;    // nearest leaf mates per point (as on the host)
constexpr uint32_t kMstDegMax = 4;

// The host builders' generator (tools::Rng): the SplitMix64 stream.
struct DevRng {
    uint64_t s;
    __device__ uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    __device__ uint32_t below(uint32_t m) { return (uint32_t)((next() >> 11) % m); }
};

// Row i as eight 16-byte slices; the norm bytes 124..127 read as zero when rows carry it
// (so products and norms are over the row's dimensions only, as the host's distI8).
__device__ __forceinline__ void loadRow(const int8_t* vec, uint32_t i, bool normsInRow, int4 (&r)[8]) {
    const int4* p = reinterpret_cast<const int4*>(vec + (size_t)i * kVecBytes);
#pragma unroll
    for (int s = 0; s < 8; s++) r[s] = p[s];
    if (normsInRow) r[7].w = 0;
}

__device__ __forceinline__ int rowDot(const int4 (&a)[8], const int4 (&b)[8]) {
    int s = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        s = sdot4(a[k].x, b[k].x, s);
        s = sdot4(a[k].y, b[k].y, s);
        s = sdot4(a[k].z, b[k].z, s);
        s = sdot4(a[k].w, b[k].w, s);
    }
    return s;
}

struct SplitSeg {
    uint32_t b, e;   // the segment ids[b..e)
    uint64_t seed;   // tools::mixSeed(tree seed, round << 32 ^ b)
};

// One workgroup per segment: ids[b..e) partitioned around two pivots, cut point in cuts.
__global__ __launch_bounds__(kSplitThreads) void treeSplitKernel(const int8_t* vec, int normsInRow,
                                                                uint32_t* ids, uint32_t* tmp,
                                                                uint8_t* flags, const SplitSeg* segs,
                                                                uint32_t* cuts) {
    __shared__ uint32_t piv[2];
    __shared__ uint32_t nearTotal;
    __shared__ uint32_t waveTot[kSplitThreads / 64];
    const SplitSeg sg = segs[blockIdx.x];
    const uint32_t b = sg.b, e = sg.e, m = e - b;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) {
        DevRng r{sg.seed};
        const uint32_t pa = ids[b + r.below(m)];
        uint32_t pb = ids[b + r.below(m)];
        for (int tries = 0; pb == pa && tries < 8; tries++) pb = ids[b + r.below(m)];
        piv[0] = pa;
        piv[1] = pb;
        nearTotal = 0;
    }
    __syncthreads();
    int4 A[8], B[8];
    loadRow(vec, piv[0], normsInRow, A);
    loadRow(vec, piv[1], normsInRow, B);
    const int na = rowDot(A, A), nb = rowDot(B, B);

    // pass 1: which side each point goes to
    uint32_t cnt = 0;
    for (uint32_t p = b + tid; p < e; p += kSplitThreads) {
        int4 X[8];
        loadRow(vec, ids[p], normsInRow, X);
        const int nx = rowDot(X, X);
        const int da = nx + na - 2 * rowDot(X, A);
        const int db = nx + nb - 2 * rowDot(X, B);
        const uint8_t f = da <= db ? 1 : 0;
        flags[p] = f;
        cnt += f;
    }
    atomicAdd(&nearTotal, cnt);
    __syncthreads();
    const uint32_t T = nearTotal;

    // pass 2: near side first, each side in input order
    const uint64_t lanesBelow = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t before = 0;  // near points before this chunk
    for (uint32_t cs = b; cs < e; cs += kSplitThreads) {
        const uint32_t p = cs + tid;
        const bool in = p < e;
        const bool f = in && flags[p];
        const uint64_t bal = ballot(f);
        if (lane == 0) waveTot[w] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < kSplitThreads / 64; k++) {
            pre += k < w ? waveTot[k] : 0u;
            tot += waveTot[k];
        }
        __syncthreads();
        if (in) {
            const uint32_t nbf = before + pre + (uint32_t)__popcll(bal & lanesBelow);
            tmp[f ? b + nbf : b + T + (p - b - nbf)] = ids[p];
        }
        before += tot;
    }
    __threadfence();
    __syncthreads();
    for (uint32_t p = b + tid; p < e; p += kSplitThreads)
        ids[p] = __hip_atomic_load(tmp + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (tid == 0) {
        uint32_t cut = b + T;
        if (cut == b || cut == e) cut = b + m / 2;
        cuts[blockIdx.x] = cut;
    }
}

struct LeafLds {
    union {
        int4 rows[kLeafMax][8];                  // the leaf's rows
        uint64_t edges[kLeafMax * kLeafNbrs];    // then its candidate edges
    };
    union {
        int32_t norms[kLeafMax];
        struct {
            uint16_t parent[kLeafMax];
            uint8_t degree[kLeafMax];
        } dsu;
    };
    uint16_t list[kLeafMax][kMstDegMax];  // accepted edges per point, in acceptance order
    uint8_t cnt[kLeafMax];
};

__device__ __forceinline__ uint32_t findRoot(const uint16_t* parent, uint32_t x) {
    while (parent[x] != x) x = parent[x];
    return x;
}

__device__ __forceinline__ uint32_t findHalving(uint16_t* parent, uint32_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

// One workgroup per leaf (at least 2 points).  outAdj: n x rowStride, this tree's
// entries at column col0; outCnt: n x trees bytes.
__global__ __launch_bounds__(kLeafThreads) void hcnngLeafKernel(const int8_t* vec, int normsInRow,
                                                               const uint32_t* ids, const uint32_t* leafB,
                                                               const uint32_t* leafE, uint32_t mstDeg,
                                                               uint32_t* outAdj, uint32_t rowStride,
                                                               uint32_t col0, uint8_t* outCnt,
                                                               uint32_t trees, uint32_t tree) {
    __shared__ LeafLds S;
    const uint32_t b = leafB[blockIdx.x], m = leafE[blockIdx.x] - b;
    const uint32_t* L = ids + b;
    const uint32_t tid = threadIdx.x;

    for (uint32_t x = tid; x < m * 8; x += kLeafThreads) {
        const uint32_t i = x >> 3, s = x & 7;
        int4 v = *reinterpret_cast<const int4*>(vec + (size_t)L[i] * kVecBytes + s * 16);
        if (normsInRow && s == 7) v.w = 0;
        S.rows[i][s] = v;
    }
    __syncthreads();
    for (uint32_t i = tid; i < m; i += kLeafThreads) {
        int4 X[8];
#pragma unroll
        for (int s = 0; s < 8; s++) X[s] = S.rows[i][s];
        S.norms[i] = rowDot(X, X);
    }
    __syncthreads();

    // each point's nearest leaf mates, (distance, leaf index) keys kept sorted
    const uint32_t K = min(kLeafNbrs, m - 1);
    uint64_t top[kLeafNbrs];
#pragma unroll
    for (int t = 0; t < (int)kLeafNbrs; t++) top[t] = ~0ull;
    const uint32_t i = tid;
    if (i < m) {
        int4 X[8];
#pragma unroll
        for (int s = 0; s < 8; s++) X[s] = S.rows[i][s];
        const int ni = S.norms[i];
        for (uint32_t j = 0; j < m; j++) {
            int4 Y[8];
#pragma unroll
            for (int s = 0; s < 8; s++) Y[s] = S.rows[j][s];
            const int d = ni + S.norms[j] - 2 * rowDot(X, Y);
            uint64_t key = j == i ? ~0ull : (((uint64_t)(uint32_t)d << 32) | j);
            if (key < top[kLeafNbrs - 1]) {
#pragma unroll
                for (int t = 0; t < (int)kLeafNbrs; t++) {
                    const uint64_t lo = key < top[t] ? key : top[t];
                    key = key < top[t] ? top[t] : key;
                    top[t] = lo;
                }
            }
        }
    }
    __syncthreads();  // rows and norms are dead from here

    // edges (distance, smaller index, larger index) as 43-bit keys, then sorted
    const uint32_t E = m * K;
    uint32_t P = 2;
    while (P < E) P <<= 1;
    if (i < m) {
        for (uint32_t t = 0; t < K; t++) {
            const uint32_t j = (uint32_t)top[t], d = (uint32_t)(top[t] >> 32);
            const uint32_t lo = min(i, j), hi = max(i, j);
            S.edges[i * K + t] = ((uint64_t)d << 20) | ((uint64_t)lo << 10) | hi;
        }
        S.dsu.parent[i] = (uint16_t)i;
        S.dsu.degree[i] = 0;
        S.cnt[i] = 0;
    }
    for (uint32_t x = E + tid; x < P; x += kLeafThreads) S.edges[x] = ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t x = tid; x < P; x += kLeafThreads) {
                const uint32_t y = x ^ j;
                if (y > x) {
                    const uint64_t a = S.edges[x], c = S.edges[y];
                    if ((a > c) == ((x & k) == 0)) {
                        S.edges[x] = c;
                        S.edges[y] = a;
                    }
                }
            }
            __syncthreads();
        }

    // Kruskal with the degree bound, in (distance, endpoints) order; repeated edges
    // (equal keys, adjacent after the sort) count once
    if (tid < 64) {
        const int lane = (int)tid;
        uint32_t accepted = 0;
        for (uint32_t base = 0; base < E && accepted + 1 < m; base += 64) {
            const uint32_t idx = base + lane;
            const uint64_t key = idx < E ? S.edges[idx] : ~0ull;
            const bool repeat = idx < E && idx > 0 && S.edges[idx - 1] == key;
            const uint32_t a = (uint32_t)(key >> 10) & 1023u, c = (uint32_t)key & 1023u;
            bool live = idx < E && !repeat;
            if (live) live = S.dsu.degree[a] < mstDeg && S.dsu.degree[c] < mstDeg &&
                             findRoot(S.dsu.parent, a) != findRoot(S.dsu.parent, c);
            uint64_t sv = ballot(live);
            while (sv) {
                const int l = __ffsll((unsigned long long)sv) - 1;
                sv &= sv - 1;
                const uint32_t aa = (uint32_t)readlane((int)a, l), cc = (uint32_t)readlane((int)c, l);
                if (lane == 0 && S.dsu.degree[aa] < mstDeg && S.dsu.degree[cc] < mstDeg) {
                    const uint32_t ra = findHalving(S.dsu.parent, aa), rb = findHalving(S.dsu.parent, cc);
                    if (ra != rb) {
                        S.dsu.parent[ra] = (uint16_t)rb;
                        S.dsu.degree[aa]++;
                        S.dsu.degree[cc]++;
                        S.list[aa][S.cnt[aa]++] = (uint16_t)cc;
                        S.list[cc][S.cnt[cc]++] = (uint16_t)aa;
                        accepted++;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
            accepted = (uint32_t)readlane((int)accepted, 0);
        }
    }
    __syncthreads();
    if (i < m) {
        const uint32_t g = L[i], c = S.cnt[i];
        outCnt[(size_t)g * trees + tree] = (uint8_t)c;
        for (uint32_t s = 0; s < c; s++) outAdj[(size_t)g * rowStride + col0 + s] = L[S.list[i][s]];
    }
}

static_assert(sizeof(LeafLds) <= 160 * 1024, "leaf LDS");

// ---- PyNNDescent ----

constexpr int kPynnThreads = 512;
constexpr uint32_t kPynnKMax = 64;        // K of the K-NN lists, at most
constexpr uint32_t kJoinRows = 160;       // |new| + |reverse new| + |old| + |reverse old| <= 2.5 K
static_assert(kJoinRows >= kPynnKMax * 5 / 2, "join rows");
constexpr int kJoinWaves = 2;
constexpr uint32_t kBridgeCapDev = 24;    // bridges per node at most (the host's kBridgeCap)

struct PynnLeafLds {
    int4 rows[kLeafMax][8];
    int32_t norms[kLeafMax];
    uint32_t ids[kLeafMax];
};

// Each point of a leaf: its KMAX nearest leaf mates by (distance, id), as sorted keys
// d << 32 | id; the first K go to outKeys (n x K), ~0 where the leaf has fewer mates.
template <uint32_t KMAX>
__global__ __launch_bounds__(kPynnThreads) void pynnLeafKernel(const int8_t* vec, int normsInRow,
                                                              const uint32_t* ids, const uint32_t* leafB,
                                                              const uint32_t* leafE, uint32_t K,
                                                              uint64_t* outKeys) {
    __shared__ PynnLeafLds S;
    const uint32_t b = leafB[blockIdx.x], m = leafE[blockIdx.x] - b;
    const uint32_t tid = threadIdx.x;
    for (uint32_t x = tid; x < m * 8; x += kPynnThreads) {
        const uint32_t i = x >> 3, s = x & 7;
        const uint32_t g = ids[b + i];
        int4 v = *reinterpret_cast<const int4*>(vec + (size_t)g * kVecBytes + s * 16);
        if (normsInRow && s == 7) v.w = 0;
        S.rows[i][s] = v;
        if (s == 0) S.ids[i] = g;
    }
    __syncthreads();
    for (uint32_t i = tid; i < m; i += kPynnThreads) {
        int4 X[8];
#pragma unroll
        for (int s = 0; s < 8; s++) X[s] = S.rows[i][s];
        S.norms[i] = rowDot(X, X);
    }
    __syncthreads();
    for (uint32_t i = tid; i < m; i += kPynnThreads) {
        uint64_t top[KMAX];
#pragma unroll
        for (int t = 0; t < (int)KMAX; t++) top[t] = ~0ull;
        int4 X[8];
#pragma unroll
        for (int s = 0; s < 8; s++) X[s] = S.rows[i][s];
        const int ni = S.norms[i];
        for (uint32_t j = 0; j < m; j++) {
            int4 Y[8];
#pragma unroll
            for (int s = 0; s < 8; s++) Y[s] = S.rows[j][s];
            const int d = ni + S.norms[j] - 2 * rowDot(X, Y);
            uint64_t key = j == i ? ~0ull : (((uint64_t)(uint32_t)d << 32) | S.ids[j]);
            if (key < top[KMAX - 1]) {
#pragma unroll
                for (int t = 0; t < (int)KMAX; t++) {
                    const uint64_t lo = key < top[t] ? key : top[t];
                    key = key < top[t] ? top[t] : key;
                    top[t] = lo;
                }
            }
        }
        uint64_t* o = outKeys + (size_t)S.ids[i] * K;
#pragma unroll
        for (int t = 0; t < (int)KMAX; t++)
            if ((uint32_t)t < K) o[t] = top[t];
    }
}

// dst := the K smallest distinct keys of src and add (each n x K sorted keys, ~0
// unused); every valid entry is new (the state after the cluster trees).
__global__ void pynnMergeInitKernel(const uint64_t* src, const uint64_t* add, uint64_t* dst,
                                    uint8_t* dstNew, uint32_t n, uint32_t K) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const uint64_t* A = src + (size_t)x * K;
    const uint64_t* B = add + (size_t)x * K;
    uint64_t* D = dst + (size_t)x * K;
    uint8_t* DF = dstNew + (size_t)x * K;
    uint32_t ia = 0, ib = 0, c = 0;
    while (c < K) {
        const uint64_t a = ia < K ? A[ia] : ~0ull, bb = ib < K ? B[ib] : ~0ull;
        const uint64_t v = a < bb ? a : bb;
        if (v == ~0ull) break;
        if (a <= bb) ia++;
        if (bb <= a) ib++;
        D[c] = v;
        DF[c] = 1;
        c++;
    }
    for (uint32_t t = c; t < K; t++) {
        D[t] = ~0ull;
        DF[t] = 0;
    }
}

// NN-descent sampling of node i: the first S new entries (their flags cleared) and every
// old entry, with (target << 32 | i) pairs for the reverse lists.
__global__ void pynnSampleKernel(const uint64_t* lists, uint8_t* isNew, uint32_t n, uint32_t K,
                                 uint32_t S, uint32_t* nw, uint32_t* nwN, uint32_t* old,
                                 uint32_t* oldN, uint64_t* pairsNw, uint64_t* pairsOld) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t cn = 0, co = 0;
    for (uint32_t j = 0; j < K; j++) {
        const uint64_t key = lists[(size_t)i * K + j];
        const uint32_t id = (uint32_t)key;
        if (id == kEmpty) continue;
        uint8_t& f = isNew[(size_t)i * K + j];
        if (f && cn < S) {
            nw[(size_t)i * S + cn] = id;
            pairsNw[(size_t)i * S + cn] = ((uint64_t)id << 32) | i;
            cn++;
            f = 0;
        } else if (!f) {
            old[(size_t)i * K + co] = id;
            pairsOld[(size_t)i * K + co] = ((uint64_t)id << 32) | i;
            co++;
        }
    }
    for (uint32_t t = cn; t < S; t++) pairsNw[(size_t)i * S + t] = ~0ull;
    for (uint32_t t = co; t < K; t++) pairsOld[(size_t)i * K + t] = ~0ull;
    nwN[i] = cn;
    oldN[i] = co;
}

// Sorted (target << 32 | source) pairs: first[target] = index of its run's start.
__global__ void pynnRunHeadsKernel(const uint64_t* sorted, size_t count, uint32_t* first) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const uint64_t k = sorted[p];
    if (k == ~0ull) return;
    if (p == 0 || (sorted[p - 1] >> 32) != (k >> 32)) first[k >> 32] = (uint32_t)p;
}

// rev[x][0..S) = the first S sources of x's run (ascending), revN[x] = their count.
__global__ void pynnRunFillKernel(const uint64_t* sorted, size_t count, const uint32_t* first,
                                  uint32_t S, uint32_t* rev, uint32_t* revN) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const uint64_t k = sorted[p];
    if (k == ~0ull) return;
    const uint32_t x = (uint32_t)(k >> 32);
    const uint32_t r = (uint32_t)(p - first[x]);
    if (r < S) {
        rev[(size_t)x * S + r] = (uint32_t)k;
        atomicMax(&revN[x], r + 1);
    }
}

struct JoinArgs {
    const int8_t* vec;
    int32_t normsInRow;
    uint32_t K, S;
    const uint64_t* before;   // n x K keys at the start of the iteration
    const uint32_t *nw, *nwN, *rnw, *rnwN, *old, *oldN, *rold, *roldN;
    uint32_t u0, u1;          // nodes of this pass
    uint32_t* counts;         // count pass: proposals per node
    const uint64_t* offsets;  // write pass: first slot per node (relative to u0's)
    uint64_t* propKey;        // d << 32 | id
    uint32_t* propTgt;
};

struct JoinLds {
    uint32_t ids[kJoinRows];
    int32_t norms[kJoinRows];
    int4 rows[kJoinRows][8];
};

// The local join of node u (one wavefront): A = new ++ reverse new, B = old ++ reverse
// old; every pair (A[a], A[b>a]) and (A[a], B[*]) of distinct ids gives two proposals,
// kept when better than the target's worst entry at the start of the iteration (a
// proposal no better can never enter the list).  WRITE: stored at the node's offset.
template <bool WRITE>
__global__ __launch_bounds__(64 * kJoinWaves) void pynnJoinKernel(JoinArgs a) {
    __shared__ JoinLds lds[kJoinWaves];
    const int lane = laneId(), w = threadIdx.x >> 6;
    JoinLds& L = lds[w];
    const uint32_t u = a.u0 + blockIdx.x * kJoinWaves + w;
    if (u >= a.u1) return;
    const uint32_t n1 = a.nwN[u], n2 = a.rnwN[u], o1 = a.oldN[u], o2 = a.roldN[u];
    const uint32_t nA = n1 + n2, nB = o1 + o2, nT = nA + nB;
    for (uint32_t t = lane; t < nT; t += 64) {
        uint32_t id;
        if (t < n1) id = a.nw[(size_t)u * a.S + t];
        else if (t < nA) id = a.rnw[(size_t)u * a.S + (t - n1)];
        else if (t < nA + o1) id = a.old[(size_t)u * a.K + (t - nA)];
        else id = a.rold[(size_t)u * a.S + (t - nA - o1)];
        L.ids[t] = id;
    }
    __builtin_amdgcn_wave_barrier();
    for (uint32_t x = lane; x < nT * 8; x += 64) {
        const uint32_t t = x >> 3, s = x & 7;
        int4 v = *reinterpret_cast<const int4*>(a.vec + (size_t)L.ids[t] * kVecBytes + s * 16);
        if (a.normsInRow && s == 7) v.w = 0;
        L.rows[t][s] = v;
    }
    __builtin_amdgcn_wave_barrier();
    for (uint32_t t = lane; t < nT; t += 64) {
        int4 X[8];
#pragma unroll
        for (int s = 0; s < 8; s++) X[s] = L.rows[t][s];
        L.norms[t] = rowDot(X, X);
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t cnt = 0;
    uint64_t pos = WRITE ? a.offsets[u - a.u0] : 0;
    for (uint32_t ai = 0; ai < nA; ai++) {
        const uint32_t x = L.ids[ai];
        int4 X[8];
#pragma unroll
        for (int s = 0; s < 8; s++) X[s] = L.rows[ai][s];
        const int nx = L.norms[ai];
        const uint64_t wx = a.before[(size_t)x * a.K + a.K - 1];
        // partners: A[ai+1..nA) then B
        const uint32_t np = nT - ai - 1;
        for (uint32_t c0 = 0; c0 < np; c0 += 64) {
            const uint32_t t = ai + 1 + c0 + lane;
            bool p1 = false, p2 = false;
            uint64_t k1 = 0, k2 = 0;
            uint32_t y = 0;
            if (c0 + lane < np) {
                y = L.ids[t];
                if (y != x) {
                    int4 Y[8];
#pragma unroll
                    for (int s = 0; s < 8; s++) Y[s] = L.rows[t][s];
                    const uint32_t d = (uint32_t)(nx + L.norms[t] - 2 * rowDot(X, Y));
                    k1 = ((uint64_t)d << 32) | y;
                    k2 = ((uint64_t)d << 32) | x;
                    p1 = k1 < wx;
                    p2 = k2 < a.before[(size_t)y * a.K + a.K - 1];
                }
            }
            if constexpr (WRITE) {
                const uint64_t m1 = ballot(p1), m2 = ballot(p2);
                if (p1) {
                    const uint64_t q = pos + mbcnt(m1);
                    a.propKey[q] = k1;
                    a.propTgt[q] = x;
                }
                pos += (uint64_t)__popcll(m1);
                if (p2) {
                    const uint64_t q = pos + mbcnt(m2);
                    a.propKey[q] = k2;
                    a.propTgt[q] = y;
                }
                pos += (uint64_t)__popcll(m2);
            } else {
                cnt += (uint32_t)__popcll(ballot(p1)) + (uint32_t)__popcll(ballot(p2));
            }
        }
    }
    if (!WRITE && lane == 0) a.counts[u - a.u0] = cnt;
}

// Node x's list := the K smallest distinct keys of its list and its proposals (the run
// of x in the (target, key)-sorted proposals); entries already listed keep their flag,
// new ones are flagged new.  src -> dst (lists and flags).
__global__ void pynnMergeKernel(const uint64_t* src, const uint8_t* srcNew, uint64_t* dst, uint8_t* dstNew,
                                const uint32_t* tgt, const uint64_t* keys, size_t count, uint32_t n,
                                uint32_t K) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    size_t lo = 0, hi = count;
    while (lo < hi) {
        const size_t mid = (lo + hi) >> 1;
        if (tgt[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    size_t pe = lo;
    while (pe < count && tgt[pe] == x) pe++;
    const uint64_t* A = src + (size_t)x * K;
    const uint8_t* F = srcNew + (size_t)x * K;
    uint64_t* D = dst + (size_t)x * K;
    uint8_t* DF = dstNew + (size_t)x * K;
    uint32_t ia = 0, c = 0;
    size_t ib = lo;
    uint64_t last = ~0ull;
    while (c < K) {
        const uint64_t av = ia < K ? A[ia] : ~0ull, bv = ib < pe ? keys[ib] : ~0ull;
        if (av == ~0ull && bv == ~0ull) break;
        uint64_t v;
        uint8_t f;
        if (av <= bv) {
            v = av;
            f = F[ia];
            ia++;
            if (bv == av) ib++;
        } else {
            v = bv;
            f = 1;
            ib++;
        }
        if (v == last) continue;
        D[c] = v;
        DF[c] = f;
        c++;
        last = v;
    }
    for (uint32_t t = c; t < K; t++) {
        D[t] = ~0ull;
        DF[t] = 0;
    }
}

// Number of list positions whose key changed (one atomic per block).
__global__ void pynnChangedKernel(const uint64_t* a, const uint64_t* b, size_t count,
                                  unsigned long long* changed) {
    __shared__ uint32_t tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    uint32_t c = 0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < count; p += (size_t)gridDim.x * blockDim.x)
        c += a[p] != b[p] ? 1u : 0u;
    atomicAdd(&tot, c);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(changed, (unsigned long long)tot);
}

// (x << 32 | i) for every listed x of node i; ~0 for unused entries.
__global__ void pynnReversePairsKernel(const uint64_t* lists, uint32_t n, uint32_t K, uint64_t* pairs) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (size_t)n * K) return;
    const uint32_t id = (uint32_t)lists[p];
    pairs[p] = id == kEmpty ? ~0ull : (((uint64_t)id << 32) | (uint32_t)(p / K));
}

struct PynnPruneArgs {
    const int8_t* vec;
    int32_t normsInRow;
    uint32_t n, K, R;
    float alpha;
    const uint64_t* lists;    // n x K
    const uint64_t* rev;      // (x << 32 | source) sorted
    size_t nrev;
    uint32_t* adj;            // n x R
    uint32_t* deg;
    uint32_t* deferred;
    uint32_t* ndeferred;
    uint32_t* err;
};

// Node i: its listed entries plus its reverse sources, robust-pruned to R.  items ==
// nullptr: every node, the wider prunes deferred; else the deferred nodes.
template <uint32_t CAP, int WPB>
__global__ __launch_bounds__(64 * WPB) void pynnPruneKernel(PynnPruneArgs a, const uint32_t* items) {
    __shared__ PruneLds<CAP> lds[WPB];
    const int lane = laneId(), w = threadIdx.x >> 6;
    PruneLds<CAP>& S = lds[w];
    const uint32_t total = items ? *a.ndeferred : a.n;
    for (uint32_t it = blockIdx.x * WPB + w; it < total; it += gridDim.x * WPB) {
        const uint32_t i = items ? items[it] : it;
        size_t lo = 0, hi = a.nrev;
        while (lo < hi) {
            const size_t mid = (lo + hi) >> 1;
            if ((a.rev[mid] >> 32) < i) lo = mid + 1;
            else hi = mid;
        }
        size_t pe = lo;
        while (pe < a.nrev && (a.rev[pe] >> 32) == i) pe++;
        uint32_t nl = 0;
        for (uint32_t t = 0; t < a.K; t++) nl += (uint32_t)a.lists[(size_t)i * a.K + t] != kEmpty ? 1u : 0u;
        const uint32_t nr = (uint32_t)(pe - lo), nc = nl + nr;
        if (nc > CAP) {
            if (lane == 0) {
                if (items == nullptr) a.deferred[atomicAdd(a.ndeferred, 1u)] = i;
                else atomicOr(a.err, 2u);
            }
            continue;
        }
        for (uint32_t t = lane; t < nl; t += 64) S.keys[t] = a.lists[(size_t)i * a.K + t];
        for (uint32_t t = lane; t < nr; t += 64) S.list[t] = (uint32_t)a.rev[lo + t];
        __builtin_amdgcn_wave_barrier();
        keysFromIds(a.vec, a.normsInRow, i, S.list, nr, S.keys + nl, lane);
        const uint32_t outc = robustPruneWave<CAP, false>(a.vec, a.normsInRow, a.R, i, a.alpha, S, nc, lane);
        for (uint32_t r = lane; r < a.R; r += 64) a.adj[(size_t)i * a.R + r] = r < outc ? S.out[r] : kEmpty;
        if (lane == 0) a.deg[i] = outc;
        __builtin_amdgcn_wave_barrier();
    }
}

struct BridgeLds {
    union {
        int4 rows[kLeafMax][8];
        struct {
            uint64_t edges[kLeafMax];
            uint16_t list[kLeafMax][kBridgeCapDev];
        } k;
    };
    uint32_t comp[kLeafMax];
    uint16_t parent[kLeafMax];
    uint8_t used[kLeafMax];
    uint8_t room[kLeafMax];   // bridges left when the tree started
    int32_t norms[kLeafMax];
    uint32_t nedges;
};

// Leaf bridges of one tree (the host builder's step 4): each point with room proposes its
// nearest leaf mate with room in another component, by (distance, leaf position); Kruskal
// over the components in (distance, lower, upper position) order, skipping pairs whose
// endpoint filled up in this leaf.  Accepted pairs go to both endpoints' lists, in
// acceptance order; bridges[] counts what each node took.
__global__ __launch_bounds__(kLeafThreads) void pynnBridgeKernel(const int8_t* vec, int normsInRow,
                                                                const uint32_t* ids, const uint32_t* leafB,
                                                                const uint32_t* leafE, const uint32_t* comp,
                                                                uint32_t* bridges, uint32_t* outList,
                                                                uint8_t* outCnt) {
    __shared__ BridgeLds S;
    const uint32_t b = leafB[blockIdx.x], m = leafE[blockIdx.x] - b;
    const uint32_t* Lg = ids + b;
    const uint32_t tid = threadIdx.x;
    for (uint32_t x = tid; x < m * 8; x += kLeafThreads) {
        const uint32_t i = x >> 3, s = x & 7;
        int4 v = *reinterpret_cast<const int4*>(vec + (size_t)Lg[i] * kVecBytes + s * 16);
        if (normsInRow && s == 7) v.w = 0;
        S.rows[i][s] = v;
    }
    for (uint32_t i = tid; i < m; i += kLeafThreads) {
        S.comp[i] = comp[Lg[i]];
        const uint32_t br = bridges[Lg[i]];
        S.room[i] = (uint8_t)(br < kBridgeCapDev ? kBridgeCapDev - br : 0);
        S.used[i] = 0;
    }
    if (tid == 0) S.nedges = 0;
    __syncthreads();
    for (uint32_t i = tid; i < m; i += kLeafThreads) {
        int4 X[8];
#pragma unroll
        for (int s = 0; s < 8; s++) X[s] = S.rows[i][s];
        S.norms[i] = rowDot(X, X);
    }
    __syncthreads();
    uint64_t best = ~0ull;
    const uint32_t i = tid;
    if (i < m) {
        int4 X[8];
#pragma unroll
        for (int s = 0; s < 8; s++) X[s] = S.rows[i][s];
        const int ni = S.norms[i];
        const uint32_t ci = S.comp[i];
        const bool roomI = S.room[i] > 0;
        uint32_t rep = i;
        for (uint32_t j = 0; j < m; j++) {
            const uint32_t cj = S.comp[j];
            if (cj == ci) {
                rep = min(rep, j);
                continue;
            }
            if (!roomI || S.room[j] == 0) continue;
            int4 Y[8];
#pragma unroll
            for (int s = 0; s < 8; s++) Y[s] = S.rows[j][s];
            const int d = ni + S.norms[j] - 2 * rowDot(X, Y);
            const uint64_t key = ((uint64_t)(uint32_t)d << 32) | j;
            best = key < best ? key : best;
        }
        S.parent[i] = (uint16_t)rep;
    }
    __syncthreads();  // rows are dead from here
    if (i < m && best != ~0ull) {
        const uint32_t j = (uint32_t)best;
        const uint64_t e = ((best >> 32) << 20) | ((uint64_t)min(i, j) << 10) | max(i, j);
        S.k.edges[atomicAdd(&S.nedges, 1u)] = e;
    }
    __syncthreads();
    const uint32_t E = S.nedges;
    uint32_t P = 2;
    while (P < E) P <<= 1;
    for (uint32_t x = E + tid; x < P; x += kLeafThreads) S.k.edges[x] = ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t x = tid; x < P; x += kLeafThreads) {
                const uint32_t y = x ^ j;
                if (y > x) {
                    const uint64_t p = S.k.edges[x], q = S.k.edges[y];
                    if ((p > q) == ((x & k) == 0)) {
                        S.k.edges[x] = q;
                        S.k.edges[y] = p;
                    }
                }
            }
            __syncthreads();
        }
    if (tid == 0) {
        uint64_t prev = ~0ull;
        for (uint32_t t = 0; t < E; t++) {
            const uint64_t e = S.k.edges[t];
            if (e == prev) continue;
            prev = e;
            const uint32_t x = (uint32_t)(e >> 10) & 1023u, y = (uint32_t)e & 1023u;
            const uint32_t rx = findHalving(S.parent, x), ry = findHalving(S.parent, y);
            if (rx == ry) continue;
            if (S.used[x] >= S.room[x] || S.used[y] >= S.room[y]) continue;
            S.parent[rx] = (uint16_t)ry;
            S.k.list[x][S.used[x]++] = (uint16_t)y;
            S.k.list[y][S.used[y]++] = (uint16_t)x;
        }
    }
    __syncthreads();
    if (i < m) {
        const uint32_t g = Lg[i], c = S.used[i];
        outCnt[g] = (uint8_t)c;
        for (uint32_t s = 0; s < c; s++) outList[(size_t)g * kBridgeCapDev + s] = Lg[S.k.list[i][s]];
        bridges[g] += c;
    }
}

static_assert(sizeof(PynnLeafLds) <= 160 * 1024, "PyNNDescent leaf LDS");
static_assert(sizeof(BridgeLds) <= 160 * 1024, "bridge LDS");


}  // namespace

// Random cluster trees over the n device rows: leaves (first, last) into ids (device).
// seed is the tree's seed (tools::mixSeed(builder seed, salt)).
hipError_t launchClusterTree(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t leafSize,
                             uint64_t seed, uint32_t* ids, uint32_t* tmp, uint8_t* flags,
                             SplitSeg* dSegs, uint32_t* dCuts,
                             std::vector<std::pair<uint32_t, uint32_t>>& leaves, hipStream_t s) {
    std::vector<uint32_t> iota(n);
    std::iota(iota.begin(), iota.end(), 0u);
    BUILD_TRY(hipMemcpyAsync(ids, iota.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    leaves.clear();
    std::vector<std::pair<uint32_t, uint32_t>> todo{{0u, n}};
    std::vector<SplitSeg> split;
    std::vector<uint32_t> cuts;
    uint64_t round = 0;
    while (!todo.empty()) {
        split.clear();
        for (auto sg : todo) {
            if (sg.second - sg.first <= leafSize) leaves.push_back(sg);
            else split.push_back({sg.first, sg.second, tools::mixSeed(seed, (round << 32) ^ sg.first)});
        }
        todo.clear();
        if (split.empty()) break;
        cuts.resize(split.size());
        BUILD_TRY(hipMemcpyAsync(dSegs, split.data(), split.size() * sizeof(SplitSeg), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(treeSplitKernel, dim3((uint32_t)split.size()), dim3(kSplitThreads), 0, s, vec,
                           normsInRow ? 1 : 0, ids, tmp, flags, static_cast<const SplitSeg*>(dSegs), dCuts);
        BUILD_TRY(hipGetLastError());
        BUILD_TRY(hipMemcpyAsync(cuts.data(), dCuts, split.size() * 4, hipMemcpyDeviceToHost, s));
        BUILD_TRY(hipStreamSynchronize(s));
        for (size_t k = 0; k < split.size(); k++) {
            todo.push_back({split[k].b, cuts[k]});
            todo.push_back({cuts[k], split[k].e});
        }
        round++;
    }
    std::sort(leaves.begin(), leaves.end());
    return hipSuccess;
}

// HCNNG over the n device rows of `vec`: adj[i] = node i's list (before any bridging).
hipError_t launchBuildHcnng(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t numTrees,
                            uint32_t leafSize, uint32_t mstDeg, uint64_t seed,
                            std::vector<std::vector<uint32_t>>& adj, hipStream_t s) {
    adj.assign(n, {});
    const uint32_t stride = numTrees * mstDeg;
    DevMem m;
    uint32_t *ids, *tmp, *cuts, *leafB, *leafE, *outAdj;
    uint8_t *flags, *outCnt;
    SplitSeg* segs;
    BUILD_TRY(m.alloc(&ids, n));
    BUILD_TRY(m.alloc(&tmp, n));
    BUILD_TRY(m.alloc(&flags, n));
    BUILD_TRY(m.alloc(&segs, n / 2 + 1));
    BUILD_TRY(m.alloc(&cuts, n / 2 + 1));
    BUILD_TRY(m.alloc(&leafB, n));
    BUILD_TRY(m.alloc(&leafE, n));
    BUILD_TRY(m.alloc(&outAdj, (size_t)n * stride));
    BUILD_TRY(m.alloc(&outCnt, (size_t)n * numTrees));
    BUILD_TRY(hipMemsetAsync(outCnt, 0, (size_t)n * numTrees, s));
    std::vector<std::pair<uint32_t, uint32_t>> leaves;
    std::vector<uint32_t> lb, le;
    for (uint32_t t = 0; t < numTrees; t++) {
        BUILD_TRY(launchClusterTree(vec, normsInRow, n, leafSize, tools::mixSeed(seed, 100 + t), ids, tmp,
                                    flags, segs, cuts, leaves, s));
        lb.clear();
        le.clear();
        for (auto& l : leaves)
            if (l.second - l.first >= 2) {
                lb.push_back(l.first);
                le.push_back(l.second);
            }
        if (lb.empty()) continue;
        BUILD_TRY(hipMemcpyAsync(leafB, lb.data(), lb.size() * 4, hipMemcpyHostToDevice, s));
        BUILD_TRY(hipMemcpyAsync(leafE, le.data(), le.size() * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(hcnngLeafKernel, dim3((uint32_t)lb.size()), dim3(kLeafThreads), 0, s, vec,
                           normsInRow ? 1 : 0, static_cast<const uint32_t*>(ids),
                           static_cast<const uint32_t*>(leafB), static_cast<const uint32_t*>(leafE), mstDeg,
                           outAdj, stride, t * mstDeg, outCnt, numTrees, t);
        BUILD_TRY(hipGetLastError());
        BUILD_TRY(hipStreamSynchronize(s));  // lb / le are reused by the next tree
    }
    std::vector<uint32_t> hAdj((size_t)n * stride);
    std::vector<uint8_t> hCnt((size_t)n * numTrees);
    BUILD_TRY(hipMemcpyAsync(hAdj.data(), outAdj, hAdj.size() * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(hCnt.data(), outCnt, hCnt.size(), hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t t = 0; t < numTrees; t++) {
            const uint32_t* r = hAdj.data() + (size_t)i * stride + t * mstDeg;
            adj[i].insert(adj[i].end(), r, r + hCnt[(size_t)i * numTrees + t]);
        }
    return hipSuccess;
}


namespace {

template <uint32_t KMAX>
hipError_t launchPynnLeaves(const int8_t* vec, bool normsInRow, const uint32_t* ids, const uint32_t* leafB,
                            const uint32_t* leafE, uint32_t nleaves, uint32_t K, uint64_t* out, hipStream_t s) {
    hipLaunchKernelGGL((pynnLeafKernel<KMAX>), dim3(nleaves), dim3(kPynnThreads), 0, s, vec, normsInRow ? 1 : 0,
                       ids, leafB, leafE, K, out);
    return hipGetLastError();
}

uint32_t bitsFor(uint32_t n) {
    uint32_t b = 1;
    while (b < 32 && (1ull << b) < (uint64_t)n) b++;
    return b;
}

// Sorted (target << 32 | source) pairs -> rev[x][0..S): the first S sources of x's run.
hipError_t reverseLists(const uint64_t* sorted, size_t count, uint32_t n, uint32_t S, uint32_t* first,
                        uint32_t* rev, uint32_t* revN, hipStream_t s) {
    BUILD_TRY(hipMemsetAsync(revN, 0, (size_t)n * 4, s));
    const uint32_t blocks = (uint32_t)((count + 255) / 256);
    hipLaunchKernelGGL(pynnRunHeadsKernel, dim3(blocks), dim3(256), 0, s, sorted, count, first);
    BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(pynnRunFillKernel, dim3(blocks), dim3(256), 0, s, sorted, count,
                       static_cast<const uint32_t*>(first), S, rev, revN);
    return hipGetLastError();
}

}  // namespace

// PyNNDescent over the n device rows of `vec`: adj[i] = node i's list (before the final
// bridging from node 0).  *status: 0, or 2 when a prune met more than kSortCap candidates.
hipError_t launchBuildPynndescent(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t K,
                                  uint32_t numTrees, uint32_t leafSize, float alpha, float delta,
                                  uint32_t maxDeg, uint64_t seed, int cus,
                                  std::vector<std::vector<uint32_t>>& adj, uint32_t* status, hipStream_t s) {
    *status = 0;
    K = std::min(K, n - 1);
    const uint32_t S = std::max<uint32_t>(1, K / 2);
    const size_t nK = (size_t)n * K;
    // proposals of one join pass: the per-node bound, and at most 2^28 at once
    const uint64_t nA = 2ull * S, nB = (uint64_t)K + S;
    const uint64_t perNode = nA * (nA > 0 ? nA - 1 : 0) + 2 * nA * nB;
    const uint64_t cap = std::max<uint64_t>(perNode, std::min<uint64_t>((uint64_t)n * perNode, 1ull << 28));

    DevMem m;
    uint32_t *ids, *tmp, *cuts, *leafB, *leafE, *nw, *nwN, *old, *oldN, *rnw, *rnwN, *rold, *roldN, *first;
    uint32_t *counts, *propTgt, *propTgt2, *adjD, *degD, *deferred, *ndeferred, *err, *comp, *bridges,
        *bList;
    uint8_t *flags, *newA, *newB, *bCnt;
    uint64_t *listA, *listB, *treeKeys, *before, *pairsNw, *pairsNwS, *pairsOld, *pairsOldS, *offsets,
        *propKey, *propKey2;
    unsigned long long* changed;
    SplitSeg* segs;
    BUILD_TRY(m.alloc(&ids, n));
    BUILD_TRY(m.alloc(&tmp, n));
    BUILD_TRY(m.alloc(&flags, n));
    BUILD_TRY(m.alloc(&segs, n / 2 + 1));
    BUILD_TRY(m.alloc(&cuts, n / 2 + 1));
    BUILD_TRY(m.alloc(&leafB, n));
    BUILD_TRY(m.alloc(&leafE, n));
    BUILD_TRY(m.alloc(&listA, nK));
    BUILD_TRY(m.alloc(&listB, nK));
    BUILD_TRY(m.alloc(&newA, nK));
    BUILD_TRY(m.alloc(&newB, nK));
    BUILD_TRY(m.alloc(&treeKeys, nK));
    BUILD_TRY(m.alloc(&before, nK));
    BUILD_TRY(m.alloc(&nw, (size_t)n * S));
    BUILD_TRY(m.alloc(&nwN, n));
    BUILD_TRY(m.alloc(&old, nK));
    BUILD_TRY(m.alloc(&oldN, n));
    BUILD_TRY(m.alloc(&rnw, (size_t)n * S));
    BUILD_TRY(m.alloc(&rnwN, n));
    BUILD_TRY(m.alloc(&rold, (size_t)n * S));
    BUILD_TRY(m.alloc(&roldN, n));
    BUILD_TRY(m.alloc(&first, n));
    BUILD_TRY(m.alloc(&pairsNw, (size_t)n * S));
    BUILD_TRY(m.alloc(&pairsNwS, (size_t)n * S));
    BUILD_TRY(m.alloc(&pairsOld, nK));
    BUILD_TRY(m.alloc(&pairsOldS, nK));
    BUILD_TRY(m.alloc(&counts, n));
    BUILD_TRY(m.alloc(&offsets, n));
    BUILD_TRY(m.alloc(&propKey, cap));
    BUILD_TRY(m.alloc(&propKey2, cap));
    BUILD_TRY(m.alloc(&propTgt, cap));
    BUILD_TRY(m.alloc(&propTgt2, cap));
    BUILD_TRY(m.alloc(&changed, 1));
    size_t tempBytes = 0, tb = 0;
    BUILD_TRY(rocprim::radix_sort_keys(nullptr, tb, pairsOld, pairsOldS, nK, 0, 64, s));
    tempBytes = std::max(tempBytes, tb);
    tb = 0;
    BUILD_TRY(rocprim::radix_sort_pairs(nullptr, tb, propKey, propKey2, propTgt, propTgt2, (size_t)cap, 0, 55, s));
    tempBytes = std::max(tempBytes, tb);
    tb = 0;
    BUILD_TRY(rocprim::radix_sort_pairs(nullptr, tb, propTgt2, propTgt, propKey2, propKey, (size_t)cap, 0,
                                        bitsFor(n), s));
    tempBytes = std::max(tempBytes, tb);
    unsigned char* temp;
    BUILD_TRY(m.alloc(&temp, tempBytes));

    // 1. K-NN lists from the cluster trees
    BUILD_TRY(hipMemsetAsync(listA, 0xFF, nK * 8, s));
    BUILD_TRY(hipMemsetAsync(newA, 0, nK, s));
    std::vector<std::pair<uint32_t, uint32_t>> leaves;
    std::vector<uint32_t> lb, le;
    const uint32_t nodeBlocks = (n + 255) / 256;
    auto leafRanges = [&](bool needTwo) {
        lb.clear();
        le.clear();
        for (auto& l : leaves)
            if (!needTwo || l.second - l.first >= 2) {
                lb.push_back(l.first);
                le.push_back(l.second);
            }
    };
    for (uint32_t t = 0; t < numTrees; t++) {
        BUILD_TRY(launchClusterTree(vec, normsInRow, n, leafSize, tools::mixSeed(seed, 200 + t), ids, tmp, flags,
                                    segs, cuts, leaves, s));
        leafRanges(true);
        BUILD_TRY(hipMemsetAsync(treeKeys, 0xFF, nK * 8, s));
        if (!lb.empty()) {
            BUILD_TRY(hipMemcpyAsync(leafB, lb.data(), lb.size() * 4, hipMemcpyHostToDevice, s));
            BUILD_TRY(hipMemcpyAsync(leafE, le.data(), le.size() * 4, hipMemcpyHostToDevice, s));
            const uint32_t nl = (uint32_t)lb.size();
            if (K <= 16) BUILD_TRY(launchPynnLeaves<16>(vec, normsInRow, ids, leafB, leafE, nl, K, treeKeys, s));
            else if (K <= 40) BUILD_TRY(launchPynnLeaves<40>(vec, normsInRow, ids, leafB, leafE, nl, K, treeKeys, s));
            else BUILD_TRY(launchPynnLeaves<64>(vec, normsInRow, ids, leafB, leafE, nl, K, treeKeys, s));
        }
        hipLaunchKernelGGL(pynnMergeInitKernel, dim3(nodeBlocks), dim3(256), 0, s,
                           static_cast<const uint64_t*>(listA), static_cast<const uint64_t*>(treeKeys), listB,
                           newB, n, K);
        BUILD_TRY(hipGetLastError());
        std::swap(listA, listB);
        std::swap(newA, newB);
        BUILD_TRY(hipStreamSynchronize(s));  // lb / le are reused by the next tree
    }

    // 2. NN-descent
    std::vector<uint32_t> hCounts(n);
    std::vector<uint64_t> hOff;
    for (int it = 0; it < 12; it++) {
        hipLaunchKernelGGL(pynnSampleKernel, dim3(nodeBlocks), dim3(256), 0, s, static_cast<const uint64_t*>(listA),
                           newA, n, K, S, nw, nwN, old, oldN, pairsNw, pairsOld);
        BUILD_TRY(hipGetLastError());
        tb = tempBytes;
        BUILD_TRY(rocprim::radix_sort_keys(temp, tb, pairsNw, pairsNwS, (size_t)n * S, 0, 64, s));
        BUILD_TRY(reverseLists(pairsNwS, (size_t)n * S, n, S, first, rnw, rnwN, s));
        tb = tempBytes;
        BUILD_TRY(rocprim::radix_sort_keys(temp, tb, pairsOld, pairsOldS, nK, 0, 64, s));
        BUILD_TRY(reverseLists(pairsOldS, nK, n, S, first, rold, roldN, s));
        BUILD_TRY(hipMemcpyAsync(before, listA, nK * 8, hipMemcpyDeviceToDevice, s));

        JoinArgs ja = {};
        ja.vec = vec;
        ja.normsInRow = normsInRow ? 1 : 0;
        ja.K = K;
        ja.S = S;
        ja.before = before;
        ja.nw = nw;
        ja.nwN = nwN;
        ja.rnw = rnw;
        ja.rnwN = rnwN;
        ja.old = old;
        ja.oldN = oldN;
        ja.rold = rold;
        ja.roldN = roldN;
        ja.u0 = 0;
        ja.u1 = n;
        ja.counts = counts;
        ja.offsets = offsets;
        ja.propKey = propKey;
        ja.propTgt = propTgt;
        hipLaunchKernelGGL((pynnJoinKernel<false>), dim3((n + kJoinWaves - 1) / kJoinWaves), dim3(64 * kJoinWaves),
                           0, s, ja);
        BUILD_TRY(hipGetLastError());
        BUILD_TRY(hipMemcpyAsync(hCounts.data(), counts, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        BUILD_TRY(hipStreamSynchronize(s));
        // passes over node ranges whose proposals fit the buffers
        uint32_t u0 = 0;
        while (u0 < n) {
            uint32_t u1 = u0;
            uint64_t tot = 0;
            hOff.clear();
            while (u1 < n && tot + hCounts[u1] <= cap) {
                hOff.push_back(tot);
                tot += hCounts[u1];
                u1++;
            }
            if (u1 == u0) return hipErrorOutOfMemory;  // cannot happen: cap >= perNode
            if (tot > 0) {
                BUILD_TRY(hipMemcpyAsync(offsets, hOff.data(), hOff.size() * 8, hipMemcpyHostToDevice, s));
                ja.u0 = u0;
                ja.u1 = u1;
                hipLaunchKernelGGL((pynnJoinKernel<true>), dim3((u1 - u0 + kJoinWaves - 1) / kJoinWaves),
                                   dim3(64 * kJoinWaves), 0, s, ja);
                BUILD_TRY(hipGetLastError());
                tb = tempBytes;
                BUILD_TRY(rocprim::radix_sort_pairs(temp, tb, propKey, propKey2, propTgt, propTgt2, (size_t)tot, 0,
                                                    55, s));
                tb = tempBytes;
                BUILD_TRY(rocprim::radix_sort_pairs(temp, tb, propTgt2, propTgt, propKey2, propKey, (size_t)tot, 0,
                                                    bitsFor(n), s));
                hipLaunchKernelGGL(pynnMergeKernel, dim3(nodeBlocks), dim3(256), 0, s,
                                   static_cast<const uint64_t*>(listA), static_cast<const uint8_t*>(newA), listB,
                                   newB, static_cast<const uint32_t*>(propTgt),
                                   static_cast<const uint64_t*>(propKey), (size_t)tot, n, K);
                BUILD_TRY(hipGetLastError());
                std::swap(listA, listB);
                std::swap(newA, newB);
                BUILD_TRY(hipStreamSynchronize(s));  // hOff is reused by the next pass
            }
            u0 = u1;
        }
        BUILD_TRY(hipMemsetAsync(changed, 0, 8, s));
        hipLaunchKernelGGL(pynnChangedKernel, dim3(1024), dim3(256), 0, s, static_cast<const uint64_t*>(listA),
                           static_cast<const uint64_t*>(before), nK, changed);
        BUILD_TRY(hipGetLastError());
        unsigned long long hChanged = 0;
        BUILD_TRY(hipMemcpyAsync(&hChanged, changed, 8, hipMemcpyDeviceToHost, s));
        BUILD_TRY(hipStreamSynchronize(s));
        if ((double)hChanged < (double)delta * n * K) break;
    }

    // 3. reverse edges + robust prune to maxDeg
    BUILD_TRY(m.alloc(&adjD, (size_t)n * maxDeg));
    BUILD_TRY(m.alloc(&degD, n));
    BUILD_TRY(m.alloc(&deferred, n));
    BUILD_TRY(m.alloc(&ndeferred, 1));
    BUILD_TRY(m.alloc(&err, 1));
    BUILD_TRY(hipMemsetAsync(ndeferred, 0, 4, s));
    BUILD_TRY(hipMemsetAsync(err, 0, 4, s));
    hipLaunchKernelGGL(pynnReversePairsKernel, dim3((uint32_t)((nK + 255) / 256)), dim3(256), 0, s,
                       static_cast<const uint64_t*>(listA), n, K, pairsOld);
    BUILD_TRY(hipGetLastError());
    tb = tempBytes;
    BUILD_TRY(rocprim::radix_sort_keys(temp, tb, pairsOld, pairsOldS, nK, 0, 64, s));
    PynnPruneArgs pa = {};
    pa.vec = vec;
    pa.normsInRow = normsInRow ? 1 : 0;
    pa.n = n;
    pa.K = K;
    pa.R = maxDeg;
    pa.alpha = alpha;
    pa.lists = listA;
    pa.rev = pairsOldS;
    pa.nrev = nK;
    pa.adj = adjD;
    pa.deg = degD;
    pa.deferred = deferred;
    pa.ndeferred = ndeferred;
    pa.err = err;
    const uint32_t pruneBlocks = (uint32_t)std::max(1, cus) * 5;
    hipLaunchKernelGGL((pynnPruneKernel<kSmallCap, kPruneWaves>), dim3(pruneBlocks), dim3(64 * kPruneWaves), 0, s,
                       pa, nullptr);
    BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL((pynnPruneKernel<kSortCap, 1>), dim3((uint32_t)std::max(1, cus)), dim3(64), 0, s, pa,
                       static_cast<const uint32_t*>(deferred));
    BUILD_TRY(hipGetLastError());
    std::vector<uint32_t> hAdj((size_t)n * maxDeg), hDeg(n);
    BUILD_TRY(hipMemcpyAsync(hAdj.data(), adjD, hAdj.size() * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(hDeg.data(), degD, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(status, err, 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipStreamSynchronize(s));
    if (*status) return hipSuccess;
    adj.assign(n, {});
    for (uint32_t i = 0; i < n; i++)
        adj[i].assign(hAdj.data() + (size_t)i * maxDeg, hAdj.data() + (size_t)i * maxDeg + hDeg[i]);

    // 4. leaf bridges between the pruned graph's components
    std::vector<uint32_t> hComp(n);
    {
        std::vector<uint32_t> p(n);
        std::iota(p.begin(), p.end(), 0u);
        auto find = [&](uint32_t x) {
            while (p[x] != x) x = p[x] = p[p[x]];
            return x;
        };
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t a : adj[i]) {
                const uint32_t ra = find(i), rb = find(a);
                if (ra != rb) p[std::max(ra, rb)] = std::min(ra, rb);
            }
        for (uint32_t i = 0; i < n; i++) hComp[i] = find(i);
    }
    BUILD_TRY(m.alloc(&comp, n));
    BUILD_TRY(m.alloc(&bridges, n));
    BUILD_TRY(m.alloc(&bList, (size_t)n * kBridgeCapDev));
    BUILD_TRY(m.alloc(&bCnt, n));
    BUILD_TRY(hipMemcpyAsync(comp, hComp.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    BUILD_TRY(hipMemsetAsync(bridges, 0, (size_t)n * 4, s));
    std::vector<uint32_t> hList((size_t)n * kBridgeCapDev);
    std::vector<uint8_t> hCnt(n);
    for (uint32_t t = 0; t < numTrees; t++) {
        BUILD_TRY(launchClusterTree(vec, normsInRow, n, leafSize, tools::mixSeed(seed, 200 + t), ids, tmp, flags,
                                    segs, cuts, leaves, s));
        leafRanges(true);
        BUILD_TRY(hipMemsetAsync(bCnt, 0, n, s));
        if (!lb.empty()) {
            BUILD_TRY(hipMemcpyAsync(leafB, lb.data(), lb.size() * 4, hipMemcpyHostToDevice, s));
            BUILD_TRY(hipMemcpyAsync(leafE, le.data(), le.size() * 4, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(pynnBridgeKernel, dim3((uint32_t)lb.size()), dim3(kLeafThreads), 0, s, vec,
                               normsInRow ? 1 : 0, static_cast<const uint32_t*>(ids),
                               static_cast<const uint32_t*>(leafB), static_cast<const uint32_t*>(leafE),
                               static_cast<const uint32_t*>(comp), bridges, bList, bCnt);
            BUILD_TRY(hipGetLastError());
        }
        BUILD_TRY(hipMemcpyAsync(hList.data(), bList, hList.size() * 4, hipMemcpyDeviceToHost, s));
        BUILD_TRY(hipMemcpyAsync(hCnt.data(), bCnt, n, hipMemcpyDeviceToHost, s));
        BUILD_TRY(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < n; i++)
            adj[i].insert(adj[i].end(), hList.data() + (size_t)i * kBridgeCapDev,
                          hList.data() + (size_t)i * kBridgeCapDev + hCnt[i]);
    }
    return hipSuccess;
}

}  // namespace pann
