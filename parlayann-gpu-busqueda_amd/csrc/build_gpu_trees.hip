// build_gpu_trees.hip -- HCNNG graph construction on gfx950.
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
#include "devmem.h"
#include "device_common.h"
#include "pann_internal.h"
#include "tools_common.h"

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

}  // namespace pann
