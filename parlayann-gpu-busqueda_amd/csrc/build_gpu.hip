// build_gpu.hip -- Vamana (DiskANN) graph construction on gfx950.
//
// The algorithm is the host builder's (build_vamana.cpp), which follows ParlayANN's
// published batch insertion (README.md:346-349 names it as the GPU build to add):
// points in a seeded random order, batches doubling from 1 to 2% of n; per batch
//   1. buildSearchKernel: each point runs a greedy beam search of width L from node 0
//      on the pre-batch graph, with an exact visited set, and records the nodes it
//      expanded with their distances;
//   2. buildPruneKernel: each point robust-prunes (alpha) its expanded nodes plus its
//      current row to at most R out-neighbours and writes its own row, emitting
//      (target << 32 | source) for every new edge;
//   3. the edges are radix-sorted (rocPRIM), so each target's new in-edges form one
//      run in source order, and buildReverseKernel gives each run to one wavefront,
//      which appends the absent sources to the target's row or, past R, robust-prunes
//      the row together with them.
// Every phase writes only rows it owns and reads the graph as the previous phase left
// it, so the result is the host builder's graph, bit for bit.
//
// Search state of one wavefront: the frontier is the best L (by (distance, id)) of all
// nodes marked so far, as 64-bit keys in C registers (element e = c*64 + lane) plus an
// expanded bitmask per register.  A hop expands the first unexpanded entry, marks its
// unmarked neighbours in a per-wavefront bitmap over all n ids (one atomic OR each,
// which also answers "was it marked"), computes their distances (8 lanes per row,
// v_dot4), sorts them and merges them in by rank.  The host search inserts neighbours
// one at a time and drops a neighbour no better than a full frontier's last entry; the
// result is the same top-L set, and the next node it expands is the same first
// unexpanded entry.  At the end the wavefront clears exactly the bitmap words it set
// (the start node and the rows of the expanded nodes), so bitmaps never need a sweep.
//
// Prunes run in two tiers: four wavefronts per block with room for 512 candidates
// each, and the rare wider prunes (up to 2048 candidates) deferred to a second launch
// with one wavefront per block.
#include "build_prune.h"
#include "devmem.h"
#include "device_common.h"
#include "hnsw_plan.h"
#include "tools_common.h"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

namespace pann {
namespace {

using namespace dev;

constexpr int kSearchWaves = 4;            // wavefronts per search block
constexpr uint32_t kVisitCap = 1024;       // expanded nodes recorded per search
constexpr uint32_t kErrVisit = 1, kErrPrune = 2;
constexpr size_t kMarkBudget = 16ull << 30; // bytes of search bitmaps at most


struct BuildArgs {
    const int8_t* vec;      // n rows, kVecBytes pitch, zero padded
    int32_t normsInRow;     // rows hold |p|^2 in bytes 124..127
    uint32_t n, R, L;
    uint32_t* adj;          // n x R, kEmpty padded
    uint32_t* deg;          // n
    const uint32_t* batch;  // the batch's points
    uint32_t size;
    uint64_t* visited;      // size x kVisitCap keys (distance << 32 | id)
    uint32_t* nvisited;     // size; kVisitCap + 1 marks an overflow
    uint32_t* marks;        // per search wavefront: markWords bitmap words
    uint32_t markWords;
    uint32_t searchWaves;   // wavefronts of the search launch (the stride over the batch)
    uint64_t* pairs;        // size x R new edges (target << 32 | source), ~0 unused
    const uint64_t* sorted; // the pairs, sorted
    uint32_t npairs;
    const uint32_t* heads;  // first index of each target's run in sorted
    const uint32_t* nheads;
    uint32_t* deferred;     // items (batch or head indices) left for the wide tier
    uint32_t* ndeferred;
    uint32_t* err;
};

// ---- the search ----

// Number of keys of the sorted register `ck` (64 lanes, ~0 padded) below x.
__device__ __forceinline__ uint32_t countBelow64(uint64_t ck, uint64_t x) {
    uint32_t pos = 0;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        const uint64_t probe = shfl64(ck, (int)(pos + s - 1));
        if (probe < x) pos += s;
    }
    if (pos == 63 && readlane64(ck, 63) < x) pos = 64;
    return pos;
}

// Number of the F frontier keys below y (frontier sorted over C registers).
template <int C>
__device__ __forceinline__ uint32_t countBelowFrontier(const uint64_t (&fk)[C], uint32_t F, uint64_t y) {
    uint32_t pos = 0;
#pragma unroll
    for (int s = 64 * C; s >= 1; s >>= 1) {
        const uint32_t idx = pos + s - 1;   // probe index; lanes differ
        uint64_t probe = ~0ull;
#pragma unroll
        for (int c = 0; c < C; c++) {
            const uint64_t v = shfl64(fk[c], (int)(idx & 63));
            if ((idx >> 6) == (uint32_t)c) probe = v;
        }
        if (pos + s <= F && probe < y) pos += s;
    }
    return pos;
}

// Merges the candidate keys ck (unsorted, ~0 padded, none already in the frontier) into
// the frontier, keeping its best L; expanded flags travel with their entries.
template <int C>
__device__ void mergeFrontier(uint64_t (&fk)[C], uint64_t (&ex)[C], uint32_t& F, uint64_t ck,
                              uint32_t L, uint64_t* mkey, unsigned char* mflag, int lane) {
    uint64_t s1[1] = {ck};
    bitonicSort<1>(s1, lane);
    ck = s1[0];
    const uint32_t nc = (uint32_t)__popcll(ballot(ck != ~0ull));
    if (nc == 0) return;
#pragma unroll
    for (int c = 0; c < C; c++) {
        const uint32_t e = c * 64 + lane;
        const uint32_t np = e + countBelow64(ck, fk[c]);
        if (e < F && np < L) {
            mkey[np] = fk[c];
            mflag[np] = (unsigned char)((ex[c] >> lane) & 1);
        }
    }
    {
        const uint32_t np = (uint32_t)lane + countBelowFrontier<C>(fk, F, ck);
        if ((uint32_t)lane < nc && np < L) {
            mkey[np] = ck;
            mflag[np] = 0;
        }
    }
    F = min(L, F + nc);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int c = 0; c < C; c++) {
        const uint32_t e = c * 64 + lane;
        fk[c] = e < F ? mkey[e] : ~0ull;
        const bool fl = e < F && mflag[e] != 0;
        ex[c] = ballot(fl);
    }
    __builtin_amdgcn_wave_barrier();
}

// Rows of one graph layer: node u's row is adj + index(u) * width, its length deg[index(u)];
// index(u) = u, or, for an HNSW upper layer, upStart[u] + layer - 1.
struct LayerRows {
    const uint32_t* adj;
    const uint32_t* deg;
    uint32_t width;
    const uint32_t* upStart;  // nullptr: index(u) = u
    uint32_t layer;
    __device__ uint32_t index(uint32_t u) const { return upStart ? upStart[u] + layer - 1 : u; }
};

// Marks the ids pend[0..np) already known fresh, scores them against the query slice and
// merges them into the frontier.
template <int C>
__device__ void scoreAndMerge(const int8_t* vec, const uint32_t* pend, uint32_t np, const int4& qv, int qq,
                              int normMask, uint64_t (&fk)[C], uint64_t (&ex)[C], uint32_t& F, uint32_t L,
                              uint64_t* mkey, unsigned char* mflag, int lane) {
    const int sub = lane & 7, grp = lane >> 3;
    for (uint32_t b0 = 0; b0 < np; b0 += 64) {
        const uint32_t nb = min(64u, np - b0);
        int4 v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            v[r] = make_int4(0, 0, 0, 0);
            const uint32_t t = (uint32_t)r * 8 + grp;
            if (t < nb) v[r] = *reinterpret_cast<const int4*>(vec + (size_t)pend[b0 + t] * kVecBytes + sub * 16);
        }
        int part8[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            v[r].w &= ~normMask;
            part8[r] = slicePartial(v[r], qv);
        }
        const int tot = transposeReduce8(part8, lane);
        const int mine = __shfl(tot, (lane & 7) * 8 + (lane >> 3), 64);
        const uint64_t ck = (uint32_t)lane < nb ? keyOf(qq + mine, pend[b0 + lane]) : ~0ull;
        mergeFrontier<C>(fk, ex, F, ck, L, mkey, mflag, lane);
    }
}

// The host builders' search (BuildSearcher::search, tools_common.h) on one layer: starts
// (LDS, nstarts ids) marked and kept as the best L, then hops until every frontier entry
// is expanded.  The expanded keys go to vis[0..*nv) (HBM); *nv = visCap + 1 on overflow.
// On return the frontier is fk (F entries) and every bitmap word the search set is clear.
template <int C>
__device__ uint32_t exactSearch(const int8_t* vec, const LayerRows& g, const uint32_t* starts,
                                uint32_t nstarts, uint32_t L, const int4& qv, int qq, int normMask,
                                uint32_t* marks, uint64_t* mkey, unsigned char* mflag, uint32_t* pend,
                                uint64_t* vis, uint32_t visCap, uint64_t (&fk)[C], uint32_t* nvOut, int lane) {
    uint64_t ex[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        fk[c] = ~0ull;
        ex[c] = 0;
    }
    uint32_t F = 0;
    for (uint32_t b0 = 0; b0 < nstarts; b0 += 64) {
        const uint32_t t = b0 + lane;
        const uint32_t x = t < nstarts ? starts[t] : kEmpty;
        bool fresh = false;
        if (x != kEmpty) {
            const uint32_t bit = 1u << (x & 31);
            fresh = (atomicOr(&marks[x >> 5], bit) & bit) == 0;
        }
        const uint64_t fm = ballot(fresh);
        if (fresh) pend[mbcnt(fm)] = x;
        __builtin_amdgcn_wave_barrier();
        scoreAndMerge<C>(vec, pend, (uint32_t)__popcll(fm), qv, qq, normMask, fk, ex, F, L, mkey, mflag, lane);
        __builtin_amdgcn_wave_barrier();
    }
    uint32_t nv = 0;
    bool overflow = false;
    while (true) {
        int e = -1;
#pragma unroll
        for (int c = 0; c < C; c++) {
            const uint64_t m = ballot((uint32_t)(c * 64 + lane) < F) & ~ex[c];
            if (e < 0 && m) e = c * 64 + __ffsll((unsigned long long)m) - 1;
        }
        if (e < 0) break;
        if (nv >= visCap) {
            overflow = true;
            break;
        }
        uint64_t cur = 0;
#pragma unroll
        for (int c = 0; c < C; c++)
            if (c == (e >> 6)) {
                cur = readlane64(fk[c], e & 63);
                ex[c] |= 1ull << (e & 63);
            }
        if (lane == 0) vis[nv] = cur;
        nv++;
        const uint32_t u = (uint32_t)cur;
        const uint32_t ri = g.index(u);
        const uint32_t du = g.deg[ri];
        const uint32_t* row = g.adj + (size_t)ri * g.width;
        uint32_t np = 0;
        for (uint32_t ch = 0; ch * 64 < du; ch++) {
            const uint32_t j = ch * 64 + lane;
            const uint32_t x = j < du ? row[j] : kEmpty;
            bool fresh = false;
            if (x != kEmpty) {
                const uint32_t bit = 1u << (x & 31);
                fresh = (atomicOr(&marks[x >> 5], bit) & bit) == 0;
            }
            const uint64_t fm = ballot(fresh);
            if (fresh) pend[np + mbcnt(fm)] = x;
            np += (uint32_t)__popcll(fm);
        }
        __builtin_amdgcn_wave_barrier();
        scoreAndMerge<C>(vec, pend, np, qv, qq, normMask, fk, ex, F, L, mkey, mflag, lane);
    }
    // clear the bitmap words this search set: those of the starts and of the expanded
    // rows, whose ids are read back past the vector L1 (it does not see this wave's stores)
    *nvOut = overflow ? visCap + 1 : nv;
    for (uint32_t t = lane; t < nstarts; t += 64) marks[starts[t] >> 5] = 0;
    __builtin_amdgcn_s_waitcnt(0);
    for (uint32_t t = 0; t < nv; t++) {
        const uint32_t u = __hip_atomic_load(reinterpret_cast<const uint32_t*>(vis + t), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t ri = g.index(u);
        const uint32_t du = g.deg[ri];
        const uint32_t* row = g.adj + (size_t)ri * g.width;
        for (uint32_t j = lane; j < du; j += 64) {
            const uint32_t x = row[j];
            if (x != kEmpty) marks[x >> 5] = 0;
        }
    }
    // the next search's atomics must find these words cleared
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    return F;
}

// Vamana: each batch point's search from node 0; its expanded nodes are the prune's input.
template <int C>
__global__ __launch_bounds__(64 * kSearchWaves) void buildSearchKernel(BuildArgs a) {
    __shared__ uint64_t mkey[kSearchWaves][64 * C];
    __shared__ unsigned char mflag[kSearchWaves][64 * C];
    __shared__ uint32_t pend[kSearchWaves][128];
    __shared__ uint32_t start[kSearchWaves][1];
    const int lane = laneId(), w = threadIdx.x >> 6;
    const uint32_t wave = blockIdx.x * kSearchWaves + w;
    uint32_t* marks = a.marks + (size_t)wave * a.markWords;
    const LayerRows g{a.adj, a.deg, a.R, nullptr, 0};
    const int sub = lane & 7;
    const int normMask = (a.normsInRow && sub == 7) ? -1 : 0;
    if (lane == 0) start[w][0] = 0;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = wave; i < a.size; i += a.searchWaves) {
        const uint32_t p = a.batch[i];
        int4 qv = *reinterpret_cast<const int4*>(a.vec + (size_t)p * kVecBytes + sub * 16);
        qv.w &= ~normMask;
        const int qq = groupSum8(sliceNorm(qv));
        uint64_t fk[C];
        uint32_t nv = 0;
        exactSearch<C>(a.vec, g, start[w], 1, a.L, qv, qq, normMask, marks, mkey[w], mflag[w], pend[w],
                       a.visited + (size_t)i * kVisitCap, kVisitCap, fk, &nv, lane);
        if (lane == 0) {
            a.nvisited[i] = nv;
            if (nv > kVisitCap) atomicOr(a.err, kErrVisit);
        }
    }
}

// Each batch point: its expanded nodes plus its current row, pruned to its new row.
// items == nullptr: every batch index, wider prunes deferred; else the deferred indices.
template <uint32_t CAP, int WPB>
__global__ __launch_bounds__(64 * WPB) void buildPruneKernel(BuildArgs a, float alpha, const uint32_t* items) {
    __shared__ PruneLds<CAP> lds[WPB];
    const int lane = laneId(), w = threadIdx.x >> 6;
    PruneLds<CAP>& S = lds[w];
    const uint32_t count = items ? *a.ndeferred : a.size;
    for (uint32_t it = blockIdx.x * WPB + w; it < count; it += gridDim.x * WPB) {
        const uint32_t i = items ? items[it] : it;
        const uint32_t p = a.batch[i];
        const uint32_t nv = a.nvisited[i];
        uint64_t* pr = a.pairs + (size_t)i * a.R;
        if (nv > kVisitCap) {
            for (uint32_t t = lane; t < a.R; t += 64) pr[t] = ~0ull;
            continue;
        }
        const uint32_t dp = a.deg[p];
        const uint32_t nc = nv + dp;
        if (nc > CAP) {
            if (items == nullptr) {
                if (lane == 0) a.deferred[atomicAdd(a.ndeferred, 1u)] = i;
            } else {
                if (lane == 0) atomicOr(a.err, kErrPrune);
                for (uint32_t t = lane; t < a.R; t += 64) pr[t] = ~0ull;
            }
            continue;
        }
        for (uint32_t t = lane; t < nv; t += 64) S.keys[t] = a.visited[(size_t)i * kVisitCap + t];
        for (uint32_t t = lane; t < dp; t += 64) S.rowT[t] = a.adj[(size_t)p * a.R + t];
        __builtin_amdgcn_wave_barrier();
        keysFromIds(a.vec, a.normsInRow, p, S.rowT, dp, S.keys + nv, lane);
        const uint32_t outc = robustPruneWave<CAP, false>(a.vec, a.normsInRow, a.R, p, alpha, S, nc, lane);
        for (uint32_t t = lane; t < a.R; t += 64) {
            const uint32_t x = t < outc ? S.out[t] : kEmpty;
            a.adj[(size_t)p * a.R + t] = x;
            pr[t] = t < outc ? (((uint64_t)x << 32) | p) : ~0ull;
        }
        if (lane == 0) a.deg[p] = outc;
        __builtin_amdgcn_wave_barrier();
    }
}

// First index of every target's run in the sorted edge list.
__global__ void buildHeadsKernel(const uint64_t* sorted, uint32_t npairs, uint32_t* heads, uint32_t* nheads) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npairs) return;
    const uint64_t x = sorted[i];
    if (x != ~0ull && (i == 0 || (sorted[i - 1] >> 32) != (x >> 32))) heads[atomicAdd(nheads, 1u)] = i;
}

// Each target with new in-edges: absent sources appended in source order, or, past R,
// the row and those sources robust-pruned together.  items == nullptr: every run, runs
// whose prune is wider than CAP deferred; else the deferred head indices.
template <uint32_t CAP, int WPB>
__global__ __launch_bounds__(64 * WPB) void buildReverseKernel(BuildArgs a, float alpha, const uint32_t* items) {
    __shared__ PruneLds<CAP> lds[WPB];
    const int lane = laneId(), w = threadIdx.x >> 6;
    PruneLds<CAP>& S = lds[w];
    const uint32_t count = items ? *a.ndeferred : *a.nheads;
    for (uint32_t it = blockIdx.x * WPB + w; it < count; it += gridDim.x * WPB) {
        const uint32_t h = items ? items[it] : it;
        const uint32_t start = a.heads[h];
        const uint32_t t = (uint32_t)(a.sorted[start] >> 32);
        const uint32_t dt = a.deg[t];
        for (uint32_t r = lane; r < dt; r += 64) S.rowT[r] = a.adj[(size_t)t * a.R + r];
        __builtin_amdgcn_wave_barrier();
        uint32_t nadd = 0;
        for (uint32_t b = start;; b += 64) {
            const uint32_t idx = b + lane;
            const uint64_t x = idx < a.npairs ? a.sorted[idx] : ~0ull;
            const bool inRun = x != ~0ull && (uint32_t)(x >> 32) == t;
            const uint32_t s = (uint32_t)x;
            bool present = false;
            if (inRun)
                for (uint32_t r = 0; r < dt; r++) present |= S.rowT[r] == s;
            const bool keep = inRun && !present && s != t;
            const uint64_t km = ballot(keep);
            const uint32_t pos = nadd + mbcnt(km);
            if (keep && pos < CAP) S.list[pos] = s;
            nadd += (uint32_t)__popcll(km);
            if (ballot(inRun) != ~0ull) break;
        }
        __builtin_amdgcn_wave_barrier();
        if (nadd == 0) continue;
        if (dt + nadd <= a.R) {
            for (uint32_t r = lane; r < nadd; r += 64) a.adj[(size_t)t * a.R + dt + r] = S.list[r];
            if (lane == 0) a.deg[t] = dt + nadd;
            continue;
        }
        if (nadd + dt > CAP) {
            if (lane == 0) {
                if (items == nullptr) a.deferred[atomicAdd(a.ndeferred, 1u)] = h;
                else atomicOr(a.err, kErrPrune);
            }
            continue;
        }
        keysFromIds(a.vec, a.normsInRow, t, S.list, nadd, S.keys, lane);
        keysFromIds(a.vec, a.normsInRow, t, S.rowT, dt, S.keys + nadd, lane);
        const uint32_t outc = robustPruneWave<CAP, false>(a.vec, a.normsInRow, a.R, t, alpha, S, nadd + dt, lane);
        for (uint32_t r = lane; r < a.R; r += 64) a.adj[(size_t)t * a.R + r] = r < outc ? S.out[r] : kEmpty;
        if (lane == 0) a.deg[t] = outc;
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- HNSW ----

struct HnswBuildArgs {
    const int8_t* vec;
    int32_t normsInRow;
    uint32_t n, M, M0, efc, top, layers, entry;
    uint32_t* adj0;             // n x M0
    uint32_t* deg0;             // n
    uint32_t* adjUp;            // blocks x M
    uint32_t* degUp;            // blocks
    const uint32_t* upStart;    // n: first upper-layer block of each node
    const uint32_t* level;      // n
    const uint32_t* blockNode;  // blocks: the node owning each upper-layer block
    const uint32_t* batch;
    uint32_t size;
    uint64_t* W;                // size x layers x efc: each layer's final frontier (keys)
    uint32_t* Wn;               // size x layers: its length; kEmpty where not searched
    uint64_t* wvis;             // per search wavefront: kVisitCap expanded keys
    uint32_t* marks;
    uint32_t markWords;
    uint32_t searchWaves;
    uint64_t* pairs;            // size x pairStride new edges (row id << 32 | source), ~0 unused
    uint32_t pairStride;        // M0 + top M
    const uint64_t* sorted;
    uint32_t npairs;
    const uint32_t* heads;
    const uint32_t* nheads;
    uint32_t* deferred;
    uint32_t* ndeferred;
    uint32_t* err;
};

__device__ __forceinline__ LayerRows hnswRows(const HnswBuildArgs& a, uint32_t l) {
    return l == 0 ? LayerRows{a.adj0, a.deg0, a.M0, nullptr, 0} : LayerRows{a.adjUp, a.degUp, a.M, a.upStart, l};
}

// Row id of node t's list on layer l: t on layer 0, n + its block above.
__device__ __forceinline__ uint32_t hnswRowId(const HnswBuildArgs& a, uint32_t t, uint32_t l) {
    return l == 0 ? t : a.n + a.upStart[t] + l - 1;
}

// Each batch point: greedy descent (width 1) from the entry down to its level, then a
// width-efc search on every layer from there to 0, each layer starting from the previous
// layer's final frontier; the frontiers are the selection's input.
template <int C>
__global__ __launch_bounds__(64 * kSearchWaves) void hnswBuildSearchKernel(HnswBuildArgs a) {
    __shared__ uint64_t mkey[kSearchWaves][64 * C];
    __shared__ unsigned char mflag[kSearchWaves][64 * C];
    __shared__ uint32_t pend[kSearchWaves][128];
    __shared__ uint32_t eps[kSearchWaves][64 * C];
    const int lane = laneId(), w = threadIdx.x >> 6;
    const uint32_t wave = blockIdx.x * kSearchWaves + w;
    uint32_t* marks = a.marks + (size_t)wave * a.markWords;
    uint64_t* vis = a.wvis + (size_t)wave * kVisitCap;
    const int sub = lane & 7;
    const int normMask = (a.normsInRow && sub == 7) ? -1 : 0;
    for (uint32_t i = wave; i < a.size; i += a.searchWaves) {
        const uint32_t p = a.batch[i];
        const uint32_t lp = min(a.level[p], a.top);
        int4 qv = *reinterpret_cast<const int4*>(a.vec + (size_t)p * kVecBytes + sub * 16);
        qv.w &= ~normMask;
        const int qq = groupSum8(sliceNorm(qv));
        if (lane == 0) eps[w][0] = a.entry;
        __builtin_amdgcn_wave_barrier();
        uint32_t neps = 1;
        bool bad = false;
        uint64_t fk[C];
        for (uint32_t l = a.top; l > lp && !bad; l--) {
            uint32_t nv = 0;
            exactSearch<C>(a.vec, hnswRows(a, l), eps[w], neps, 1, qv, qq, normMask, marks, mkey[w], mflag[w],
                           pend[w], vis, kVisitCap, fk, &nv, lane);
            if (nv > kVisitCap) {
                bad = true;
                break;
            }
            const uint32_t best = (uint32_t)readlane64(fk[0], 0);
            if (lane == 0) eps[w][0] = best;
            __builtin_amdgcn_wave_barrier();
            neps = 1;
        }
        for (int l = (int)lp; l >= 0 && !bad; l--) {
            uint32_t nv = 0;
            const uint32_t F = exactSearch<C>(a.vec, hnswRows(a, (uint32_t)l), eps[w], neps, a.efc, qv, qq, normMask,
                                              marks, mkey[w], mflag[w], pend[w], vis, kVisitCap, fk, &nv, lane);
            if (nv > kVisitCap) {
                bad = true;
                break;
            }
            uint64_t* Wl = a.W + ((size_t)i * a.layers + l) * a.efc;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const uint32_t idx = c * 64 + lane;
                if (idx < F) {
                    Wl[idx] = fk[c];
                    eps[w][idx] = (uint32_t)fk[c];
                }
            }
            if (lane == 0) a.Wn[(size_t)i * a.layers + l] = F;
            __builtin_amdgcn_wave_barrier();
            neps = F;
        }
        if (bad) {
            for (uint32_t l = lane; l < a.layers; l += 64) a.Wn[(size_t)i * a.layers + l] = kEmpty;
            if (lane == 0) atomicOr(a.err, kErrVisit);
        }
    }
}

// Each batch point and each of its layers: the frontier selected with the heuristic to
// M0 (layer 0) or M neighbours, written as the point's list; every pair slot is written.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void hnswBuildSelectKernel(HnswBuildArgs a, float alpha) {
    __shared__ PruneLds<kSmallCap> lds[WPB];
    const int lane = laneId(), w = threadIdx.x >> 6;
    PruneLds<kSmallCap>& S = lds[w];
    for (uint32_t i = blockIdx.x * WPB + w; i < a.size; i += gridDim.x * WPB) {
        const uint32_t p = a.batch[i];
        const uint32_t lp = min(a.level[p], a.top);
        uint64_t* pr = a.pairs + (size_t)i * a.pairStride;
        for (uint32_t l = 0; l <= a.top; l++) {
            const uint32_t m = l == 0 ? a.M0 : a.M;
            const uint32_t off = l == 0 ? 0 : a.M0 + (l - 1) * a.M;
            const uint32_t nc = l <= lp ? a.Wn[(size_t)i * a.layers + l] : kEmpty;
            if (nc == kEmpty || nc > a.efc) {
                for (uint32_t t = lane; t < m; t += 64) pr[off + t] = ~0ull;
                continue;
            }
            const uint64_t* Wl = a.W + ((size_t)i * a.layers + l) * a.efc;
            for (uint32_t t = lane; t < nc; t += 64) S.keys[t] = Wl[t];
            __builtin_amdgcn_wave_barrier();
            const uint32_t outc = robustPruneWave<kSmallCap, true>(a.vec, a.normsInRow, m, p, alpha, S, nc, lane);
            uint32_t* row = l == 0 ? a.adj0 + (size_t)p * a.M0 : a.adjUp + ((size_t)a.upStart[p] + l - 1) * a.M;
            for (uint32_t t = lane; t < m; t += 64) {
                const uint32_t x = t < outc ? S.out[t] : kEmpty;
                row[t] = x;
                pr[off + t] = t < outc ? (((uint64_t)hnswRowId(a, x, l) << 32) | p) : ~0ull;
            }
            if (lane == 0) {
                if (l == 0) a.deg0[p] = outc;
                else a.degUp[a.upStart[p] + l - 1] = outc;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Each (layer, target) with new in-edges: absent sources appended in source order, or,
// past the layer's width, the list and those sources selected together.  items == nullptr:
// every run, runs wider than CAP deferred; else the deferred head indices.
template <uint32_t CAP, int WPB>
__global__ __launch_bounds__(64 * WPB) void hnswBuildReverseKernel(HnswBuildArgs a, float alpha, const uint32_t* items) {
    __shared__ PruneLds<CAP> lds[WPB];
    const int lane = laneId(), w = threadIdx.x >> 6;
    PruneLds<CAP>& S = lds[w];
    const uint32_t count = items ? *a.ndeferred : *a.nheads;
    for (uint32_t it = blockIdx.x * WPB + w; it < count; it += gridDim.x * WPB) {
        const uint32_t h = items ? items[it] : it;
        const uint32_t start = a.heads[h];
        const uint32_t rid = (uint32_t)(a.sorted[start] >> 32);
        uint32_t t, cap;
        uint32_t* row;
        uint32_t* degp;
        if (rid < a.n) {
            t = rid;
            cap = a.M0;
            row = a.adj0 + (size_t)t * a.M0;
            degp = a.deg0 + t;
        } else {
            const uint32_t b = rid - a.n;
            t = a.blockNode[b];
            cap = a.M;
            row = a.adjUp + (size_t)b * a.M;
            degp = a.degUp + b;
        }
        const uint32_t dt = *degp;
        for (uint32_t r = lane; r < dt; r += 64) S.rowT[r] = row[r];
        __builtin_amdgcn_wave_barrier();
        uint32_t nadd = 0;
        for (uint32_t b = start;; b += 64) {
            const uint32_t idx = b + lane;
            const uint64_t x = idx < a.npairs ? a.sorted[idx] : ~0ull;
            const bool inRun = x != ~0ull && (uint32_t)(x >> 32) == rid;
            const uint32_t s = (uint32_t)x;
            bool present = false;
            if (inRun)
                for (uint32_t r = 0; r < dt; r++) present |= S.rowT[r] == s;
            const bool keep = inRun && !present && s != t;
            const uint64_t km = ballot(keep);
            const uint32_t pos = nadd + mbcnt(km);
            if (keep && pos < CAP) S.list[pos] = s;
            nadd += (uint32_t)__popcll(km);
            if (ballot(inRun) != ~0ull) break;
        }
        __builtin_amdgcn_wave_barrier();
        if (nadd == 0) continue;
        if (dt + nadd <= cap) {
            for (uint32_t r = lane; r < nadd; r += 64) row[dt + r] = S.list[r];
            if (lane == 0) *degp = dt + nadd;
            continue;
        }
        if (nadd + dt > CAP) {
            if (lane == 0) {
                if (items == nullptr) a.deferred[atomicAdd(a.ndeferred, 1u)] = h;
                else atomicOr(a.err, kErrPrune);
            }
            continue;
        }
        keysFromIds(a.vec, a.normsInRow, t, S.rowT, dt, S.keys, lane);
        keysFromIds(a.vec, a.normsInRow, t, S.list, nadd, S.keys + dt, lane);
        const uint32_t outc = robustPruneWave<CAP, true>(a.vec, a.normsInRow, cap, t, alpha, S, nadd + dt, lane);
        for (uint32_t r = lane; r < cap; r += 64) row[r] = r < outc ? S.out[r] : kEmpty;
        if (lane == 0) *degp = outc;
        __builtin_amdgcn_wave_barrier();
    }
}


hipError_t launchSearch(const BuildArgs& a, int C, uint32_t blocks, hipStream_t s) {
    switch (C) {
        case 1: hipLaunchKernelGGL((buildSearchKernel<1>), dim3(blocks), dim3(64 * kSearchWaves), 0, s, a); break;
        case 2: hipLaunchKernelGGL((buildSearchKernel<2>), dim3(blocks), dim3(64 * kSearchWaves), 0, s, a); break;
        default: hipLaunchKernelGGL((buildSearchKernel<4>), dim3(blocks), dim3(64 * kSearchWaves), 0, s, a); break;
    }
    return hipGetLastError();
}

}  // namespace

// Builds the graph over the n device rows of `vec` into host arrays outDeg (n) and
// outAdj (n x R, kEmpty padded).  *status: 0, or kErrVisit / kErrPrune when a search
// expanded more than kVisitCap nodes or a prune met more than kSortCap candidates.
hipError_t launchBuildVamana(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t R, uint32_t L,
                             float alpha, bool twoPass, uint64_t seed, int cus, uint32_t* outDeg,
                             uint32_t* outAdj, uint32_t* status, hipStream_t s) {
    *status = 0;
    const std::vector<uint32_t> order = tools::permutation(n, tools::mixSeed(seed, 1));
    const uint32_t maxBatch = std::max<uint32_t>(1, (uint32_t)(0.02 * n));
    const uint32_t markWords = (n + 31) / 32;
    const int C = L <= 64 ? 1 : L <= 128 ? 2 : 4;
    // up to six search wavefronts per SIMD, within the bitmap budget
    uint32_t searchBlocksMax = (uint32_t)std::max(1, cus) * 6;
    while (searchBlocksMax > 1 && (size_t)searchBlocksMax * kSearchWaves * markWords * 4 > kMarkBudget)
        searchBlocksMax /= 2;
    const uint32_t searchWavesMax = searchBlocksMax * kSearchWaves;
    const uint32_t pruneBlocks = (uint32_t)std::max(1, cus) * 5;
    const uint32_t wideBlocks = (uint32_t)std::max(1, cus);

    DevMem m;
    uint32_t *adj, *deg, *dOrder, *nvisited, *marks, *heads, *nheads, *deferred, *ndeferred, *err;
    uint64_t *visited, *pairs, *sorted;
    BUILD_TRY(m.alloc(&adj, (size_t)n * R));
    BUILD_TRY(m.alloc(&deg, n));
    BUILD_TRY(m.alloc(&dOrder, n));
    BUILD_TRY(m.alloc(&nvisited, maxBatch));
    BUILD_TRY(m.alloc(&visited, (size_t)maxBatch * kVisitCap));
    BUILD_TRY(m.alloc(&marks, (size_t)searchWavesMax * markWords));
    BUILD_TRY(m.alloc(&pairs, (size_t)maxBatch * R));
    BUILD_TRY(m.alloc(&sorted, (size_t)maxBatch * R));
    BUILD_TRY(m.alloc(&heads, (size_t)maxBatch * R));
    BUILD_TRY(m.alloc(&deferred, (size_t)maxBatch * R));
    BUILD_TRY(m.alloc(&nheads, 1));
    BUILD_TRY(m.alloc(&ndeferred, 1));
    BUILD_TRY(m.alloc(&err, 1));
    size_t tempBytes = 0;
    BUILD_TRY(rocprim::radix_sort_keys(nullptr, tempBytes, pairs, sorted, (size_t)maxBatch * R, 0, 64, s));
    unsigned char* temp;
    BUILD_TRY(m.alloc(&temp, tempBytes));
    BUILD_TRY(hipMemsetAsync(adj, 0xFF, (size_t)n * R * 4, s));
    BUILD_TRY(hipMemsetAsync(deg, 0, (size_t)n * 4, s));
    BUILD_TRY(hipMemsetAsync(marks, 0, (size_t)searchWavesMax * markWords * 4, s));
    BUILD_TRY(hipMemsetAsync(err, 0, 4, s));
    BUILD_TRY(hipMemcpyAsync(dOrder, order.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));

    BuildArgs a = {};
    a.vec = vec;
    a.normsInRow = normsInRow ? 1 : 0;
    a.n = n;
    a.R = R;
    a.L = L;
    a.adj = adj;
    a.deg = deg;
    a.visited = visited;
    a.nvisited = nvisited;
    a.marks = marks;
    a.markWords = markWords;
    a.pairs = pairs;
    a.sorted = sorted;
    a.heads = heads;
    a.nheads = nheads;
    a.deferred = deferred;
    a.ndeferred = ndeferred;
    a.err = err;

    for (int pass = twoPass ? 0 : 1; pass < 2; pass++) {
        const float al = pass == 0 ? 1.0f : alpha;
        uint32_t count = 0, bsize = 1;
        while (count < n) {
            const uint32_t size = std::min(bsize, n - count);
            bsize = std::min(maxBatch, bsize * 2);
            a.batch = dOrder + count;
            a.size = size;
            count += size;
            const uint32_t blocks = std::min(searchBlocksMax, (size + kSearchWaves - 1) / kSearchWaves);
            a.searchWaves = blocks * kSearchWaves;
            BUILD_TRY(launchSearch(a, C, blocks, s));
            // prunes of the batch points, the wide ones after the others
            BUILD_TRY(hipMemsetAsync(ndeferred, 0, 4, s));
            const uint32_t pb = std::min(pruneBlocks, (size + kPruneWaves - 1) / kPruneWaves);
            hipLaunchKernelGGL((buildPruneKernel<kSmallCap, kPruneWaves>), dim3(pb), dim3(64 * kPruneWaves), 0,
                               s, a, al, nullptr);
            BUILD_TRY(hipGetLastError());
            hipLaunchKernelGGL((buildPruneKernel<kSortCap, 1>), dim3(wideBlocks), dim3(64), 0, s, a, al,
                               static_cast<const uint32_t*>(deferred));
            BUILD_TRY(hipGetLastError());
            // new edges grouped by target, then the reverse phase
            a.npairs = size * R;
            size_t tb = tempBytes;
            BUILD_TRY(rocprim::radix_sort_keys(temp, tb, pairs, sorted, (size_t)a.npairs, 0, 64, s));
            BUILD_TRY(hipMemsetAsync(nheads, 0, 4, s));
            BUILD_TRY(hipMemsetAsync(ndeferred, 0, 4, s));
            hipLaunchKernelGGL(buildHeadsKernel, dim3((a.npairs + 255) / 256), dim3(256), 0, s, sorted,
                               a.npairs, heads, nheads);
            BUILD_TRY(hipGetLastError());
            hipLaunchKernelGGL((buildReverseKernel<kSmallCap, kPruneWaves>), dim3(pruneBlocks),
                               dim3(64 * kPruneWaves), 0, s, a, al, nullptr);
            BUILD_TRY(hipGetLastError());
            hipLaunchKernelGGL((buildReverseKernel<kSortCap, 1>), dim3(wideBlocks), dim3(64), 0, s, a, al,
                               static_cast<const uint32_t*>(deferred));
            BUILD_TRY(hipGetLastError());
        }
    }
    BUILD_TRY(hipMemcpyAsync(outDeg, deg, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(outAdj, adj, (size_t)n * R * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(status, err, 4, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}


hipError_t launchBuildHnsw(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t M, uint32_t efc, float ml,
                           float alpha, uint64_t seed, int cus, HnswBuildOut* out, uint32_t* status,
                           hipStream_t s) {
    *status = 0;
    tools::HnswPlan plan = tools::hnswPlan(n, ml, seed);
    const uint32_t M0 = 2 * M;
    const uint64_t blocks = plan.blocks;
    if ((uint64_t)n + blocks >= 0xFFFFFFFFull) return hipErrorInvalidValue;
    const uint32_t layers = plan.top + 1;
    const uint32_t pairStride = M0 + plan.top * M;
    const uint32_t markWords = (n + 31) / 32;
    const int C = efc <= 64 ? 1 : efc <= 128 ? 2 : 4;
    uint32_t searchBlocksMax = (uint32_t)std::max(1, cus) * 6;
    while (searchBlocksMax > 1 && (size_t)searchBlocksMax * kSearchWaves * markWords * 4 > kMarkBudget)
        searchBlocksMax /= 2;
    const uint32_t searchWavesMax = searchBlocksMax * kSearchWaves;
    const uint32_t pruneBlocks = (uint32_t)std::max(1, cus) * 5;
    const uint32_t wideBlocks = (uint32_t)std::max(1, cus);
    const uint32_t maxBatch = plan.maxBatch;

    std::vector<uint32_t> up32(n), blockNode(std::max<uint64_t>(1, blocks));
    for (uint32_t i = 0; i < n; i++) {
        up32[i] = (uint32_t)plan.upStart[i];
        for (uint32_t l = 0; l < plan.level[i]; l++) blockNode[plan.upStart[i] + l] = i;
    }

    DevMem m;
    uint32_t *adj0, *deg0, *adjUp, *degUp, *dUp, *dLevel, *dBlockNode, *dOrder, *Wn, *marks, *heads, *nheads,
        *deferred, *ndeferred, *err;
    uint64_t *W, *wvis, *pairs, *sorted;
    const size_t nBlocks = std::max<uint64_t>(1, blocks);
    BUILD_TRY(m.alloc(&adj0, (size_t)n * M0));
    BUILD_TRY(m.alloc(&deg0, n));
    BUILD_TRY(m.alloc(&adjUp, nBlocks * M));
    BUILD_TRY(m.alloc(&degUp, nBlocks));
    BUILD_TRY(m.alloc(&dUp, n));
    BUILD_TRY(m.alloc(&dLevel, n));
    BUILD_TRY(m.alloc(&dBlockNode, nBlocks));
    BUILD_TRY(m.alloc(&dOrder, std::max<size_t>(1, plan.order.size())));
    BUILD_TRY(m.alloc(&W, (size_t)maxBatch * layers * efc));
    BUILD_TRY(m.alloc(&Wn, (size_t)maxBatch * layers));
    BUILD_TRY(m.alloc(&wvis, (size_t)searchWavesMax * kVisitCap));
    BUILD_TRY(m.alloc(&marks, (size_t)searchWavesMax * markWords));
    BUILD_TRY(m.alloc(&pairs, (size_t)maxBatch * pairStride));
    BUILD_TRY(m.alloc(&sorted, (size_t)maxBatch * pairStride));
    BUILD_TRY(m.alloc(&heads, (size_t)maxBatch * pairStride));
    BUILD_TRY(m.alloc(&deferred, (size_t)maxBatch * pairStride));
    BUILD_TRY(m.alloc(&nheads, 1));
    BUILD_TRY(m.alloc(&ndeferred, 1));
    BUILD_TRY(m.alloc(&err, 1));
    size_t tempBytes = 0;
    BUILD_TRY(rocprim::radix_sort_keys(nullptr, tempBytes, pairs, sorted, (size_t)maxBatch * pairStride, 0, 64, s));
    unsigned char* temp;
    BUILD_TRY(m.alloc(&temp, tempBytes));
    BUILD_TRY(hipMemsetAsync(adj0, 0xFF, (size_t)n * M0 * 4, s));
    BUILD_TRY(hipMemsetAsync(deg0, 0, (size_t)n * 4, s));
    BUILD_TRY(hipMemsetAsync(adjUp, 0xFF, nBlocks * M * 4, s));
    BUILD_TRY(hipMemsetAsync(degUp, 0, nBlocks * 4, s));
    BUILD_TRY(hipMemsetAsync(marks, 0, (size_t)searchWavesMax * markWords * 4, s));
    BUILD_TRY(hipMemsetAsync(err, 0, 4, s));
    BUILD_TRY(hipMemcpyAsync(dUp, up32.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    BUILD_TRY(hipMemcpyAsync(dLevel, plan.level.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    BUILD_TRY(hipMemcpyAsync(dBlockNode, blockNode.data(), nBlocks * 4, hipMemcpyHostToDevice, s));
    if (!plan.order.empty())
        BUILD_TRY(hipMemcpyAsync(dOrder, plan.order.data(), plan.order.size() * 4, hipMemcpyHostToDevice, s));

    HnswBuildArgs a = {};
    a.vec = vec;
    a.normsInRow = normsInRow ? 1 : 0;
    a.n = n;
    a.M = M;
    a.M0 = M0;
    a.efc = efc;
    a.top = plan.top;
    a.layers = layers;
    a.entry = plan.entry;
    a.adj0 = adj0;
    a.deg0 = deg0;
    a.adjUp = adjUp;
    a.degUp = degUp;
    a.upStart = dUp;
    a.level = dLevel;
    a.blockNode = dBlockNode;
    a.W = W;
    a.Wn = Wn;
    a.wvis = wvis;
    a.marks = marks;
    a.markWords = markWords;
    a.pairs = pairs;
    a.pairStride = pairStride;
    a.sorted = sorted;
    a.heads = heads;
    a.nheads = nheads;
    a.deferred = deferred;
    a.ndeferred = ndeferred;
    a.err = err;

    const size_t total = plan.order.size();
    size_t count = 0;
    uint32_t bsize = 1;
    while (count < total) {
        const uint32_t size = (uint32_t)std::min<size_t>(bsize, total - count);
        a.batch = dOrder + count;
        a.size = size;
        count += size;
        bsize = std::min(maxBatch, bsize * 2);
        const uint32_t blocksS = std::min(searchBlocksMax, (size + kSearchWaves - 1) / kSearchWaves);
        a.searchWaves = blocksS * kSearchWaves;
        switch (C) {
            case 1: hipLaunchKernelGGL((hnswBuildSearchKernel<1>), dim3(blocksS), dim3(64 * kSearchWaves), 0, s, a); break;
            case 2: hipLaunchKernelGGL((hnswBuildSearchKernel<2>), dim3(blocksS), dim3(64 * kSearchWaves), 0, s, a); break;
            default: hipLaunchKernelGGL((hnswBuildSearchKernel<4>), dim3(blocksS), dim3(64 * kSearchWaves), 0, s, a); break;
        }
        BUILD_TRY(hipGetLastError());
        const uint32_t pb = std::min(pruneBlocks, (size + kPruneWaves - 1) / kPruneWaves);
        hipLaunchKernelGGL((hnswBuildSelectKernel<kPruneWaves>), dim3(pb), dim3(64 * kPruneWaves), 0, s, a, alpha);
        BUILD_TRY(hipGetLastError());
        a.npairs = size * pairStride;
        size_t tb = tempBytes;
        BUILD_TRY(rocprim::radix_sort_keys(temp, tb, pairs, sorted, (size_t)a.npairs, 0, 64, s));
        BUILD_TRY(hipMemsetAsync(nheads, 0, 4, s));
        BUILD_TRY(hipMemsetAsync(ndeferred, 0, 4, s));
        hipLaunchKernelGGL(buildHeadsKernel, dim3((a.npairs + 255) / 256), dim3(256), 0, s, sorted, a.npairs,
                           heads, nheads);
        BUILD_TRY(hipGetLastError());
        hipLaunchKernelGGL((hnswBuildReverseKernel<kSmallCap, kPruneWaves>), dim3(pruneBlocks),
                           dim3(64 * kPruneWaves), 0, s, a, alpha, nullptr);
        BUILD_TRY(hipGetLastError());
        hipLaunchKernelGGL((hnswBuildReverseKernel<kSortCap, 1>), dim3(wideBlocks), dim3(64), 0, s, a, alpha,
                           static_cast<const uint32_t*>(deferred));
        BUILD_TRY(hipGetLastError());
    }
    out->level = std::move(plan.level);
    out->upStart = std::move(plan.upStart);
    out->entry = plan.entry;
    out->M0 = M0;
    out->deg0.resize(n);
    out->adj0.resize((size_t)n * M0);
    out->degUp.resize(nBlocks);
    out->adjUp.resize(nBlocks * M);
    BUILD_TRY(hipMemcpyAsync(out->deg0.data(), deg0, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(out->adj0.data(), adj0, (size_t)n * M0 * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(out->degUp.data(), degUp, nBlocks * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(out->adjUp.data(), adjUp, nBlocks * M * 4, hipMemcpyDeviceToHost, s));
    BUILD_TRY(hipMemcpyAsync(status, err, 4, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

}  // namespace pann
