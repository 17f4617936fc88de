// device_common.h -- wave64 building blocks shared by the search kernels (gfx950).
#pragma once

#include "pann_internal.h"

namespace pann {
namespace dev {

// ---- per-section cycle accounting (compiled in only with -DPANN_PROFILE) ----
// Sections of the flat search; each wavefront sums clock64() deltas and adds them to
// g_profile once when it exits (the reference's -DPERFIL idea, cuda/busqueda.cu:15-32).
enum ProfSection {
    kProfInit = 0, kProfVisit, kProfFilter, kProfDist, kProfSort, kProfUnion, kProfMerge,
    kProfCut, kProfNext, kProfOutput, kProfHops, kProfQueries,
    kProfCands,    // candidates below the cutoff, summed over hops
    kProfNoCands,  // hops with no candidate below the cutoff
    kProfSmall,    // hops with 1..8 candidates below the cutoff
    kProfPredFC,   // hops whose next node was the prefetched one
    kProfSections
};
#ifdef PANN_PROFILE
extern __device__ unsigned long long g_profile[kProfSections];
struct Prof {
    unsigned long long acc[kProfSections];
    long long t;
    __device__ void start() {
        for (int i = 0; i < kProfSections; i++) acc[i] = 0;
        t = clock64();
    }
    __device__ void mark(int s) {
        const long long n = clock64();
        acc[s] += (unsigned long long)(n - t);
        t = n;
    }
    __device__ void count(int s) { acc[s]++; }
    __device__ void add(int s, unsigned long long v) { acc[s] += v; }
    __device__ void flush(int lane) {
        if (lane == 0)
            for (int i = 0; i < kProfSections; i++) atomicAdd(&g_profile[i], acc[i]);
    }
};
#else
struct Prof {
    __device__ void start() {}
    __device__ void mark(int) {}
    __device__ void count(int) {}
    __device__ void add(int, unsigned long long) {}
    __device__ void flush(int) {}
};
#endif

// parlaylib hash64_2 (cuda/utiles_gpu.cuh:12-17): slot of ParlayANN's seen filter.
__device__ __forceinline__ uint64_t hash64_2(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// parlaylib hash64 (cuda/utiles_gpu.cuh:21-31): HNSW marks its entry points with it.
__device__ __forceinline__ uint64_t hash64(uint64_t u) {
    uint64_t v = u * 3935559000370003845ull + 2691343689449507681ull;
    v ^= v >> 21;
    v ^= v << 37;
    v ^= v >> 4;
    v *= 4768777513237032717ull;
    v ^= v << 20;
    v ^= v >> 41;
    v ^= v << 5;
    return v;
}

__device__ __forceinline__ int laneId() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0));
}

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shflXor64(uint64_t v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shflUp64(uint64_t v, int d) {
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, d, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shflDown64(uint64_t v, int d) {
    const uint32_t lo = (uint32_t)__shfl_down((int)(uint32_t)v, d, 64);
    const uint32_t hi = (uint32_t)__shfl_down((int)(uint32_t)(v >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
}

// DPP moves inside 8-lane groups.
__device__ __forceinline__ int dppHalfMirror(int x) {  // lane i <- lane 7-i
    return __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false);
}
__device__ __forceinline__ int dppXor2(int x) {  // quad_perm [2,3,0,1]
    return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
}
__device__ __forceinline__ int dppXor1(int x) {  // quad_perm [1,0,3,2]
    return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
}

// Sum of x over the 8-lane group, in every lane of the group.
__device__ __forceinline__ int groupSum8(int x) {
    x += dppXor1(x);
    x += dppXor2(x);
    x += dppHalfMirror(x);
    return x;
}

// v[r] holds the lane's partial for round r.  Returns, in lane l, the 8-lane group total
// of round (l & 7): a 3-step transpose-reduce (7 DPP moves instead of 24).
__device__ __forceinline__ int transposeReduce8(const int (&v)[8], int lane) {
    const bool b2 = lane & 4, b1 = lane & 2, b0 = lane & 1;
    int w[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int send = b2 ? v[j] : v[j + 4];
        const int keep = b2 ? v[j + 4] : v[j];
        w[j] = keep + dppHalfMirror(send);
    }
    int x[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int send = b1 ? w[j] : w[j + 2];
        const int keep = b1 ? w[j + 2] : w[j];
        x[j] = keep + dppXor2(send);
    }
    const int send = b0 ? x[0] : x[1];
    const int keep = b0 ? x[1] : x[0];
    return keep + dppXor1(send);
}

__device__ __forceinline__ int sdot4(int a, int b, int c) {
    return __builtin_amdgcn_sdot4(a, b, c, false);
}

// p.p - 2 q.p over one 16-byte slice (exact int32).
__device__ __forceinline__ int slicePartial(const int4& v, const int4& q) {
    int pp = sdot4(v.x, v.x, 0);
    pp = sdot4(v.y, v.y, pp);
    pp = sdot4(v.z, v.z, pp);
    pp = sdot4(v.w, v.w, pp);
    int qp = sdot4(v.x, q.x, 0);
    qp = sdot4(v.y, q.y, qp);
    qp = sdot4(v.z, q.z, qp);
    qp = sdot4(v.w, q.w, qp);
    return pp - 2 * qp;
}

// With rows that carry their squared norm in bytes 124..127 (the last slice, held by
// lane 7 of each 8-lane group) and a query whose bytes 124..127 are zero, the partial of
// one slice is (p.p from the row, in lane 7) - 2 q.p: normMask is -1 in lane 7, else 0.
__device__ __forceinline__ int slicePartialN(const int4& v, const int4& q, int normMask) {
    int qp = sdot4(v.x, q.x, 0);
    qp = sdot4(v.y, q.y, qp);
    qp = sdot4(v.z, q.z, qp);
    qp = sdot4(v.w, q.w, qp);
    return (v.w & normMask) - 2 * qp;
}

__device__ __forceinline__ int sliceNorm(const int4& q) {
    int s = sdot4(q.x, q.x, 0);
    s = sdot4(q.y, q.y, s);
    s = sdot4(q.z, q.z, s);
    return sdot4(q.w, q.w, s);
}

// Distance from the query slice (8 lanes x 16 B) to one row, in every lane of the group.
// normMask: -1 in lane 7 when rows carry their norm in bytes 124..127 (q's are zero).
__device__ __forceinline__ int groupDistance(const int8_t* rows, uint32_t id, int sub,
                                             const int4& qv, int qq, int normMask) {
    int4 v = *reinterpret_cast<const int4*>(rows + (size_t)id * kVecBytes + sub * 16);
    v.w &= ~normMask;
    return qq + groupSum8(slicePartial(v, qv));
}

// DPP move with bound_ctrl set: lanes whose source falls outside the row read 0, so no
// "old" value has to be materialised.
template <int CTRL>
__device__ __forceinline__ uint32_t dppMov(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}

// Value of lane (lane ^ j) with VALU cross-lane moves only (no LDS round trip):
// j = 1, 2 quad permutes; 4 row shifts; 8 row rotate; 16 / 32 the gfx950 permlane swaps.
__device__ __forceinline__ uint32_t xorLane32(uint32_t x, int j, int lane) {
    switch (j) {
        case 1: return dppMov<0xB1>(x);
        case 2: return dppMov<0x4E>(x);
        case 4: {
            // both moves run with every lane active: a DPP read of a lane that is off
            // in EXEC yields 0, so they must not sit inside the lane-dependent select
            const uint32_t up = dppMov<0x104>(x);
            const uint32_t dn = dppMov<0x114>(x);
            return (lane & 4) ? dn : up;
        }
        case 8: return dppMov<0x128>(x);
        case 16: {
            const auto p = __builtin_amdgcn_permlane16_swap(x, x, false, false);
            return (lane & 16) ? p[0] : p[1];
        }
        default: {
            const auto p = __builtin_amdgcn_permlane32_swap(x, x, false, false);
            return (lane & 32) ? p[0] : p[1];
        }
    }
}

__device__ __forceinline__ uint64_t xorLane64(uint64_t v, int j, int lane) {
    const uint32_t lo = xorLane32((uint32_t)v, j, lane);
    const uint32_t hi = xorLane32((uint32_t)(v >> 32), j, lane);
    return ((uint64_t)hi << 32) | lo;
}

// Value of lane - 1 (lane 0 gets `first`), a whole-wave DPP shift.
__device__ __forceinline__ uint64_t prevLane64(uint64_t v, uint64_t first, int lane) {
    const uint32_t lo = dppMov<0x138>((uint32_t)v);
    const uint32_t hi = dppMov<0x138>((uint32_t)(v >> 32));
    return lane == 0 ? first : (((uint64_t)hi << 32) | lo);
}

// Value of lane + 1 (lane 63 gets `last`), a whole-wave DPP shift.
__device__ __forceinline__ uint64_t nextLane64(uint64_t v, uint64_t last, int lane) {
    const uint32_t lo = dppMov<0x130>((uint32_t)v);
    const uint32_t hi = dppMov<0x130>((uint32_t)(v >> 32));
    return lane == 63 ? last : (((uint64_t)hi << 32) | lo);
}

// Value of lane 8 * (lane / 8) + r (ds_swizzle bit mode: and 0x18, or r).
__device__ __forceinline__ uint32_t groupBroadcast8(uint32_t x, int r) {
    switch (r) {
        case 0: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (0 << 5));
        case 1: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (1 << 5));
        case 2: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (2 << 5));
        case 3: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (3 << 5));
        case 4: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (4 << 5));
        case 5: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (5 << 5));
        case 6: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (6 << 5));
        default: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x18 | (7 << 5));
    }
}

// One compare-exchange stage at distance j < 64 inside a register: the lane keeps the
// smaller of itself and its partner iff keepMin (one 64-bit compare).
__device__ __forceinline__ uint64_t cmpXchg(uint64_t x, int j, bool keepMin, int lane) {
    const uint64_t p = xorLane64(x, j, lane);
    return ((x < p) == keepMin) ? x : p;
}

// Ascending bitonic sort of 64*C keys, element e = c*64 + lane in key[c].
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
                        const int e = c * 64 + lane;
                        const bool asc = (e & k) == 0;
                        const uint64_t a = key[c], b = key[c2];
                        const bool sw = asc ? (a > b) : (a < b);
                        key[c] = sw ? b : a;
                        key[c2] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int e = c * 64 + lane;
                    const bool asc = (e & k) == 0;
                    key[c] = cmpXchg(key[c], j, ((lane & j) == 0) == asc, lane);
                }
            }
        }
    }
}

// Turns two ascending-sorted registers' worth of a bitonic sequence (lo = elements 0..63,
// hi = 64..127, lo ascending and hi descending) into one ascending run of 128.
__device__ __forceinline__ void bitonicMerge128(uint64_t& lo, uint64_t& hi, int lane) {
    const uint64_t mn = lo < hi ? lo : hi;
    const uint64_t mx = lo < hi ? hi : lo;
    lo = mn;
    hi = mx;
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) {
        const bool keepMin = (lane & j) == 0;
        lo = cmpXchg(lo, j, keepMin, lane);
        hi = cmpXchg(hi, j, keepMin, lane);
    }
}

// Seen-filter accesses: LDS filters use plain ds ops; a per-wave HBM filter is read and
// written at agent scope so the wave's own earlier writes are visible to it (the vector
// L1 is not coherent with stores).
template <bool G>
__device__ __forceinline__ uint32_t filtLoad(const uint32_t* f, uint32_t h) {
    if constexpr (G) return __hip_atomic_load(f + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return f[h];
}
template <bool G>
__device__ __forceinline__ void filtStore(uint32_t* f, uint32_t h, uint32_t v) {
    if constexpr (G) __hip_atomic_store(f + h, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else f[h] = v;
}

// Clears 2^bits filter slots to kEmpty.
template <bool G>
__device__ __forceinline__ void filtClear(uint32_t* f, int bits, int lane) {
    const uint32_t slots = 1u << bits;
    if constexpr (G) {
        for (uint32_t i = lane; i < slots; i += 64) filtStore<true>(f, i, kEmpty);
        __builtin_amdgcn_s_waitcnt(0);
    } else {
        if (slots >= 4) {
            const uint4 e4 = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
            uint4* f4 = reinterpret_cast<uint4*>(f);
            for (uint32_t i = lane; i < slots / 4; i += 64) f4[i] = e4;
        } else if ((uint32_t)lane < slots) {
            f[lane] = kEmpty;
        }
    }
}

// ParlayANN's lossy seen filter over one 64-entry chunk of an adjacency row, with the
// sequential semantics of has_been_seen: entry a (participating) is "seen" iff the slot
// held a just before its turn, then the slot holds a.  Slot reads happen for the whole
// chunk at once; lanes whose byte tag was overwritten share a slot (or a tag) with
// another lane and are resolved by a scalar loop over those groups.  Returns keep.
//
// The work is split in two so that callers can issue loads between the halves:
// seenFilterBegin reads slots and tags; seenFilterEnd resolves the lanes that share a
// slot and writes the slots.  A lane alone in its slot is kept iff part && prev != a; a
// lane sharing a slot may be kept although prev == a (an earlier lane of the row wrote
// another id there first), and only the losers of the tag race see coll.
struct SeenProbe {
    uint32_t h;     // filter slot
    uint32_t prev;  // slot value before this chunk
    bool coll;      // shares a slot (or its byte tag) with another lane of the chunk
};

template <bool G>
__device__ __forceinline__ SeenProbe seenFilterBegin(const uint32_t* filt, unsigned char* tag,
                                                     uint32_t fmask, uint32_t a, bool part,
                                                     int lane) {
    SeenProbe s;
    s.h = (uint32_t)hash64_2(a) & fmask;
    const uint32_t ts = s.h & (kTagSlots - 1);
    if (part) tag[ts] = (unsigned char)lane;
    __builtin_amdgcn_wave_barrier();
    s.prev = kEmpty;
    s.coll = false;
    if (part) {
        s.coll = tag[ts] != (unsigned char)lane;
        s.prev = filtLoad<G>(filt, s.h);
    }
    return s;
}

template <bool G>
__device__ __forceinline__ bool seenFilterEnd(uint32_t* filt, const SeenProbe& s, uint32_t a,
                                              bool part, int lane, uint64_t lanesBelow) {
    const uint32_t h = s.h;
    uint32_t prev = s.prev;
    uint64_t cm = ballot(s.coll);
    bool last = true;
    while (cm) {
        const int Lr = __ffsll((unsigned long long)cm) - 1;
        const uint32_t hL = (uint32_t)readlane((int)h, Lr);
        const uint64_t M = ballot(part && h == hL);
        if (__popcll(M) == 2) {
            // the usual group, two lanes: the upper one sees the lower one's id
            const int lo = __ffsll((unsigned long long)M) - 1;
            const int hi = 63 - __clzll((long long)M);
            const uint32_t aLo = (uint32_t)readlane((int)a, lo);
            if (lane == hi) prev = aLo;
            if (lane == lo) last = false;
        } else {
            const bool inM = (M >> lane) & 1ull;
            const uint64_t below = M & lanesBelow;
            const uint64_t above = M & ~(lanesBelow | (1ull << lane));
            const int src = (inM && below) ? (63 - __clzll((long long)below)) : lane;
            const uint32_t x = (uint32_t)__shfl((int)a, src, 64);
            if (inM && below) prev = x;
            if (inM && above) last = false;
        }
        cm &= ~M;
    }
    const bool seen = part && (prev == a);
    // an HBM slot that already holds a is not written again (most entries of a long
    // beam's rows were seen before, and each write is a partial-line HBM write).  The
    // wave does not wait for its HBM stores here: its own later loads of a slot come after
    // them in its memory request stream (one wave's accesses to one address keep their
    // order, as for any plain store and load), and the next wait on a later load covers
    // them.
    if (part && last && (!G || s.prev != a)) filtStore<G>(filt, h, a);
    __builtin_amdgcn_wave_barrier();
    return part && !seen;
}

template <bool G>
__device__ __forceinline__ bool seenFilterChunk(uint32_t* filt, unsigned char* tag, uint32_t fmask,
                                                uint32_t a, bool part, int lane,
                                                uint64_t lanesBelow) {
    const SeenProbe s = seenFilterBegin<G>(filt, tag, fmask, a, part, lane);
    return seenFilterEnd<G>(filt, s, a, part, lane, lanesBelow);
}

// A query whose parameters fall outside what the launch was sized for: every output
// column holds the missing-neighbour sentinel, the counters are zero and the flag says why.
template <class Launch>
__device__ __forceinline__ void rejectQuery(const Launch& L, uint32_t q, int lane) {
    uint32_t* oi = L.outIds + (size_t)q * L.outStride;
    int32_t* od = L.outDists + (size_t)q * L.outStride;
    for (uint32_t j = lane; j < L.outStride; j += 64) {
        oi[j] = kEmpty;
        od[j] = 0x7FFFFFFF;
    }
    if (lane == 0) {
        if (L.outComps) L.outComps[q] = 0;
        if (L.outFlags) L.outFlags[q] = kFlagConfig;
        if (L.outEdges) L.outEdges[q] = 0;
        if (L.outHops) L.outHops[q] = 0;
    }
}

// What takeQuery returns once the queue is drained.  The order array may hold a range of
// a larger batch, so a query index says nothing about the queue's end.
constexpr uint32_t kQueueEmpty = 0xFFFFFFFFu;

// Persistent work queue: queue[0] is the next ticket, queue[1] counts finished waves.
// Tickets below nq map to queries through `order` (NULL: identity).
__device__ __forceinline__ uint32_t takeQuery(uint32_t* queue, const uint32_t* order, uint32_t nq,
                                              int lane) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(&queue[0], 1u);
    t = (uint32_t)readlane((int)t, 0);
    if (t >= nq) return kQueueEmpty;
    return order ? order[t] : t;
}

// The last wave to finish resets the queue, so the same launch (or a captured graph of
// it) can run again without a host-side clear.
__device__ __forceinline__ void finishQueue(uint32_t* queue, uint32_t totalWaves, int lane) {
    if (lane == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&queue[1], 1u);
        if (done + 1 == totalWaves) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

}  // namespace dev
}  // namespace pann
