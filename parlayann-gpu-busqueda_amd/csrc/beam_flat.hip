// beam_flat.hip -- ParlayANN flat-graph beam search (Vamana / HCNNG / PyNNDescent) for
// gfx950.  One 64-lane wavefront owns one query; all per-query state lives in LDS.
//
// Semantics follow ParlayANN's beam_search as restated by the reference kernel
// (cuda/busqueda.cu:95-506): start at node 0, lossy seen filter written in row order,
// neighbours equal to the query index skipped, multiset set_union of frontier and
// candidates, trim to Ls, cut at (float)(cut * (double)F[k].d), next = first frontier
// copy not yet visited.  The data movement is this file's own:
//   * adjacency row: one coalesced dword per lane (64 entries per instruction);
//   * seen filter: slot reads for the whole row at once, in-row slot collisions found
//     with a byte tag table and resolved with a scalar loop over the (rare) groups;
//   * distances: 8 lanes x 16 B per 128-B row, 8 rows per load instruction,
//     v_dot4_i32_i8 for q.p and p.p (no norm array), DPP transpose-reduce;
//   * candidates: in-register bitonic sort of (d << 32 | id) keys;
//   * merge: every merged position computes its source with a ballot rank (no shifts).
#include "pann_internal.h"

namespace pann {
namespace {

__device__ __forceinline__ uint64_t hash64_2(uint64_t x) {
    // parlaylib hash64_2 (cuda/utiles_gpu.cuh:12-17)
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ int laneId() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shflXor64(uint64_t v, int m) {
    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
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

// v[r] holds lane's partial for round r.  Returns, in lane l, the 8-lane group total of
// round (l & 7): a 3-step transpose-reduce (7 DPP moves instead of 24).
__device__ __forceinline__ int transposeReduce8(const int (&v)[8], int lane) {
    const bool b2 = lane & 4, b1 = lane & 2, b0 = lane & 1;
    int w[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int send = b2 ? v[j] : v[j + 4];
        int keep = b2 ? v[j + 4] : v[j];
        w[j] = keep + dppHalfMirror(send);
    }
    int x[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int send = b1 ? w[j] : w[j + 2];
        int keep = b1 ? w[j + 2] : w[j];
        x[j] = keep + dppXor2(send);
    }
    int send = b0 ? x[0] : x[1];
    int keep = b0 ? x[1] : x[0];
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
                        uint64_t a = key[c], b = key[c2];
                        bool sw = asc ? (a > b) : (a < b);
                        key[c] = sw ? b : a;
                        key[c2] = sw ? a : b;
                    }
                }
            } else {
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int e = c * 64 + lane;
                    const bool asc = (e & k) == 0;
                    const bool lower = (lane & j) == 0;
                    uint64_t p = shflXor64(key[c], j);
                    uint64_t mn = key[c] < p ? key[c] : p;
                    uint64_t mx = key[c] < p ? p : key[c];
                    key[c] = (lower == asc) ? mn : mx;
                }
            }
        }
    }
}

__device__ __forceinline__ uint32_t visitedSlot(uint32_t id, uint32_t logCap) {
    return (id * 2654435761u) >> (32 - logCap);
}

// Distance from the query slice to one base row, summed over the 8-lane group.
__device__ __forceinline__ int groupDistance(const int8_t* base, uint32_t id, int sub,
                                             const int4& qv, int qq) {
    const int4 v = *reinterpret_cast<const int4*>(base + (size_t)id * kVecBytes + sub * 16);
    return qq + groupSum8(slicePartial(v, qv));
}

// Seen-filter accesses.  An LDS filter uses plain ds ops; a per-wave HBM filter is read
// and written at agent scope so the wave's own earlier writes are visible to it (the
// vector L1 is not coherent with its stores).
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

struct QueryCfg {
    uint32_t qid;
    int Ls, k;
    int64_t limit;
    double cut;
};

template <int MAXCH, bool FGLOBAL>
__device__ void searchOne(const FlatLaunch& L, unsigned char* lds, uint32_t q, const QueryCfg& c,
                          uint32_t* gfilt) {
    const int lane = laneId();
    const int grp = lane >> 3, sub = lane & 7;
    const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    uint64_t* fkey[2] = {reinterpret_cast<uint64_t*>(lds + L.lds.fkey[0]),
                         reinterpret_cast<uint64_t*>(lds + L.lds.fkey[1])};
    unsigned char* fvis[2] = {lds + L.lds.fvis[0], lds + L.lds.fvis[1]};
    unsigned char* occ = lds + L.lds.occ;
    uint64_t* kkey = reinterpret_cast<uint64_t*>(lds + L.lds.kkey);
    unsigned char* kvis = lds + L.lds.kvis;
    uint32_t* pend = reinterpret_cast<uint32_t*>(lds + L.lds.pend);
    unsigned char* tag = lds + L.lds.tag;
    uint32_t* vkey = reinterpret_cast<uint32_t*>(lds + L.lds.vkey);
    unsigned char* vcnt = lds + L.lds.vcnt;
    uint32_t* filt = FGLOBAL ? gfilt : reinterpret_cast<uint32_t*>(lds + L.lds.filt);

    const int Ls = c.Ls;
    const int fbits = flatFilterBits(Ls);
    const uint32_t fmask = (1u << fbits) - 1;
    // visited capacity for this query (never above what the launch holds)
    uint32_t capV = flatVisitedCap(Ls);
    if (capV > L.capV) capV = L.capV;
    const uint32_t logCap = (uint32_t)ceilLog2(capV);
    const uint32_t vmask = capV - 1;

    // ---- query slice: lane holds bytes [16*sub, 16*sub+16) ----
    const int4 qv = *reinterpret_cast<const int4*>(L.queries + (size_t)q * kVecBytes + sub * 16);
    int qq = sdot4(qv.x, qv.x, 0);
    qq = sdot4(qv.y, qv.y, qq);
    qq = sdot4(qv.z, qv.z, qq);
    qq = sdot4(qv.w, qv.w, qq);
    qq = groupSum8(qq);

    // ---- fresh tables for this query (the CPU creates them per query) ----
    {
        uint4 e4 = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
        if constexpr (FGLOBAL) {
            for (uint32_t i = lane; i < (1u << fbits); i += 64) filtStore<true>(filt, i, kEmpty);
            __builtin_amdgcn_s_waitcnt(0);
        } else {
            uint4* f4 = reinterpret_cast<uint4*>(filt);
            for (uint32_t i = lane; i < (1u << fbits) / 4; i += 64) f4[i] = e4;
        }
        uint4* v4 = reinterpret_cast<uint4*>(vkey);
        for (uint32_t i = lane; i < capV / 4; i += 64) v4[i] = e4;
        for (int i = lane; i < Ls; i += 64) occ[i] = 0;
    }

    // ---- start point: node 0 (cuda/busqueda.cu:166-173) ----
    const int d0 = groupDistance(L.base, 0u, sub, qv, qq);
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        fkey[0][0] = (uint64_t)(uint32_t)d0 << 32;
        fvis[0][0] = 0;
    }
    __builtin_amdgcn_wave_barrier();

    int cur = 0;
    int fsize = 1;
    int u = 0;
    int64_t visits = 0;
    int comps = 1, edges = 0, hops = 0, flags = 0;
    uint32_t distinctVisited = 0;
    const uint32_t G = L.G;
    const uint32_t nch = (G + 63) / 64;

    while (u >= 0 && visits < c.limit) {
        const uint64_t ukey = fkey[cur][u];
        const uint32_t uid = (uint32_t)ukey;

        // ---- adjacency row, one coalesced dword per lane per 64 entries ----
        uint32_t a[MAXCH];
        const uint32_t* row = L.adj + (size_t)uid * G;
#pragma unroll
        for (int ch = 0; ch < MAXCH; ch++) {
            const uint32_t j = ch * 64 + lane;
            a[ch] = (ch < (int)nch && j < G) ? row[j] : kEmpty;
        }

        // ---- visit u: frontier flag and visited multiset count ----
        bool overflow = false;
        if (lane == 0) {
            fvis[cur][u] = 1;
            uint32_t h = visitedSlot(uid, logCap);
            while (vkey[h] != kEmpty && vkey[h] != uid) h = (h + 1) & vmask;
            if (vkey[h] == uid) {
                vcnt[h] = (unsigned char)(vcnt[h] + 1);
            } else if (distinctVisited + 1 < capV) {
                vkey[h] = uid;
                vcnt[h] = 1;
                distinctVisited++;
            } else {
                overflow = true;
            }
        }
        distinctVisited = (uint32_t)readlane((int)distinctVisited, 0);
        if (readlane((int)overflow, 0)) {
            flags |= kFlagOverflow;
            break;
        }
        visits++;
        hops++;

        const int cutoff = (fsize < Ls) ? 0x7FFFFFFF : (int)(fkey[cur][fsize - 1] >> 32);

        // ---- lossy seen filter in row order (ParlayANN has_been_seen) ----
        uint32_t nP = 0;
#pragma unroll
        for (int ch = 0; ch < MAXCH; ch++) {
            if (ch >= (int)nch) break;
            const uint32_t ai = a[ch];
            edges += __popcll(ballot(ai != kEmpty));
            const bool part = (ai != kEmpty) && (ai != c.qid);
            const uint32_t h = (uint32_t)hash64_2(ai) & fmask;
            const uint32_t ts = h & (kTagSlots - 1);
            if (part) tag[ts] = (unsigned char)lane;
            __builtin_amdgcn_wave_barrier();
            uint32_t prev = kEmpty;
            bool coll = false;
            if (part) {
                coll = tag[ts] != (unsigned char)lane;
                prev = filtLoad<FGLOBAL>(filt, h);
            }
            uint64_t cm = ballot(coll);
            bool last = true;
            while (cm) {
                const int Lr = __ffsll((unsigned long long)cm) - 1;
                const uint32_t hL = (uint32_t)readlane((int)h, Lr);
                const uint64_t M = ballot(part && h == hL);
                const bool inM = (M >> lane) & 1ull;
                const uint64_t below = M & lanesBelow;
                const uint64_t above = M & ~(lanesBelow | (1ull << lane));
                const int src = (inM && below) ? (63 - __clzll((long long)below)) : lane;
                const uint32_t x = (uint32_t)__shfl((int)ai, src, 64);
                if (inM && below) prev = x;
                if (inM && above) last = false;
                cm &= ~M;
            }
            const bool seen = part && (prev == ai);
            if (part && last) filtStore<FGLOBAL>(filt, h, ai);
            if constexpr (FGLOBAL) __builtin_amdgcn_s_waitcnt(0);
            const bool keep = part && !seen;
            const uint64_t km = ballot(keep);
            if (keep) pend[nP + mbcnt(km)] = ai;
            nP += (uint32_t)__popcll(km);
            __builtin_amdgcn_wave_barrier();
        }
        comps += (int)nP;

        // ---- distances: 8 rows per round, 8 lanes x 16 B each ----
        uint64_t key[MAXCH];
#pragma unroll
        for (int b = 0; b < MAXCH; b++) {
            key[b] = ~0ull;
            if ((uint32_t)b * 64 >= nP) continue;
            const uint32_t nb = min(64u, nP - (uint32_t)b * 64);
            const uint32_t rounds = (nb + 7) / 8;
            int4 v[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                v[r] = make_int4(0, 0, 0, 0);
                const uint32_t i = (uint32_t)r * 8 + grp;
                if ((uint32_t)r < rounds && i < nb) {
                    const uint32_t id = pend[b * 64 + i];
                    v[r] = *reinterpret_cast<const int4*>(L.base + (size_t)id * kVecBytes + sub * 16);
                }
            }
            int part8[8];
#pragma unroll
            for (int r = 0; r < 8; r++) part8[r] = slicePartial(v[r], qv);
            const int tot = transposeReduce8(part8, lane);
            const uint32_t il = (uint32_t)sub * 8 + grp;  // candidate held by this lane
            if (il < nb) {
                const int d = qq + tot;
                const uint32_t id = pend[b * 64 + il];
                if (d < cutoff) key[b] = ((uint64_t)(uint32_t)d << 32) | id;
            }
        }

        // ---- sort candidates by (d, id) ----
        bitonicSort<MAXCH>(key, lane);

        // ---- multiset union with the frontier ----
        const uint64_t* F = fkey[cur];
        uint64_t runStarts[MAXCH];
#pragma unroll
        for (int b = 0; b < MAXCH; b++) {
            uint64_t up = (uint64_t)__shfl_up((int)(uint32_t)key[b], 1, 64) |
                          ((uint64_t)(uint32_t)__shfl_up((int)(uint32_t)(key[b] >> 32), 1, 64) << 32);
            if (lane == 0) up = (b == 0) ? ~key[b] : readlane64(key[b - 1], 63);
            runStarts[b] = ballot(up != key[b]);
        }
        uint32_t kBase = 0;
        int jPos[MAXCH];
        bool enter[MAXCH];
#pragma unroll
        for (int b = 0; b < MAXCH; b++) {
            const uint64_t kv = key[b];
            const bool valid = kv != ~0ull;
            int rep = 0, lb = 0, cF = 0;
            if (valid) {
                const int e = b * 64 + lane;
                uint64_t m = runStarts[b] & (lanesBelow | (1ull << lane));
                int start;
                if (m) {
                    start = b * 64 + 63 - __clzll((long long)m);
                } else {
                    start = 0;
                    for (int bb = b - 1; bb >= 0; bb--)
                        if (runStarts[bb]) {
                            start = bb * 64 + 63 - __clzll((long long)runStarts[bb]);
                            break;
                        }
                }
                rep = e - start;
                int lo = 0, hi = fsize;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (F[mid] < kv) lo = mid + 1;
                    else hi = mid;
                }
                lb = lo;
                while (lb + cF < fsize && F[lb + cF] == kv) cF++;
            }
            enter[b] = valid && rep >= cF;
            const uint64_t em = ballot(enter[b]);
            const uint32_t jK = kBase + mbcnt(em);
            kBase += (uint32_t)__popcll(em);
            jPos[b] = lb + cF + (int)jK;
            if (enter[b]) {
                // this copy is visited if the id was expanded more than `rep` times
                const uint32_t id = (uint32_t)kv;
                uint32_t h = visitedSlot(id, logCap);
                int cnt = 0;
                while (true) {
                    const uint32_t kk = vkey[h];
                    if (kk == id) { cnt = vcnt[h]; break; }
                    if (kk == kEmpty) break;
                    h = (h + 1) & vmask;
                }
                kkey[jK] = kv;
                kvis[jK] = (unsigned char)(rep < cnt ? 1 : 0);
                if (jPos[b] < Ls) occ[jPos[b]] = 1;
            }
        }
        const int nK = (int)kBase;
        __builtin_amdgcn_wave_barrier();

        // ---- merged frontier: each position finds its source by rank ----
        const int nxt = cur ^ 1;
        const int newSize = min(Ls, fsize + nK);
        {
            uint32_t rankBase = 0;
            for (int p0 = 0; p0 < newSize; p0 += 64) {
                const int p = p0 + lane;
                const bool inb = p < newSize;
                const bool o = inb && occ[p];
                const uint64_t ob = ballot(o);
                const uint32_t rank = rankBase + mbcnt(ob);
                if (inb) {
                    uint64_t nk;
                    unsigned char nv;
                    if (o) {
                        nk = kkey[rank];
                        nv = kvis[rank];
                    } else {
                        nk = F[p - rank];
                        nv = fvis[cur][p - rank];
                    }
                    fkey[nxt][p] = nk;
                    fvis[nxt][p] = nv;
                    occ[p] = 0;
                }
                rankBase += (uint32_t)__popcll(ob);
            }
        }
        __builtin_amdgcn_wave_barrier();
        cur = nxt;
        fsize = newSize;

        // ---- cut: upper_bound(F, (0, (float)(cut * (double)F[k].d))) ----
        if (c.k > 0 && fsize > c.k) {
            const int dk = (int)(fkey[cur][c.k] >> 32);
            const float t = (float)(c.cut * (double)dk);
            int nf = fsize;
            for (int p0 = 0; p0 < fsize; p0 += 64) {
                const int p = p0 + lane;
                bool pred = false;
                if (p < fsize) {
                    const uint64_t kv = fkey[cur][p];
                    const float df = (float)(int)(kv >> 32);
                    pred = (t < df) || (t == df && (uint32_t)kv > 0u);
                }
                const uint64_t m = ballot(pred);
                if (m) {
                    nf = p0 + __ffsll((unsigned long long)m) - 1;
                    break;
                }
            }
            fsize = nf;
        }

        // ---- next: first frontier copy not visited ----
        u = -1;
        for (int p0 = 0; p0 < fsize; p0 += 64) {
            const int p = p0 + lane;
            const bool nv = (p < fsize) && !fvis[cur][p];
            const uint64_t m = ballot(nv);
            if (m) {
                u = p0 + __ffsll((unsigned long long)m) - 1;
                break;
            }
        }
    }

    // ---- output: first k of the frontier (cuda/busqueda.cu:487-497) ----
    if (fsize < c.k) flags |= kFlagShort;
    uint32_t* oi = L.outIds + (size_t)q * L.outStride;
    int32_t* od = L.outDists + (size_t)q * L.outStride;
    for (int j = lane; j < c.k; j += 64) {
        if (j < fsize) {
            const uint64_t kv = fkey[cur][j];
            oi[j] = (uint32_t)kv;
            od[j] = (int32_t)(kv >> 32);
        } else {
            oi[j] = kEmpty;
            od[j] = 0x7FFFFFFF;
        }
    }
    if (lane == 0) {
        if (L.outComps) L.outComps[q] = comps;
        if (L.outFlags) L.outFlags[q] = flags;
        if (L.outEdges) L.outEdges[q] = edges;
        if (L.outHops) L.outHops[q] = hops;
    }
    __builtin_amdgcn_wave_barrier();
}

template <int MAXCH>
__global__ __launch_bounds__(256) void beamFlatKernel(FlatLaunch L) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = laneId();
    const int waveInBlock = threadIdx.x >> 6;
    const uint32_t wavesPerBlock = blockDim.x >> 6;
    const uint32_t waveGlobal = blockIdx.x * wavesPerBlock + waveInBlock;
    unsigned char* lds = smem + (size_t)waveInBlock * L.ldsPerWave;
    uint32_t* gfilt = L.gfilt ? L.gfilt + ((size_t)waveGlobal << L.gbitsMax) : nullptr;

    uint32_t q = waveGlobal;
    if (L.queue) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&L.queue[0], 1u);
        q = (uint32_t)readlane((int)t, 0);
    }
    while (q < L.nq) {
        QueryCfg c;
        if (L.qp) {
            const QueryParams& P = L.qp[q];
            c.qid = P.qid;
            c.Ls = P.Ls;
            c.k = P.k;
            c.limit = P.visit_limit > 0 ? P.visit_limit : (int64_t)L.n;
            c.cut = P.cut;
        } else {
            c.qid = L.qids ? L.qids[q] : q;
            c.Ls = L.Ls;
            c.k = L.k;
            c.limit = L.limit;
            c.cut = L.cut;
        }
        const int fb = flatFilterBits(c.Ls);
        if (c.Ls < 1 || c.Ls > L.LsMax || c.k < 0 || (uint32_t)c.k > L.outStride ||
            (fb > L.fbitsLds && (L.gfilt == nullptr || fb > L.gbitsMax))) {
            if (lane == 0 && L.outFlags) L.outFlags[q] = kFlagConfig;
        } else if (fb <= L.fbitsLds) {
            searchOne<MAXCH, false>(L, lds, q, c, nullptr);
        } else {
            searchOne<MAXCH, true>(L, lds, q, c, gfilt);
        }
        if (L.queue) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(&L.queue[0], 1u);
            q = (uint32_t)readlane((int)t, 0);
        } else {
            q += L.totalWaves;
        }
    }
    if (L.queue && lane == 0) {
        // last wave out resets the queue so the launch can be replayed (graph capture)
        __threadfence();
        const uint32_t done = atomicAdd(&L.queue[1], 1u);
        if (done + 1 == L.totalWaves) {
            atomicExch(&L.queue[0], 0u);
            atomicExch(&L.queue[1], 0u);
        }
    }
}

}  // namespace

hipError_t launchBeamFlat(const FlatLaunch& L, int maxch, int wavesPerBlock, int blocks,
                          hipStream_t stream) {
    const size_t lds = (size_t)L.ldsPerWave * wavesPerBlock;
    const dim3 grid(blocks), block(64 * wavesPerBlock);
    switch (maxch) {
        case 1: hipLaunchKernelGGL(beamFlatKernel<1>, grid, block, lds, stream, L); break;
        case 2: hipLaunchKernelGGL(beamFlatKernel<2>, grid, block, lds, stream, L); break;
        case 4: hipLaunchKernelGGL(beamFlatKernel<4>, grid, block, lds, stream, L); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t beamFlatOccupancy(int maxch, int wavesPerBlock, size_t ldsBytes, int* blocksPerCU) {
    const int threads = 64 * wavesPerBlock;
    switch (maxch) {
        case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatKernel<1>, threads, ldsBytes);
        case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatKernel<2>, threads, ldsBytes);
        case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatKernel<4>, threads, ldsBytes);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pann
