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
//   * candidates: in-register bitonic sort of (d << 32 | id) keys, or a rank sort when
//     at most kRankSortMax are left (the usual case of long beams);
//   * merge: every merged position computes its source with a ballot rank (no shifts).
#include "device_common.h"

namespace pann {
namespace {

using namespace dev;

__device__ __forceinline__ uint32_t visitedSlot(uint32_t id, uint32_t logCap) {
    return (id * 2654435761u) >> (32 - logCap);
}

// Visited multiset of the general form: capV id slots (kEmpty: free) and their expansion
// counts (u8), in LDS or in a per-wave HBM region.  In LDS (GROUPS false) an id takes the
// first free slot from its hash slot on, one probe at a time.  In HBM (GROUPS true) it
// hashes to an aligned group of four slots and takes the first free slot from there on,
// group by group, so one 16-B key load and one 4-B count load answer most lookups.
// Inserts stop one short of capV, so every walk meets a free slot.
template <bool GROUPS>
__device__ __forceinline__ uint32_t visitedCount(const uint32_t* vkey, const unsigned char* vcnt,
                                                 uint32_t id, uint32_t logCap, uint32_t vmask) {
    if constexpr (!GROUPS) {
        uint32_t h = visitedSlot(id, logCap);
        while (true) {
            const uint32_t k = vkey[h];
            if (k == id) return vcnt[h];
            if (k == kEmpty) return 0;
            h = (h + 1) & vmask;
        }
    }
    uint32_t g = visitedSlot(id, logCap) & ~3u;
    while (true) {
        const uint4 k4 = *reinterpret_cast<const uint4*>(vkey + g);
        const uint32_t c4 = *reinterpret_cast<const uint32_t*>(vcnt + g);
        const uint32_t kk[4] = {k4.x, k4.y, k4.z, k4.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (kk[j] == id) return (c4 >> (8 * j)) & 0xFFu;
            if (kk[j] == kEmpty) return 0;
        }
        g = (g + 4) & vmask;
    }
}

// One more expansion of id (one lane).  Returns false when a new id would fill the table.
template <bool GROUPS>
__device__ __forceinline__ bool visitedInsert(uint32_t* vkey, unsigned char* vcnt, uint32_t id,
                                              uint32_t logCap, uint32_t vmask, uint32_t capV,
                                              uint32_t& distinct) {
    if constexpr (!GROUPS) {
        uint32_t h = visitedSlot(id, logCap);
        while (vkey[h] != kEmpty && vkey[h] != id) h = (h + 1) & vmask;
        if (vkey[h] == id) {
            vcnt[h] = (unsigned char)(vcnt[h] + 1);
            return true;
        }
        if (distinct + 1 >= capV) return false;
        vkey[h] = id;
        vcnt[h] = 1;
        distinct++;
        return true;
    }
    uint32_t g = visitedSlot(id, logCap) & ~3u;
    while (true) {
        const uint4 k4 = *reinterpret_cast<const uint4*>(vkey + g);
        const uint32_t kk[4] = {k4.x, k4.y, k4.z, k4.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (kk[j] == id) {
                vcnt[g + j] = (unsigned char)(vcnt[g + j] + 1);
                return true;
            }
            if (kk[j] == kEmpty) {
                if (distinct + 1 >= capV) return false;
                vkey[g + j] = id;
                vcnt[g + j] = 1;
                distinct++;
                return true;
            }
        }
        g = (g + 4) & vmask;
    }
}

struct QueryCfg {
    uint32_t qid;
    int Ls, k;
    int64_t limit;
    double cut;
};

// The parameters are the same in every lane; saying so keeps them in scalar registers
// and keeps branches on them uniform.
__device__ __forceinline__ void makeUniform(QueryCfg& c) {
    c.qid = (uint32_t)__builtin_amdgcn_readfirstlane((int)c.qid);
    c.Ls = __builtin_amdgcn_readfirstlane(c.Ls);
    c.k = __builtin_amdgcn_readfirstlane(c.k);
    const uint64_t lim = (uint64_t)c.limit;
    c.limit = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(lim >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)lim));
    const uint64_t cb = (uint64_t)__double_as_longlong(c.cut);
    c.cut = __longlong_as_double((long long)(
        ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(cb >> 32)) << 32) |
        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)cb)));
}

// Candidate keys carry the expansion count of their id in the low bits,
//     d (23 bits) | id (32 bits) | count (9 bits, saturated),
// so the count is looked up next to the candidate's vector load and travels with it
// through the sort (equal (d, id) means equal count, so the order by (d, id) stands).
// Frontier keys are (d << 32 | id).
__device__ __forceinline__ uint64_t countKey(int d, uint32_t id, uint32_t cnt) {
    return ((uint64_t)(uint32_t)d << 41) | ((uint64_t)id << 9) | (cnt < 511u ? cnt : 511u);
}
__device__ __forceinline__ uint64_t frontierOfCountKey(uint64_t k) {
    return ((k >> 41) << 32) | ((k >> 9) & 0xFFFFFFFFull);
}

// VGLOBAL: the visited table lives in this wave's HBM region (gvk / gvc) rather than in
// LDS, for beams whose table would otherwise bound how many wavefronts share a CU.
// Up to this many candidates of the general form are sorted by rank instead of by the
// bitonic network.
constexpr int kRankSortMax = 16;

// Ascending sort of one key per lane when the valid keys (the lanes of vm) are few and
// the others hold ~0: a valid key's rank counts the smaller keys and the equal ones on
// lower lanes, the others follow in lane order; a forward permute of each half moves the
// key of rank r to lane r.
__device__ __forceinline__ uint64_t rankSortFew(uint64_t key, uint64_t vm, int lane, uint64_t lanesBelow) {
    int rank = 0;
    uint64_t m = vm;
    while (m) {
        const int j = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const uint64_t kj = readlane64(key, j);
        rank += (kj < key || (kj == key && j < lane)) ? 1 : 0;
    }
    if (!((vm >> lane) & 1ull)) rank = __popcll(vm) + __popcll(~vm & lanesBelow);
    const int lo = __builtin_amdgcn_ds_permute(rank << 2, (int)(uint32_t)key);
    const int hi = __builtin_amdgcn_ds_permute(rank << 2, (int)(uint32_t)(key >> 32));
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

template <int MAXCH, bool FGLOBAL, bool VGLOBAL>
__device__ void searchOne(const FlatLaunch& L, unsigned char* lds, uint32_t q, const QueryCfg& c,
                          uint32_t* gfilt, uint32_t* gvk, unsigned char* gvc, Prof& P) {
    const int lane = laneId();
    const int grp = lane >> 3, sub = lane & 7;
    const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    // the frontier (keys and visited flags), merged in place
    uint64_t* F = reinterpret_cast<uint64_t*>(lds + L.lds.fkey);
    unsigned char* FV = lds + L.lds.fvis;
    unsigned char* occ = lds + L.lds.occ;
    uint64_t* kkey = reinterpret_cast<uint64_t*>(lds + L.lds.kkey);
    unsigned char* kvis = lds + L.lds.kvis;
    uint32_t* pend = reinterpret_cast<uint32_t*>(lds + L.lds.pend);
    unsigned char* tag = lds + L.lds.tag;
    uint32_t* vkey = VGLOBAL ? gvk : reinterpret_cast<uint32_t*>(lds + L.lds.vkey);
    unsigned char* vcnt = VGLOBAL ? gvc : lds + L.lds.vcnt;
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
    // rows may hold |p|^2 in bytes 124..127 (lane 7's .w): masked out of the products
    const int normMask = (L.normsInRow && sub == 7) ? -1 : 0;
    int4 qv = *reinterpret_cast<const int4*>(L.queries + (size_t)q * kVecBytes + sub * 16);
    qv.w &= ~normMask;
    const int qq = groupSum8(sliceNorm(qv));

    // ---- fresh tables for this query (the CPU creates them per query) ----
    {
        const uint4 e4 = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
        filtClear<FGLOBAL>(filt, fbits, lane);
        uint4* v4 = reinterpret_cast<uint4*>(vkey);
        for (uint32_t i = lane; i < capV / 4; i += 64) v4[i] = e4;
        for (int i = lane; i < Ls; i += 64) occ[i] = 0;
    }

    // ---- start point: node 0 (cuda/busqueda.cu:166-173) ----
    const int d0 = groupDistance(L.base, 0u, sub, qv, qq, normMask);
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        F[0] = (uint64_t)(uint32_t)d0 << 32;
        FV[0] = 0;
    }
    __builtin_amdgcn_wave_barrier();

    int fsize = 1;
    int u = 0;
    int64_t visits = 0;
    int comps = 1, edges = 0, hops = 0, flags = 0;
    uint32_t distinctVisited = 0;
    const uint32_t G = L.G;
    const uint32_t nch = (G + 63) / 64;
    P.mark(kProfInit);
    P.count(kProfQueries);

    uint32_t predId = kEmpty, nextPred = kEmpty;  // one-chunk rows: the node whose row was
    uint32_t aNext = kEmpty;                        // prefetched into aNext, the next guess
    while (u >= 0 && visits < c.limit) {
        const uint64_t ukey = F[u];
        const uint32_t uid = (uint32_t)ukey;
        uint64_t key[MAXCH];
        int cutoff;

        if constexpr (MAXCH == 1) {
            // ---- rows of at most 64 entries: entry e stays on lane e.  The row comes from
            // the prefetch when u is the node it guessed; the seen filter's slot reads are
            // issued before the wait on u's visit, so the two round trips overlap; every
            // entry that may pass the filter has its vector load issued once the slots are
            // read, the rest of the kept entries (an earlier lane took their slot) after ----
            const uint32_t a0 = uid == predId ? aNext
                                              : ((uint32_t)lane < G ? L.adj[(size_t)uid * G + lane] : kEmpty);
            // the row of the entry the previous scan found after u: the next node unless
            // this hop brings a closer candidate.  A wrong guess costs one row.
            predId = nextPred;
            aNext = (predId != kEmpty && (uint32_t)lane < G) ? L.adj[(size_t)predId * G + lane] : kEmpty;
            const bool part = (a0 != kEmpty) && (a0 != c.qid);
            const SeenProbe sp = seenFilterBegin<FGLOBAL>(filt, tag, fmask, a0, part, lane);

            bool overflow = false;
            if (lane == 0) {
                FV[u] = 1;
                overflow = !visitedInsert<VGLOBAL>(vkey, vcnt, uid, logCap, vmask, capV, distinctVisited);
            }
            distinctVisited = (uint32_t)readlane((int)distinctVisited, 0);
            if (readlane((int)overflow, 0)) {
                flags |= kFlagOverflow;
                break;
            }
            visits++;
            hops++;
            P.mark(kProfVisit);
            P.count(kProfHops);
            cutoff = (fsize < Ls) ? 0x7FFFFFFF : (int)(F[fsize - 1] >> 32);
            edges += __popcll(ballot(a0 != kEmpty));

            const bool spec = part && (sp.coll || sp.prev != a0);
            const uint32_t gmask = (uint32_t)(ballot(spec) >> (8 * grp)) & 0xFFu;
            // rounds of entries that are not loaded keep whatever the registers hold: their
            // distances are computed and thrown away
            int4 v[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                asm("" : "=v"(v[r].x), "=v"(v[r].y), "=v"(v[r].z), "=v"(v[r].w));
                if ((gmask >> r) & 1u)
                    v[r] = *reinterpret_cast<const int4*>(L.base + (size_t)groupBroadcast8(a0, r) * kVecBytes + sub * 16);
            }
            const bool keep = seenFilterEnd<FGLOBAL>(filt, sp, a0, part, lane, lanesBelow);
            const uint64_t keepM = ballot(keep);
            comps += __popcll(keepM);
            const uint64_t late = keepM & ~ballot(spec);
            if (late) {
                const uint32_t lmask = (uint32_t)(late >> (8 * grp)) & 0xFFu;
#pragma unroll
                for (int r = 0; r < 8; r++)
                    if ((lmask >> r) & 1u)
                        v[r] = *reinterpret_cast<const int4*>(L.base + (size_t)groupBroadcast8(a0, r) * kVecBytes + sub * 16);
            }
            P.mark(kProfFilter);
            // expansion count of each kept entry, read while the vectors are in flight
            const uint32_t cnt = keep ? visitedCount<VGLOBAL>(vkey, vcnt, a0, logCap, vmask) : 0u;
            int part8[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                v[r].w &= ~normMask;
                part8[r] = slicePartial(v[r], qv);
            }
            const int d = qq + transposeReduce8(part8, lane);  // distance of entry `lane`
            key[0] = (keep && d < cutoff) ? countKey(d, a0, cnt) : ~0ull;
        } else {
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
                FV[u] = 1;
                overflow = !visitedInsert<VGLOBAL>(vkey, vcnt, uid, logCap, vmask, capV, distinctVisited);
            }
            distinctVisited = (uint32_t)readlane((int)distinctVisited, 0);
            if (readlane((int)overflow, 0)) {
                flags |= kFlagOverflow;
                break;
            }
            visits++;
            hops++;
            P.mark(kProfVisit);
            P.count(kProfHops);

            cutoff = (fsize < Ls) ? 0x7FFFFFFF : (int)(F[fsize - 1] >> 32);

            // ---- lossy seen filter in row order (ParlayANN has_been_seen) ----
            uint32_t nP = 0;
#pragma unroll
            for (int ch = 0; ch < MAXCH; ch++) {
                if (ch >= (int)nch) break;
                const uint32_t ai = a[ch];
                edges += __popcll(ballot(ai != kEmpty));
                const bool part = (ai != kEmpty) && (ai != c.qid);
                const bool keep = seenFilterChunk<FGLOBAL>(filt, tag, fmask, ai, part, lane, lanesBelow);
                const uint64_t km = ballot(keep);
                if (keep) pend[nP + mbcnt(km)] = ai;
                nP += (uint32_t)__popcll(km);
                __builtin_amdgcn_wave_barrier();
            }
            comps += (int)nP;
            P.mark(kProfFilter);

            // ---- distances: 8 rows per round, 8 lanes x 16 B each ----
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
                // expansion count of the candidate this lane ends up holding, read while the
                // vectors are in flight
                const uint32_t il = (uint32_t)sub * 8 + grp;
                uint32_t cid = 0, cnt = 0;
                if (il < nb) {
                    cid = pend[b * 64 + il];
                    cnt = visitedCount<VGLOBAL>(vkey, vcnt, cid, logCap, vmask);
                }
                int part8[8];
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    v[r].w &= ~normMask;
                    part8[r] = slicePartial(v[r], qv);
                }
                const int tot = transposeReduce8(part8, lane);
                if (il < nb) {
                    const int d = qq + tot;
                    if (d < cutoff) key[b] = countKey(d, cid, cnt);
                }
            }
        }

        P.mark(kProfDist);
        // ---- sort candidates by (d, id); a few candidates (the usual case of long beams)
        // by rank ----
        if constexpr (MAXCH == 1) {
            const uint64_t vm = ballot(key[0] != ~0ull);
            if (__popcll(vm) <= kRankSortMax) key[0] = rankSortFew(key[0], vm, lane, lanesBelow);
            else bitonicSort<1>(key, lane);
        } else {
            bitonicSort<MAXCH>(key, lane);
        }
        P.mark(kProfSort);

        // ---- multiset union with the frontier ----
        uint64_t runStarts[MAXCH];
#pragma unroll
        for (int b = 0; b < MAXCH; b++) {
            uint64_t up = (uint64_t)__shfl_up((int)(uint32_t)key[b], 1, 64) |
                          ((uint64_t)(uint32_t)__shfl_up((int)(uint32_t)(key[b] >> 32), 1, 64) << 32);
            if (lane == 0) up = (b == 0) ? ~key[b] : readlane64(key[b - 1], 63);
            runStarts[b] = ballot(up != key[b]);
        }
        uint32_t kBase = 0;
        int placed = 0;  // entering copies that land inside the first Ls positions
        int pmin = -1;   // merged position of the first entering copy
        int jPos[MAXCH];
        bool enter[MAXCH];
#pragma unroll
        for (int b = 0; b < MAXCH; b++) {
            const bool valid = key[b] != ~0ull;
            const uint64_t kv = valid ? frontierOfCountKey(key[b]) : ~0ull;
            const int cnt = (int)(key[b] & 0x1FFull);
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
            if (pmin < 0 && em) pmin = readlane(jPos[b], __ffsll((unsigned long long)em) - 1);
            placed += __popcll(ballot(enter[b] && jPos[b] < Ls));
            if (enter[b]) {
                // this copy is visited if the id was expanded more than `rep` times
                kkey[jK] = kv;
                kvis[jK] = (unsigned char)(rep < cnt ? 1 : 0);
                if (jPos[b] < Ls) occ[jPos[b]] = 1;
            }
        }
        const int nK = (int)kBase;
        __builtin_amdgcn_wave_barrier();

        P.mark(kProfUnion);
        // ---- merge in place: positions from the first entering copy on, last chunk first,
        // each position taking its source by rank (sources never lie above it) ----
        const int uPrev = u;
        if (placed > 0) {
            const int newSize = min(Ls, fsize + nK);
            int rankEnd = placed;  // entering copies before the current chunk's end
            for (int p0 = pmin + ((newSize - pmin - 1) & ~63); p0 >= pmin; p0 -= 64) {
                const int p = p0 + lane;
                const bool inb = p < newSize;
                const bool o = inb && occ[p];
                const uint64_t ob = ballot(o);
                const int rank = rankEnd - __popcll(ob) + (int)mbcnt(ob);
                uint64_t nk = 0;
                unsigned char nv = 0;
                if (inb) {
                    if (o) {
                        nk = kkey[rank];
                        nv = kvis[rank];
                    } else {
                        nk = F[p - rank];
                        nv = FV[p - rank];
                    }
                }
                __builtin_amdgcn_wave_barrier();
                if (inb) {
                    F[p] = nk;
                    FV[p] = nv;
                    occ[p] = 0;
                }
                rankEnd -= __popcll(ob);
            }
            __builtin_amdgcn_wave_barrier();
            fsize = newSize;
        }

        P.mark(kProfMerge);
        // ---- cut: upper_bound(F, (0, (float)(cut * (double)F[k].d))).  The predicate is
        // monotone over the sorted frontier: lane l tests the last position of stretch l
        // (stretches of ceil(fsize / 64)), the first true lane names the stretch where it
        // turns true, a second ballot the position.  Nothing entered: nothing to cut ----
        if (placed > 0 && c.k > 0 && fsize > c.k) {
            const int dk = (int)(F[c.k] >> 32);
            const float t = (float)(c.cut * (double)dk);
            const int stride = (fsize + 63) >> 6;
            auto pred = [&](int p) {
                const uint64_t kv = F[p];
                const float df = (float)(int)(kv >> 32);
                return (t < df) || (t == df && (uint32_t)kv > 0u);
            };
            const int lo0 = lane * stride;
            const uint64_t m = ballot(lo0 < fsize && pred(min(lo0 + stride, fsize) - 1));
            if (m) {
                const int f = __ffsll((unsigned long long)m) - 1;
                const int lo = f * stride, hi = min(lo + stride, fsize);
                int nf = hi - 1;  // the predicate holds there
                for (int p0 = lo; p0 < hi; p0 += 64) {
                    const int p = p0 + lane;
                    const uint64_t m2 = ballot(p < hi && pred(p));
                    if (m2) {
                        nf = p0 + __ffsll((unsigned long long)m2) - 1;
                        break;
                    }
                }
                fsize = nf;
            }
        }

        P.mark(kProfCut);
        // ---- next: first frontier copy not visited.  Positions below both the first
        // entering copy and u + 1 hold old, visited copies ----
        u = -1;
        int u2 = -1;  // the next unvisited copy after u in the same 64 positions
        for (int p0 = placed > 0 ? min(pmin, uPrev + 1) : uPrev + 1; p0 < fsize; p0 += 64) {
            const int p = p0 + lane;
            const bool nv = (p < fsize) && !FV[p];
            const uint64_t m = ballot(nv);
            if (m) {
                u = p0 + __ffsll((unsigned long long)m) - 1;
                const uint64_t m2 = m & (m - 1);
                if (m2) u2 = p0 + __ffsll((unsigned long long)m2) - 1;
                break;
            }
        }
        if constexpr (MAXCH == 1)
            nextPred = u2 >= 0 ? (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)F[u2]) : kEmpty;
        P.mark(kProfNext);
    }

    // ---- output: first k of the frontier (cuda/busqueda.cu:487-497) ----
    if (fsize < c.k) flags |= kFlagShort;
    uint32_t* oi = L.outIds + (size_t)q * L.outStride;
    int32_t* od = L.outDists + (size_t)q * L.outStride;
    for (int j = lane; j < (int)L.outStride; j += 64) {
        if (j < c.k && j < fsize) {
            const uint64_t kv = F[j];
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
    P.mark(kProfOutput);
}

// ---------------------------------------------------------------------------------------
// Cooperative general form: WPQ wavefronts (one block) share one query and its LDS.  The
// semantics and every step are searchOne's; the work of a hop is split as follows:
//   * wavefront 0 (the leader) reads the row, updates the visited multiset and runs the
//     seen filter in row order, publishing the kept ids and the cutoff;
//   * every wavefront computes the distances (and expansion counts) of the candidates
//     whose round (candidate / 8) it owns, round r to wavefront r % WPQ, into LDS keys;
//   * the leader sorts the keys and computes the union marks, publishing the counts;
//   * every wavefront merges chunks of 64 merged positions, the chunks taken from the
//     top in rounds of WPQ: a round reads all its sources, then writes (a position's
//     source never lies above it, so later rounds read positions below every write);
//   * every wavefront applies the cut and picks the next node itself: same LDS, same
//     result, no broadcast.
// Block barriers separate the phases; one query per block keeps them per query.
template <int MAXCH, bool FGLOBAL, bool VGLOBAL, int WPQ>
__device__ void searchCoop(const FlatLaunch& L, unsigned char* lds, uint32_t q, const QueryCfg& c,
                           uint32_t* gfilt, uint32_t* gvk, unsigned char* gvc, int w) {
    const int lane = laneId();
    const int grp = lane >> 3, sub = lane & 7;
    const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const bool leader = w == 0;

    uint64_t* F = reinterpret_cast<uint64_t*>(lds + L.lds.fkey);
    unsigned char* FV = lds + L.lds.fvis;
    unsigned char* occ = lds + L.lds.occ;
    uint64_t* kkey = reinterpret_cast<uint64_t*>(lds + L.lds.kkey);
    unsigned char* kvis = lds + L.lds.kvis;
    uint32_t* pend = reinterpret_cast<uint32_t*>(lds + L.lds.pend);
    unsigned char* tag = lds + L.lds.tag;
    uint64_t* ck = reinterpret_cast<uint64_t*>(lds + L.lds.ckey);
    int* scal = reinterpret_cast<int*>(lds + L.lds.cscal);
    uint32_t* vkey = VGLOBAL ? gvk : reinterpret_cast<uint32_t*>(lds + L.lds.vkey);
    unsigned char* vcnt = VGLOBAL ? gvc : lds + L.lds.vcnt;
    uint32_t* filt = FGLOBAL ? gfilt : reinterpret_cast<uint32_t*>(lds + L.lds.filt);

    const int Ls = c.Ls;
    const int fbits = flatFilterBits(Ls);
    const uint32_t fmask = (1u << fbits) - 1;
    uint32_t capV = flatVisitedCap(Ls);
    if (capV > L.capV) capV = L.capV;
    const uint32_t logCap = (uint32_t)ceilLog2(capV);
    const uint32_t vmask = capV - 1;

    const int normMask = (L.normsInRow && sub == 7) ? -1 : 0;
    int4 qv = *reinterpret_cast<const int4*>(L.queries + (size_t)q * kVecBytes + sub * 16);
    qv.w &= ~normMask;
    const int qq = groupSum8(sliceNorm(qv));

    // fresh tables and the start point (node 0), by the leader
    if (leader) {
        const uint4 e4 = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
        filtClear<FGLOBAL>(filt, fbits, lane);
        uint4* v4 = reinterpret_cast<uint4*>(vkey);
        for (uint32_t i = lane; i < capV / 4; i += 64) v4[i] = e4;
        for (int i = lane; i < Ls; i += 64) occ[i] = 0;
        const int d0 = groupDistance(L.base, 0u, sub, qv, qq, normMask);
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            F[0] = (uint64_t)(uint32_t)d0 << 32;
            FV[0] = 0;
        }
    }
    __syncthreads();

    int fsize = 1;
    int u = 0;
    int64_t visits = 0;
    int comps = 1, edges = 0, hops = 0, flags = 0;
    uint32_t distinctVisited = 0;
    const uint32_t G = L.G;
    const uint32_t nch = (G + 63) / 64;

    while (u >= 0 && visits < c.limit) {
        // ---- leader: row, visit, seen filter ----
        if (leader) {
            const uint32_t uid = (uint32_t)F[u];
            uint32_t a[MAXCH];
            const uint32_t* row = L.adj + (size_t)uid * G;
#pragma unroll
            for (int ch = 0; ch < MAXCH; ch++) {
                const uint32_t j = ch * 64 + lane;
                a[ch] = (ch < (int)nch && j < G) ? row[j] : kEmpty;
            }
            bool overflow = false;
            if (lane == 0) {
                FV[u] = 1;
                overflow = !visitedInsert<VGLOBAL>(vkey, vcnt, uid, logCap, vmask, capV, distinctVisited);
            }
            distinctVisited = (uint32_t)readlane((int)distinctVisited, 0);
            const int ovf = readlane((int)overflow, 0);
            uint32_t nP = 0;
            if (!ovf) {
#pragma unroll
                for (int ch = 0; ch < MAXCH; ch++) {
                    if (ch >= (int)nch) break;
                    const uint32_t ai = a[ch];
                    edges += __popcll(ballot(ai != kEmpty));
                    const bool part = (ai != kEmpty) && (ai != c.qid);
                    const bool keep = seenFilterChunk<FGLOBAL>(filt, tag, fmask, ai, part, lane, lanesBelow);
                    const uint64_t km = ballot(keep);
                    if (keep) pend[nP + mbcnt(km)] = ai;
                    nP += (uint32_t)__popcll(km);
                    __builtin_amdgcn_wave_barrier();
                }
            }
            if (lane == 0) {
                scal[0] = (int)nP;
                scal[1] = (fsize < Ls) ? 0x7FFFFFFF : (int)(F[fsize - 1] >> 32);
                scal[2] = ovf;
            }
        }
        __syncthreads();
        const uint32_t nP = (uint32_t)__builtin_amdgcn_readfirstlane(scal[0]);
        const int cutoff = __builtin_amdgcn_readfirstlane(scal[1]);
        if (__builtin_amdgcn_readfirstlane(scal[2])) {
            flags |= kFlagOverflow;
            break;
        }
        visits++;
        hops++;
        comps += (int)nP;

        // ---- all: distances of the rounds this wavefront owns ----
#pragma unroll
        for (int b = 0; b < MAXCH; b++) {
            if ((uint32_t)b * 64 >= nP) continue;
            const uint32_t nb = min(64u, nP - (uint32_t)b * 64);
            const uint32_t rounds = (nb + 7) / 8;
            int4 v[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                v[r] = make_int4(0, 0, 0, 0);
                const uint32_t i = (uint32_t)r * 8 + grp;
                if (r % WPQ == w && (uint32_t)r < rounds && i < nb) {
                    const uint32_t id = pend[b * 64 + i];
                    v[r] = *reinterpret_cast<const int4*>(L.base + (size_t)id * kVecBytes + sub * 16);
                }
            }
            const uint32_t il = (uint32_t)sub * 8 + grp;  // candidate this lane ends up with
            const bool mine = il < nb && (int)(il >> 3) % WPQ == w;
            uint32_t cid = 0, cnt = 0;
            if (mine) {
                cid = pend[b * 64 + il];
                cnt = visitedCount<VGLOBAL>(vkey, vcnt, cid, logCap, vmask);
            }
            int part8[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                v[r].w &= ~normMask;
                part8[r] = slicePartial(v[r], qv);
            }
            const int tot = transposeReduce8(part8, lane);
            if (mine) {
                const int d = qq + tot;
                ck[b * 64 + il] = d < cutoff ? countKey(d, cid, cnt) : ~0ull;
            }
        }
        __syncthreads();

        // ---- leader: sort, multiset union marks ----
        if (leader) {
            uint64_t key[MAXCH];
#pragma unroll
            for (int b = 0; b < MAXCH; b++) {
                const uint32_t i = (uint32_t)b * 64 + lane;
                key[b] = i < nP ? ck[i] : ~0ull;
            }
            bitonicSort<MAXCH>(key, lane);
            uint64_t runStarts[MAXCH];
#pragma unroll
            for (int b = 0; b < MAXCH; b++) {
                uint64_t up = (uint64_t)__shfl_up((int)(uint32_t)key[b], 1, 64) |
                              ((uint64_t)(uint32_t)__shfl_up((int)(uint32_t)(key[b] >> 32), 1, 64) << 32);
                if (lane == 0) up = (b == 0) ? ~key[b] : readlane64(key[b - 1], 63);
                runStarts[b] = ballot(up != key[b]);
            }
            uint32_t kBase = 0;
            int placed = 0, pmin = -1;
#pragma unroll
            for (int b = 0; b < MAXCH; b++) {
                const bool valid = key[b] != ~0ull;
                const uint64_t kv = valid ? frontierOfCountKey(key[b]) : ~0ull;
                const int cnt = (int)(key[b] & 0x1FFull);
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
                const bool enter = valid && rep >= cF;
                const uint64_t em = ballot(enter);
                const uint32_t jK = kBase + mbcnt(em);
                kBase += (uint32_t)__popcll(em);
                const int jPos = lb + cF + (int)jK;
                if (pmin < 0 && em) pmin = readlane(jPos, __ffsll((unsigned long long)em) - 1);
                placed += __popcll(ballot(enter && jPos < Ls));
                if (enter) {
                    kkey[jK] = kv;
                    kvis[jK] = (unsigned char)(rep < cnt ? 1 : 0);
                    if (jPos < Ls) occ[jPos] = 1;
                }
            }
            if (lane == 0) {
                scal[3] = placed;
                scal[4] = pmin;
                scal[5] = (int)kBase;
            }
        }
        __syncthreads();
        const int placed = __builtin_amdgcn_readfirstlane(scal[3]);
        const int pmin = __builtin_amdgcn_readfirstlane(scal[4]);
        const int nK = __builtin_amdgcn_readfirstlane(scal[5]);

        // ---- all: merge in place, chunks from the top in rounds of WPQ ----
        const int uPrev = u;
        if (placed > 0) {
            const int newSize = min(Ls, fsize + nK);
            const int top = pmin + ((newSize - pmin - 1) & ~63);
            const int nChunks = ((top - pmin) >> 6) + 1;
            for (int r0 = 0; r0 < nChunks; r0 += WPQ) {
                const int j = r0 + w;
                const bool active = j < nChunks;
                const int p0 = top - 64 * j;
                const int p = p0 + lane;
                const bool inb = active && p < newSize;
                uint64_t nk = 0;
                unsigned char nv = 0;
                if (active) {
                    // entering copies at or above this chunk's end, counted chunk by chunk
                    int above = 0;
                    for (int jj = 0; jj < j; jj++) {
                        const int pp = top - 64 * jj + lane;
                        above += __popcll(ballot(pp < newSize && occ[pp]));
                    }
                    const bool o = inb && occ[p];
                    const uint64_t ob = ballot(o);
                    const int rank = (placed - above) - __popcll(ob) + (int)mbcnt(ob);
                    if (inb) {
                        if (o) {
                            nk = kkey[rank];
                            nv = kvis[rank];
                        } else {
                            nk = F[p - rank];
                            nv = FV[p - rank];
                        }
                    }
                }
                __syncthreads();
                if (inb) {
                    F[p] = nk;
                    FV[p] = nv;
                }
                __syncthreads();
            }
            // the marks are counted by every later round, so they clear only now (the
            // leader sets new ones after the next hop's first barriers)
            for (int j = w; j < nChunks; j += WPQ) {
                const int p = top - 64 * j + lane;
                if (p < newSize) occ[p] = 0;
            }
            fsize = newSize;
        }

        // ---- all: cut, then the next node (reads only) ----
        if (placed > 0 && c.k > 0 && fsize > c.k) {
            const int dk = (int)(F[c.k] >> 32);
            const float t = (float)(c.cut * (double)dk);
            const int stride = (fsize + 63) >> 6;
            auto pred = [&](int p) {
                const uint64_t kv = F[p];
                const float df = (float)(int)(kv >> 32);
                return (t < df) || (t == df && (uint32_t)kv > 0u);
            };
            const int lo0 = lane * stride;
            const uint64_t m = ballot(lo0 < fsize && pred(min(lo0 + stride, fsize) - 1));
            if (m) {
                const int f = __ffsll((unsigned long long)m) - 1;
                const int lo = f * stride, hi = min(lo + stride, fsize);
                int nf = hi - 1;
                for (int p0 = lo; p0 < hi; p0 += 64) {
                    const int p = p0 + lane;
                    const uint64_t m2 = ballot(p < hi && pred(p));
                    if (m2) {
                        nf = p0 + __ffsll((unsigned long long)m2) - 1;
                        break;
                    }
                }
                fsize = nf;
            }
        }
        u = -1;
        for (int p0 = placed > 0 ? min(pmin, uPrev + 1) : uPrev + 1; p0 < fsize; p0 += 64) {
            const int p = p0 + lane;
            const bool nv = (p < fsize) && !FV[p];
            const uint64_t m = ballot(nv);
            if (m) {
                u = p0 + __ffsll((unsigned long long)m) - 1;
                break;
            }
        }
        // every wavefront has read F / FV before the leader's next hop writes them
        __syncthreads();
    }

    // ---- output by the leader: first k of the frontier ----
    if (leader) {
        if (fsize < c.k) flags |= kFlagShort;
        uint32_t* oi = L.outIds + (size_t)q * L.outStride;
        int32_t* od = L.outDists + (size_t)q * L.outStride;
        for (int j = lane; j < (int)L.outStride; j += 64) {
            if (j < c.k && j < fsize) {
                const uint64_t kv = F[j];
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
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------------
// Register-resident form for Ls <= 64 (128 with two keys per lane) and rows of at most
// 64 entries.  Same semantics as searchOne; lane i holds frontier entry i as one tagged
// 64-bit key
//     d (23 bits) | id (32 bits) | tag (9 bits)
// (an exact int8 squared distance is below 128 * 255^2 < 2^23).  Tag bit 8 is 0 for a
// frontier copy, whose bits 1..7 hold its position (equal keys keep their order through
// the merge) and bit 0 its visited flag; it is 1 for a candidate, whose bits 0..6 hold
// how often its id was expanded (saturated at 127; a row has at most 64 copies).
//   * candidate vectors are loaded straight from the row layout (entry e = lane e) as
//     soon as the filter slots are read, while in-row slot collisions are resolved;
//   * the candidates are sorted descending with DPP / permlane exchanges and merged
//     with the frontier by one bitonic merge.  In a run of equal keys the cF frontier
//     copies come first; candidate copy rep (0-based) enters iff rep >= cF, visited iff
//     rep < expansions(id) -- the multiset union of the reference's search.
constexpr uint64_t kNoKey = ~0ull;
constexpr int kTagBits = 9;
constexpr uint32_t kCand = 0x100u;        // tag: candidate copy
constexpr uint64_t kTagMask = 0x1FFull;
constexpr int kInsertMax = 16;  // up to this many candidates are inserted one by one

__device__ __forceinline__ uint64_t tagKey(int d, uint32_t id, uint32_t tag) {
    return ((uint64_t)(uint32_t)d << 41) | ((uint64_t)id << kTagBits) | tag;
}
__device__ __forceinline__ int keyDist(uint64_t k) { return (int)((uint32_t)(k >> 32) >> 9); }
__device__ __forceinline__ uint32_t keyId(uint64_t k) {
    return ((uint32_t)(k >> 32) << 23) | ((uint32_t)k >> 9);
}

// Row `i` of a table at byte offset i * pitch + extra: a 32-bit offset from a scalar base
// when the table is below 4 GiB (NARROW), else 64-bit address arithmetic.
template <bool NARROW, typename T>
__device__ __forceinline__ const T* rowAt(const void* base, uint32_t i, uint32_t pitch, uint32_t extra) {
    const char* b = static_cast<const char*>(base);
    if constexpr (NARROW) return reinterpret_cast<const T*>(b + (i * pitch + extra));
    else return reinterpret_cast<const T*>(b + ((size_t)i * pitch + extra));
}

template <bool DESC>
__device__ __forceinline__ uint64_t sortWave(uint64_t x, int lane) {
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const bool asc = ((lane & k) == 0) != DESC;
            x = cmpXchg(x, j, ((lane & j) == 0) == asc, lane);
        }
    }
    return x;
}

__device__ __forceinline__ uint64_t mergeHalf(uint64_t x, int lane) {
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) x = cmpXchg(x, j, (lane & j) == 0, lane);
    return x;
}

// Highest set bit of m at or below this lane (upTo = lanes 0..lane), or -1.
__device__ __forceinline__ int lastAtOrBelow(uint64_t m, uint64_t upTo) {
    const uint64_t x = m & upTo;
    return x ? 63 - __clzll((long long)x) : -1;
}

struct UnionMarks {
    uint64_t drop;  // candidate copies the union drops
    int lastRun;    // run start / first candidate copy seen by lane 63 (carried to the
    int lastC;      // next 64 positions of the merged sequence)
};

// Multiset-union marks over positions base..base+63 of the merged sequence; x holds this
// lane's entry, prev the entry before it (kNoKey before position 0).  Candidate copies
// get their visited flag in bit 0 (bits 1..6 cleared).
__device__ __forceinline__ UnionMarks unionMarks(uint64_t& x, uint64_t prev, int base,
                                                 int carryRun, int carryC, int lane,
                                                 uint64_t upTo) {
    const bool valid = x != kNoKey;
    const bool isC = valid && (x & kCand);
    const bool same = valid && ((x ^ prev) >> kTagBits) == 0;
    const bool prevC = (prev & kCand) != 0;
    int run = lastAtOrBelow(ballot(valid && !same), upTo);
    run = run < 0 ? carryRun : base + run;
    int fc = lastAtOrBelow(ballot(isC && !(same && prevC)), upTo);
    fc = fc < 0 ? carryC : base + fc;
    bool drop = false;
    if (isC) {
        const int p = base + lane;
        drop = p < 2 * fc - run;  // rep = p - fc < cF = fc - run
        const uint32_t vis = (p - fc) < (int)(x & 0x7Fu) ? 1u : 0u;
        x = (x & ~0x7Full) | vis;
    }
    UnionMarks m;
    m.drop = ballot(drop);
    m.lastRun = readlane(run, 63);
    m.lastC = readlane(fc, 63);
    return m;
}

// Expansions of id a so far, from the register-resident forms' visited hash in LDS
// (capped at 127, the width of the tag field that carries it).
__device__ __forceinline__ uint32_t fastExpansions(const uint32_t* vkey, const unsigned char* vcnt,
                                                   uint32_t a, uint32_t logCap, uint32_t vmask) {
    uint32_t hh = visitedSlot(a, logCap);
    while (true) {
        const uint32_t kk = vkey[hh];
        if (kk == a) return min(127u, (uint32_t)vcnt[hh]);
        if (kk == kEmpty) return 0;
        hh = (hh + 1) & vmask;
    }
}

// ROW2: adjacency rows of up to 128 entries, read as two 64-entry halves; each half runs
// the filter and distance steps in row order (the second half's filter sees the first's
// slot writes), and the merge takes the candidates of both, counting earlier copies of a
// key across the halves.
template <bool NARROW, int PREFETCH, bool NORMS, bool ROW2 = false>
__device__ void searchFast(const FlatLaunch& L, unsigned char* lds, uint32_t q, const QueryCfg& c,
                           Prof& P) {
    static_assert(!ROW2 || PREFETCH == 0, "rows of two halves do not prefetch");
    const int lane = laneId();
    const int grp = lane >> 3, sub = lane & 7;
    const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    uint64_t* scratch = reinterpret_cast<uint64_t*>(lds + L.lds.kkey);
    unsigned char* tag = lds + L.lds.tag;
    uint32_t* vkey = reinterpret_cast<uint32_t*>(lds + L.lds.vkey);
    unsigned char* vcnt = lds + L.lds.vcnt;
    uint32_t* filt = reinterpret_cast<uint32_t*>(lds + L.lds.filt);
    const uint64_t upTo = lanesBelow | (1ull << lane);

    const int Ls = c.Ls;
    const uint32_t fmask = (1u << kFastFilterBits) - 1;
    uint32_t capV = flatVisitedCap(Ls);
    if (capV > L.capV) capV = L.capV;
    const uint32_t logCap = (uint32_t)ceilLog2(capV);
    const uint32_t vmask = capV - 1;
    const uint32_t win = capV < 64 ? capV : 64;

    // NORMS: rows hold |p|^2 in bytes 124..127, lane 7's .w of each 8-lane group
    const int normMask = (NORMS && sub == 7) ? -1 : 0;
    int4 qv = *reinterpret_cast<const int4*>(L.queries + (size_t)q * kVecBytes + sub * 16);
    qv.w &= ~normMask;
    const int qq = groupSum8(sliceNorm(qv));

    {
        const uint4 e4 = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
        filtClear<false>(filt, kFastFilterBits, lane);
        uint4* v4 = reinterpret_cast<uint4*>(vkey);
        for (uint32_t i = lane; i < capV / 4; i += 64) v4[i] = e4;
    }

    // start point: node 0 (cuda/busqueda.cu:166-173)
    const int d0 = groupDistance(L.base, 0u, sub, qv, qq, normMask);
    uint64_t F = lane == 0 ? tagKey(d0, 0u, 0u) : kNoKey;
    int fsize = 1;
    int u = 0;
    int64_t visits = 0;
    int comps = 1, edges = 0, hops = 0, flags = 0;
    uint32_t distinct = 0;
    const uint32_t G = L.G;
    uint32_t predId = kEmpty, predId2 = kEmpty;  // nodes whose rows were prefetched into
    uint32_t aNext = kEmpty, aNext2 = kEmpty;    // aNext (and, depth 2, aNext2)
    __builtin_amdgcn_wave_barrier();
    P.mark(kProfInit);
    P.count(kProfQueries);

    while (u >= 0 && visits < c.limit) {
        const uint32_t uid = keyId(readlane64(F, u));
        uint32_t a = aNext;
        if (PREFETCH >= 2 && uid != predId && uid == predId2) a = aNext2;
        else if (PREFETCH == 0 || uid != predId)
            a = (uint32_t)lane < G ? *rowAt<NARROW, uint32_t>(L.adj, uid, 4 * G, 4 * (uint32_t)lane) : kEmpty;
        uint32_t a1 = kEmpty;
        if constexpr (ROW2)
            if (64u + (uint32_t)lane < G) a1 = *rowAt<NARROW, uint32_t>(L.adj, uid, 4 * G, 4 * (64u + (uint32_t)lane));

        // visit u: its flag, and its expansion count (64 probe slots read at once)
        if (lane == u) F |= 1u;
        // Prefetch the row of the likely next node: the first unvisited frontier entry after
        // u, which is the next node unless this hop brings a closer candidate (depth 2: the
        // second one too, the next node when the first is displaced by one candidate...).  The
        // loads have the whole hop to arrive; a wrong guess costs one row, and the next hop
        // loads its own.
        if constexpr (PREFETCH >= 1) {
            uint64_t unv = ballot(lane < fsize && !(F & 1u));
            predId = unv ? keyId(readlane64(F, __ffsll((unsigned long long)unv) - 1)) : kEmpty;
            aNext = kEmpty;
            if (predId != kEmpty && (uint32_t)lane < G)
                aNext = *rowAt<NARROW, uint32_t>(L.adj, predId, 4 * G, 4 * (uint32_t)lane);
            if constexpr (PREFETCH >= 2) {
                unv &= unv - 1;
                predId2 = unv ? keyId(readlane64(F, __ffsll((unsigned long long)unv) - 1)) : kEmpty;
                aNext2 = kEmpty;
                if (predId2 != kEmpty && (uint32_t)lane < G)
                    aNext2 = *rowAt<NARROW, uint32_t>(L.adj, predId2, 4 * G, 4 * (uint32_t)lane);
            }
        }
        bool overflow = false;
        {
            const uint32_t h0 = visitedSlot(uid, logCap);
            for (uint32_t off = 0; off < capV; off += win) {
                const uint32_t s = (h0 + off + (uint32_t)lane) & vmask;
                const bool act = (uint32_t)lane < win;
                const uint32_t kk = act ? vkey[s] : 0u;
                const uint64_t hit = ballot(act && kk == uid);
                const uint64_t m = hit | ballot(act && kk == kEmpty);
                if (m) {
                    const int p = __ffsll((unsigned long long)m) - 1;
                    if ((hit >> p) & 1ull) {
                        if (lane == p) vcnt[s] = (unsigned char)(vcnt[s] + 1);
                    } else if (distinct + 1 < capV) {
                        if (lane == p) {
                            vkey[s] = uid;
                            vcnt[s] = 1;
                        }
                        distinct++;
                    } else {
                        overflow = true;
                    }
                    break;
                }
            }
        }
        if (overflow) {
            flags |= kFlagOverflow;
            break;
        }
        visits++;
        hops++;
        P.mark(kProfVisit);
        P.count(kProfHops);

        const int cutoff = (fsize < Ls) ? 0x7FFFFFFF
                                        : (int)((uint32_t)readlane((int)(uint32_t)(F >> 32), fsize - 1) >> 9);

        uint64_t C, C1 = kNoKey;
        if constexpr (ROW2) {
            // one 64-entry half of the row: seen filter, vector loads, distances; the key
            // of entry `lane` if it is a candidate below the cutoff, else kNoKey
            auto halfKeys = [&](const uint32_t a) -> uint64_t {
                // seen filter, first half: slots and tags
                edges += __popcll(ballot(a != kEmpty));
                const bool part = (a != kEmpty) && (a != c.qid);
                const SeenProbe sp = seenFilterBegin<false>(filt, tag, fmask, a, part, lane);
                const bool spec = part && (sp.coll || sp.prev != a);

                // vectors of every entry that may pass: group g loads entry 8g + r in round r
                const uint32_t gmask = (uint32_t)(ballot(spec) >> (8 * grp)) & 0xFFu;
                uint32_t ids[8];
#pragma unroll
                for (int r = 0; r < 8; r++) ids[r] = groupBroadcast8(a, r);
                // Slices of entries that are not loaded keep whatever the registers hold: their
                // distances are computed and thrown away.
                int4 v[8];
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    // an empty asm hands the compiler registers it need not initialise
                    asm("" : "=v"(v[r].x), "=v"(v[r].y), "=v"(v[r].z), "=v"(v[r].w));
                    if ((gmask >> r) & 1u) v[r] = *rowAt<NARROW, int4>(L.base, ids[r], kVecBytes, sub * 16);
                }

                // expansions of the id so far (decide the visited flags of entering copies)
                uint32_t cnt = spec ? fastExpansions(vkey, vcnt, a, logCap, vmask) : 0u;

                // seen filter, second half: in-row collisions, slot writes
                const bool keep = seenFilterEnd<false>(filt, sp, a, part, lane, lanesBelow);
                const uint64_t keepM = ballot(keep);
                comps += __popcll(keepM);
                // A lane that won its byte tag but shares the slot with an earlier lane was not
                // speculated when the slot held its id before the row; load those now.
                // Their ids held the slot before the row, so they may have been expanded: their
                // expansion counts are read now too.
                const uint64_t late = keepM & ~ballot(spec);
                if (late) {
                    const uint32_t lmask = (uint32_t)(late >> (8 * grp)) & 0xFFu;
#pragma unroll
                    for (int r = 0; r < 8; r++)
                        if ((lmask >> r) & 1u) v[r] = *rowAt<NARROW, int4>(L.base, ids[r], kVecBytes, sub * 16);
                    if ((late >> lane) & 1ull) cnt = fastExpansions(vkey, vcnt, a, logCap, vmask);
                }
                P.mark(kProfFilter);

                int part8[8];
#pragma unroll
                for (int r = 0; r < 8; r++)
                    part8[r] = NORMS ? slicePartialN(v[r], qv, normMask) : slicePartial(v[r], qv);
                // p.p is recomputed from the row: a stored norm would cost one more random
                // cache line per candidate
                const int d = qq + transposeReduce8(part8, lane);  // distance of entry `lane`
                const uint64_t key = (keep && d < cutoff) ? tagKey(d, a, kCand | cnt) : kNoKey;
                P.mark(kProfDist);
                return key;
            };
            C = halfKeys(a);
            C1 = halfKeys(a1);
        } else {
            // seen filter, first half: slots and tags
            edges += __popcll(ballot(a != kEmpty));
            const bool part = (a != kEmpty) && (a != c.qid);
            const SeenProbe sp = seenFilterBegin<false>(filt, tag, fmask, a, part, lane);
            const bool spec = part && (sp.coll || sp.prev != a);

            // vectors of every entry that may pass: group g loads entry 8g + r in round r
            const uint32_t gmask = (uint32_t)(ballot(spec) >> (8 * grp)) & 0xFFu;
            uint32_t ids[8];
#pragma unroll
            for (int r = 0; r < 8; r++) ids[r] = groupBroadcast8(a, r);
            // Slices of entries that are not loaded keep whatever the registers hold: their
            // distances are computed and thrown away.
            int4 v[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                // an empty asm hands the compiler registers it need not initialise
                asm("" : "=v"(v[r].x), "=v"(v[r].y), "=v"(v[r].z), "=v"(v[r].w));
                if ((gmask >> r) & 1u) v[r] = *rowAt<NARROW, int4>(L.base, ids[r], kVecBytes, sub * 16);
            }

            // expansions of the id so far (decide the visited flags of entering copies)
            uint32_t cnt = spec ? fastExpansions(vkey, vcnt, a, logCap, vmask) : 0u;

            // seen filter, second half: in-row collisions, slot writes
            const bool keep = seenFilterEnd<false>(filt, sp, a, part, lane, lanesBelow);
            const uint64_t keepM = ballot(keep);
            comps += __popcll(keepM);
            // A lane that won its byte tag but shares the slot with an earlier lane was not
            // speculated when the slot held its id before the row; load those now.
            // Their ids held the slot before the row, so they may have been expanded: their
            // expansion counts are read now too.
            const uint64_t late = keepM & ~ballot(spec);
            if (late) {
                const uint32_t lmask = (uint32_t)(late >> (8 * grp)) & 0xFFu;
#pragma unroll
                for (int r = 0; r < 8; r++)
                    if ((lmask >> r) & 1u) v[r] = *rowAt<NARROW, int4>(L.base, ids[r], kVecBytes, sub * 16);
                if ((late >> lane) & 1ull) cnt = fastExpansions(vkey, vcnt, a, logCap, vmask);
            }
            P.mark(kProfFilter);

            int part8[8];
#pragma unroll
            for (int r = 0; r < 8; r++)
                part8[r] = NORMS ? slicePartialN(v[r], qv, normMask) : slicePartial(v[r], qv);
            // p.p is recomputed from the row: a stored norm would cost one more random
            // cache line per candidate
            const int d = qq + transposeReduce8(part8, lane);  // distance of entry `lane`
            C = (keep && d < cutoff) ? tagKey(d, a, kCand | cnt) : kNoKey;
            P.mark(kProfDist);
        }

        const uint64_t candM = ballot(C != kNoKey);
        const uint64_t candM1 = ROW2 ? ballot(C1 != kNoKey) : 0ull;
        const int nc = __popcll(candM) + __popcll(candM1);
        P.add(kProfCands, (unsigned long long)nc);
        if (nc == 0) P.count(kProfNoCands);
        if (nc > 0 && nc <= 8) P.count(kProfSmall);

        bool changed = false;
        // candidates from both halves of a row go through the insertion, whatever their number
        const bool bothHalves = ROW2 && candM != 0 && candM1 != 0;
        if (nc > 0 && (nc <= kInsertMax || bothHalves)) {
            // insert each at its upper bound (equal keys: frontier copies, then the
            // candidate copies already inserted), trimming to Ls as we go
            int fs = fsize;
            for (int h = 0; h < (ROW2 ? 2 : 1); h++) {
            const uint64_t Ch = h ? C1 : C;
            uint64_t mm = h ? candM1 : candM;
            while (mm) {
                const int j = __ffsll((unsigned long long)mm) - 1;
                mm &= mm - 1;
                const uint64_t x = readlane64(Ch, j);
                const uint64_t xmin = x & ~kTagMask, xmax = x | kTagMask;
                const int pos = __popcll(ballot(F <= xmax));
                if (pos >= Ls) continue;
                const int lower = __popcll(ballot(F < xmin));
                // cF frontier copies of the key; rep earlier candidate copies (inserted or
                // dropped).  Without any copy yet there is no earlier candidate copy either:
                // it would have been inserted, or dropped against a frontier copy.  The
                // entering copy keeps "expansions - rep" in its tag: visited iff positive.
                uint64_t xe = x;
                if (pos != lower) {
                    const int cF = __popcll(ballot(F >= xmin && F <= xmax && !(F & kCand)));
                    // earlier copies: in this half below lane j, and (second half) all of
                    // the first half's
                    const int rep = __popcll(ballot(Ch == x) & ((1ull << j) - 1)) +
                                    (h ? __popcll(ballot(C == x)) : 0);
                    if (rep < cF) continue;
                    xe = xmin | kCand | (uint32_t)max(0, (int)(x & 0x7Fu) - rep);
                }
                // shift the tail up one lane, then place the key (lane 0 never takes the
                // shifted value: pos >= 0)
                const uint64_t pv = ((uint64_t)dppMov<0x138>((uint32_t)(F >> 32)) << 32) |
                                    dppMov<0x138>((uint32_t)F);
                if (lane > pos) F = pv;
                if (lane == pos) F = xe;
                fs = min(Ls, fs + 1);
                changed = true;
            }
            }
            if (changed) {
                const uint32_t vis = (F & kCand) ? ((F & 0x7Fu) ? 1u : 0u) : (uint32_t)(F & 1u);
                F = lane < fs ? ((F & ~kTagMask) | ((uint64_t)lane << 1) | vis) : kNoKey;
                fsize = fs;
            }
            P.mark(kProfMerge);
        } else if (nc > kInsertMax) {
            // (rows of two halves: all candidates lie in one half here)
            const uint64_t Cd = sortWave<true>(ROW2 && candM == 0 ? C1 : C, lane);
            P.mark(kProfSort);

            // bitonic merge of frontier (ascending) and candidates (descending)
            uint64_t lo = F < Cd ? F : Cd;
            lo = mergeHalf(lo, lane);
            const UnionMarks ml = unionMarks(lo, prevLane64(lo, kNoKey, lane), 0, 0, 0, lane, upTo);
            const int validLo = __popcll(ballot(lo != kNoKey));
            int newSize;
            if (ml.drop == 0) {
                newSize = min(Ls, validLo);
            } else {
                uint64_t hi = F < Cd ? Cd : F;
                hi = mergeHalf(hi, lane);
                const UnionMarks mh = unionMarks(hi, prevLane64(hi, readlane64(lo, 63), lane), 64,
                                                 ml.lastRun, ml.lastC, lane, upTo);
                const int validHi = __popcll(ballot(hi != kNoKey));
                const int nDropLo = __popcll(ml.drop);
                const int dLo = lane - (int)__popcll(ml.drop & lanesBelow);
                const int dHi = 64 + lane - nDropLo - (int)__popcll(mh.drop & lanesBelow);
                if (lo != kNoKey && !((ml.drop >> lane) & 1ull)) scratch[dLo] = lo;
                if (hi != kNoKey && !((mh.drop >> lane) & 1ull) && dHi < 64) scratch[dHi] = hi;
                __builtin_amdgcn_wave_barrier();
                newSize = min(Ls, validLo + validHi - nDropLo - (int)__popcll(mh.drop));
                lo = lane < newSize ? scratch[lane] : kNoKey;
                __builtin_amdgcn_wave_barrier();
            }
            F = lane < newSize ? ((lo & ~kTagMask) | ((uint64_t)lane << 1) | (lo & 1u)) : kNoKey;
            fsize = newSize;
            changed = true;
            P.mark(kProfMerge);
        }

        // cut: upper_bound(F, (0, (float)(cut * (double)F[k].d))); a frontier that did
        // not change since the last cut is already cut
        if (changed && c.k > 0 && fsize > c.k) {
            const int dk = (int)((uint32_t)readlane((int)(uint32_t)(F >> 32), c.k) >> 9);
            const float t = (float)(c.cut * (double)dk);
            const float df = (float)keyDist(F);
            const bool pred = lane < fsize && ((t < df) || (t == df && keyId(F) > 0u));
            const uint64_t m = ballot(pred);
            if (m) {
                fsize = __ffsll((unsigned long long)m) - 1;
                if (lane >= fsize) F = kNoKey;
            }
        }
        P.mark(kProfCut);

        const uint64_t nv = ballot(lane < fsize && !(F & 1u));
        u = nv ? __ffsll((unsigned long long)nv) - 1 : -1;
#ifdef PANN_PROFILE
        if (u >= 0 && (keyId(readlane64(F, u)) == predId || keyId(readlane64(F, u)) == predId2))
            P.count(kProfPredFC);
#endif
        P.mark(kProfNext);
    }

    // output: first k of the frontier (cuda/busqueda.cu:487-497)
    if (fsize < c.k) flags |= kFlagShort;
    uint32_t* oi = L.outIds + (size_t)q * L.outStride;
    int32_t* od = L.outDists + (size_t)q * L.outStride;
    for (int j = lane; j < (int)L.outStride; j += 64) {
        if (j < c.k && j < fsize) {
            oi[j] = keyId(F);
            od[j] = keyDist(F);
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
    P.mark(kProfOutput);
}

// lane i <- lane 63 - i (row mirror, then the rows in reverse order)
__device__ __forceinline__ uint64_t reverseWave(uint64_t x, int lane) {
    const uint64_t m = ((uint64_t)dppMov<0x140>((uint32_t)(x >> 32)) << 32) | dppMov<0x140>((uint32_t)x);
    return xorLane64(xorLane64(m, 32, lane), 16, lane);
}

// Key at frontier position p (uniform) of a two-register frontier.
__device__ __forceinline__ uint64_t frontierKey(uint64_t F0, uint64_t F1, int p) {
    return p < 64 ? readlane64(F0, p) : readlane64(F1, p - 64);
}

// The register-resident search for 64 < Ls <= 128: positions 0..63 in F0, 64..127 in F1,
// the seen filter (2^11 or 2^12 slots) in the wavefront's HBM region.  Same steps as
// searchFast; the union of more than kInsertMax candidates merges the descending
// candidates with F0, then the upper half of that with F1: every element of the first
// half is <= every element of F1, so the two merges give the sorted union.
template <bool NARROW, bool NORMS, bool ROW2 = false>
__device__ void searchFast2(const FlatLaunch& L, unsigned char* lds, uint32_t q, const QueryCfg& c,
                            uint32_t* filt, Prof& P) {
    const int lane = laneId();
    const int grp = lane >> 3, sub = lane & 7;
    const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t upTo = lanesBelow | (1ull << lane);

    uint64_t* scratch = reinterpret_cast<uint64_t*>(lds + L.lds.kkey);  // 128 keys
    unsigned char* tag = lds + L.lds.tag;
    uint32_t* vkey = reinterpret_cast<uint32_t*>(lds + L.lds.vkey);
    unsigned char* vcnt = lds + L.lds.vcnt;

    const int Ls = c.Ls;
    const int fbits = flatFilterBits(Ls);
    const uint32_t fmask = (1u << fbits) - 1;
    uint32_t capV = flatVisitedCap(Ls);
    if (capV > L.capV) capV = L.capV;
    const uint32_t logCap = (uint32_t)ceilLog2(capV);
    const uint32_t vmask = capV - 1;
    const uint32_t win = capV < 64 ? capV : 64;

    const int normMask = (NORMS && sub == 7) ? -1 : 0;
    int4 qv = *reinterpret_cast<const int4*>(L.queries + (size_t)q * kVecBytes + sub * 16);
    qv.w &= ~normMask;
    const int qq = groupSum8(sliceNorm(qv));

    {
        const uint4 e4 = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
        filtClear<true>(filt, fbits, lane);
        uint4* v4 = reinterpret_cast<uint4*>(vkey);
        for (uint32_t i = lane; i < capV / 4; i += 64) v4[i] = e4;
    }

    const int d0 = groupDistance(L.base, 0u, sub, qv, qq, normMask);
    uint64_t F0 = lane == 0 ? tagKey(d0, 0u, 0u) : kNoKey;
    uint64_t F1 = kNoKey;
    int fsize = 1;
    int u = 0;
    int64_t visits = 0;
    int comps = 1, edges = 0, hops = 0, flags = 0;
    uint32_t distinct = 0;
    const uint32_t G = L.G;
    __builtin_amdgcn_wave_barrier();
    P.mark(kProfInit);
    P.count(kProfQueries);

    while (u >= 0 && visits < c.limit) {
        const uint32_t uid = keyId(frontierKey(F0, F1, u));
        const uint32_t a =
            (uint32_t)lane < G ? *rowAt<NARROW, uint32_t>(L.adj, uid, 4 * G, 4 * (uint32_t)lane) : kEmpty;
        uint32_t a1 = kEmpty;
        if constexpr (ROW2)
            if (64u + (uint32_t)lane < G) a1 = *rowAt<NARROW, uint32_t>(L.adj, uid, 4 * G, 4 * (64u + (uint32_t)lane));

        if (u < 64) {
            if (lane == u) F0 |= 1u;
        } else if (lane == u - 64) {
            F1 |= 1u;
        }
        bool overflow = false;
        {
            const uint32_t h0 = visitedSlot(uid, logCap);
            for (uint32_t off = 0; off < capV; off += win) {
                const uint32_t s = (h0 + off + (uint32_t)lane) & vmask;
                const bool act = (uint32_t)lane < win;
                const uint32_t kk = act ? vkey[s] : 0u;
                const uint64_t hit = ballot(act && kk == uid);
                const uint64_t m = hit | ballot(act && kk == kEmpty);
                if (m) {
                    const int p = __ffsll((unsigned long long)m) - 1;
                    if ((hit >> p) & 1ull) {
                        if (lane == p) vcnt[s] = (unsigned char)(vcnt[s] + 1);
                    } else if (distinct + 1 < capV) {
                        if (lane == p) {
                            vkey[s] = uid;
                            vcnt[s] = 1;
                        }
                        distinct++;
                    } else {
                        overflow = true;
                    }
                    break;
                }
            }
        }
        if (overflow) {
            flags |= kFlagOverflow;
            break;
        }
        visits++;
        hops++;
        P.mark(kProfVisit);
        P.count(kProfHops);

        const int cutoff = (fsize < Ls) ? 0x7FFFFFFF : keyDist(frontierKey(F0, F1, fsize - 1));

        uint64_t C, C1 = kNoKey;
        if constexpr (ROW2) {
            // one 64-entry half of the row (rows of 65 to 128 entries)
            auto halfKeys = [&](const uint32_t a) -> uint64_t {
                edges += __popcll(ballot(a != kEmpty));
                const bool part = (a != kEmpty) && (a != c.qid);
                const SeenProbe sp = seenFilterBegin<true>(filt, tag, fmask, a, part, lane);
                const bool spec = part && (sp.coll || sp.prev != a);

                const uint32_t gmask = (uint32_t)(ballot(spec) >> (8 * grp)) & 0xFFu;
                uint32_t ids[8];
#pragma unroll
                for (int r = 0; r < 8; r++) ids[r] = groupBroadcast8(a, r);
                int4 v[8];
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    asm("" : "=v"(v[r].x), "=v"(v[r].y), "=v"(v[r].z), "=v"(v[r].w));
                    if ((gmask >> r) & 1u) v[r] = *rowAt<NARROW, int4>(L.base, ids[r], kVecBytes, sub * 16);
                }

                uint32_t cnt = spec ? fastExpansions(vkey, vcnt, a, logCap, vmask) : 0u;

                const bool keep = seenFilterEnd<true>(filt, sp, a, part, lane, lanesBelow);
                const uint64_t keepM = ballot(keep);
                comps += __popcll(keepM);
                const uint64_t late = keepM & ~ballot(spec);
                if (late) {
                    const uint32_t lmask = (uint32_t)(late >> (8 * grp)) & 0xFFu;
#pragma unroll
                    for (int r = 0; r < 8; r++)
                        if ((lmask >> r) & 1u) v[r] = *rowAt<NARROW, int4>(L.base, ids[r], kVecBytes, sub * 16);
                    // ids that held their slot before the row: read their expansion counts too
                    if ((late >> lane) & 1ull) cnt = fastExpansions(vkey, vcnt, a, logCap, vmask);
                }
                P.mark(kProfFilter);

                int part8[8];
#pragma unroll
                for (int r = 0; r < 8; r++)
                    part8[r] = NORMS ? slicePartialN(v[r], qv, normMask) : slicePartial(v[r], qv);
                const int d = qq + transposeReduce8(part8, lane);
                const uint64_t key = (keep && d < cutoff) ? tagKey(d, a, kCand | cnt) : kNoKey;
                P.mark(kProfDist);
                return key;
            };
            C = halfKeys(a);
            C1 = halfKeys(a1);
        } else {
            edges += __popcll(ballot(a != kEmpty));
            const bool part = (a != kEmpty) && (a != c.qid);
            const SeenProbe sp = seenFilterBegin<true>(filt, tag, fmask, a, part, lane);
            const bool spec = part && (sp.coll || sp.prev != a);

            const uint32_t gmask = (uint32_t)(ballot(spec) >> (8 * grp)) & 0xFFu;
            uint32_t ids[8];
#pragma unroll
            for (int r = 0; r < 8; r++) ids[r] = groupBroadcast8(a, r);
            int4 v[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                asm("" : "=v"(v[r].x), "=v"(v[r].y), "=v"(v[r].z), "=v"(v[r].w));
                if ((gmask >> r) & 1u) v[r] = *rowAt<NARROW, int4>(L.base, ids[r], kVecBytes, sub * 16);
            }

            uint32_t cnt = spec ? fastExpansions(vkey, vcnt, a, logCap, vmask) : 0u;

            const bool keep = seenFilterEnd<true>(filt, sp, a, part, lane, lanesBelow);
            const uint64_t keepM = ballot(keep);
            comps += __popcll(keepM);
            const uint64_t late = keepM & ~ballot(spec);
            if (late) {
                const uint32_t lmask = (uint32_t)(late >> (8 * grp)) & 0xFFu;
#pragma unroll
                for (int r = 0; r < 8; r++)
                    if ((lmask >> r) & 1u) v[r] = *rowAt<NARROW, int4>(L.base, ids[r], kVecBytes, sub * 16);
                // ids that held their slot before the row: read their expansion counts too
                if ((late >> lane) & 1ull) cnt = fastExpansions(vkey, vcnt, a, logCap, vmask);
            }
            P.mark(kProfFilter);

            int part8[8];
#pragma unroll
            for (int r = 0; r < 8; r++)
                part8[r] = NORMS ? slicePartialN(v[r], qv, normMask) : slicePartial(v[r], qv);
            const int d = qq + transposeReduce8(part8, lane);
            C = (keep && d < cutoff) ? tagKey(d, a, kCand | cnt) : kNoKey;
            P.mark(kProfDist);
        }

        const uint64_t candM = ballot(C != kNoKey);
        const uint64_t candM1 = ROW2 ? ballot(C1 != kNoKey) : 0ull;
        const int nc = __popcll(candM) + __popcll(candM1);
        bool changed = false;
        // candidates from both halves of a row go through the insertion, whatever their number
        const bool bothHalves = ROW2 && candM != 0 && candM1 != 0;
        if (nc > 0 && (nc <= kInsertMax || bothHalves)) {
            int fs = fsize;
            for (int h = 0; h < (ROW2 ? 2 : 1); h++) {
            const uint64_t Ch = h ? C1 : C;
            uint64_t mm = h ? candM1 : candM;
            while (mm) {
                const int j = __ffsll((unsigned long long)mm) - 1;
                mm &= mm - 1;
                const uint64_t x = readlane64(Ch, j);
                const uint64_t xmin = x & ~kTagMask, xmax = x | kTagMask;
                const int pos = __popcll(ballot(F0 <= xmax)) + __popcll(ballot(F1 <= xmax));
                if (pos >= Ls) continue;
                const int lower = __popcll(ballot(F0 < xmin)) + __popcll(ballot(F1 < xmin));
                uint64_t xe = x;
                if (pos != lower) {
                    const int cF = __popcll(ballot(F0 >= xmin && F0 <= xmax && !(F0 & kCand))) +
                                   __popcll(ballot(F1 >= xmin && F1 <= xmax && !(F1 & kCand)));
                    // earlier copies: in this half below lane j, and (second half) all of
                    // the first half's
                    const int rep = __popcll(ballot(Ch == x) & ((1ull << j) - 1)) +
                                    (h ? __popcll(ballot(C == x)) : 0);
                    if (rep < cF) continue;
                    xe = xmin | kCand | (uint32_t)max(0, (int)(x & 0x7Fu) - rep);
                }
                // shift positions >= pos up by one (F0's lane 63 moves to F1's lane 0)
                const uint64_t carry = readlane64(F0, 63);
                uint64_t pv1 = ((uint64_t)dppMov<0x138>((uint32_t)(F1 >> 32)) << 32) |
                               dppMov<0x138>((uint32_t)F1);
                if (lane == 0) pv1 = carry;
                if (pos < 64) {
                    const uint64_t pv0 = ((uint64_t)dppMov<0x138>((uint32_t)(F0 >> 32)) << 32) |
                                         dppMov<0x138>((uint32_t)F0);
                    F1 = pv1;
                    if (lane > pos) F0 = pv0;
                    if (lane == pos) F0 = xe;
                } else {
                    const int p1 = pos - 64;
                    if (lane > p1) F1 = pv1;
                    if (lane == p1) F1 = xe;
                }
                fs = min(Ls, fs + 1);
                changed = true;
            }
            }
            if (changed) {
                const uint32_t vis0 = (F0 & kCand) ? ((F0 & 0x7Fu) ? 1u : 0u) : (uint32_t)(F0 & 1u);
                const uint32_t vis1 = (F1 & kCand) ? ((F1 & 0x7Fu) ? 1u : 0u) : (uint32_t)(F1 & 1u);
                F0 = lane < fs ? ((F0 & ~kTagMask) | ((uint64_t)lane << 1) | vis0) : kNoKey;
                F1 = lane + 64 < fs ? ((F1 & ~kTagMask) | ((uint64_t)(lane + 64) << 1) | vis1) : kNoKey;
                fsize = fs;
            }
            P.mark(kProfMerge);
        } else if (nc > kInsertMax) {
            // (rows of two halves: all candidates lie in one half here)
            const uint64_t Cd = sortWave<true>(ROW2 && candM == 0 ? C1 : C, lane);
            P.mark(kProfSort);
            // F0 (ascending) with the candidates (descending), then the upper half of that
            // (reversed to descending) with F1
            uint64_t lo1 = F0 < Cd ? F0 : Cd;
            uint64_t hi1 = F0 < Cd ? Cd : F0;
            lo1 = mergeHalf(lo1, lane);
            hi1 = reverseWave(mergeHalf(hi1, lane), lane);
            uint64_t lo2 = F1 < hi1 ? F1 : hi1;
            uint64_t hi2 = F1 < hi1 ? hi1 : F1;
            lo2 = mergeHalf(lo2, lane);
            const UnionMarks m1 = unionMarks(lo1, prevLane64(lo1, kNoKey, lane), 0, 0, 0, lane, upTo);
            const UnionMarks m2 = unionMarks(lo2, prevLane64(lo2, readlane64(lo1, 63), lane), 64,
                                             m1.lastRun, m1.lastC, lane, upTo);
            const int v1 = __popcll(ballot(lo1 != kNoKey)), v2 = __popcll(ballot(lo2 != kNoKey));
            int newSize;
            if ((m1.drop | m2.drop) == 0) {
                newSize = min(Ls, v1 + v2);
                F0 = lo1;
                F1 = lo2;
            } else {
                hi2 = mergeHalf(hi2, lane);
                const UnionMarks m3 = unionMarks(hi2, prevLane64(hi2, readlane64(lo2, 63), lane), 128,
                                                 m2.lastRun, m2.lastC, lane, upTo);
                const int v3 = __popcll(ballot(hi2 != kNoKey));
                const int n1 = __popcll(m1.drop), n2 = __popcll(m2.drop);
                const int d1 = lane - (int)__popcll(m1.drop & lanesBelow);
                const int d2 = 64 + lane - n1 - (int)__popcll(m2.drop & lanesBelow);
                const int d3 = 128 + lane - n1 - n2 - (int)__popcll(m3.drop & lanesBelow);
                if (lo1 != kNoKey && !((m1.drop >> lane) & 1ull)) scratch[d1] = lo1;
                if (lo2 != kNoKey && !((m2.drop >> lane) & 1ull) && d2 < 128) scratch[d2] = lo2;
                if (hi2 != kNoKey && !((m3.drop >> lane) & 1ull) && d3 < 128) scratch[d3] = hi2;
                __builtin_amdgcn_wave_barrier();
                newSize = min(Ls, v1 + v2 + v3 - n1 - n2 - (int)__popcll(m3.drop));
                F0 = lane < newSize ? scratch[lane] : kNoKey;
                F1 = lane + 64 < newSize ? scratch[64 + lane] : kNoKey;
                __builtin_amdgcn_wave_barrier();
            }
            F0 = lane < newSize ? ((F0 & ~kTagMask) | ((uint64_t)lane << 1) | (F0 & 1u)) : kNoKey;
            F1 = lane + 64 < newSize ? ((F1 & ~kTagMask) | ((uint64_t)(lane + 64) << 1) | (F1 & 1u)) : kNoKey;
            fsize = newSize;
            changed = true;
            P.mark(kProfMerge);
        }

        if (changed && c.k > 0 && fsize > c.k) {
            const int dk = keyDist(frontierKey(F0, F1, c.k));
            const float t = (float)(c.cut * (double)dk);
            const float df0 = (float)keyDist(F0), df1 = (float)keyDist(F1);
            const uint64_t m0 = ballot(lane < fsize && ((t < df0) || (t == df0 && keyId(F0) > 0u)));
            const uint64_t m1 = ballot(lane + 64 < fsize && ((t < df1) || (t == df1 && keyId(F1) > 0u)));
            if (m0 | m1) {
                fsize = m0 ? __ffsll((unsigned long long)m0) - 1 : 64 + __ffsll((unsigned long long)m1) - 1;
                if (lane >= fsize) F0 = kNoKey;
                if (lane + 64 >= fsize) F1 = kNoKey;
            }
        }
        P.mark(kProfCut);

        const uint64_t nv0 = ballot(lane < fsize && !(F0 & 1u));
        const uint64_t nv1 = ballot(lane + 64 < fsize && !(F1 & 1u));
        u = nv0 ? __ffsll((unsigned long long)nv0) - 1 : nv1 ? 64 + __ffsll((unsigned long long)nv1) - 1 : -1;
        P.mark(kProfNext);
    }

    if (fsize < c.k) flags |= kFlagShort;
    uint32_t* oi = L.outIds + (size_t)q * L.outStride;
    int32_t* od = L.outDists + (size_t)q * L.outStride;
    for (int j = lane; j < (int)L.outStride; j += 64) {
        if (j < c.k && j < fsize) {
            const uint64_t kv = j < 64 ? F0 : F1;
            oi[j] = keyId(kv);
            od[j] = keyDist(kv);
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
    P.mark(kProfOutput);
}

// FAST - 1, bit 0: tables below 4 GiB (32-bit offsets), bit 1: next-row prefetch, bit 2:
// rows carry their squared norm: every query of the launch takes searchFast (Ls <= 64,
// rows <= 64, layout flatFastLdsLayout).  FAST 9..12: searchFast2.  FAST 0: queries with
// Ls <= 64 on rows of at most 64 entries take searchFast when the launch holds a filter
// of at least kFastFilterBits in LDS; the others take searchOne.  FAST -1: searchOne with
// the visited table in HBM (L.gvkey / L.gvcnt, L.capV slots per wave).
template <int MAXCH, int FAST>
__device__ __forceinline__ void drainQueries(const FlatLaunch& L) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = laneId();
    const int waveInBlock = threadIdx.x >> 6;
    const uint32_t wavesPerBlock = blockDim.x >> 6;
    const uint32_t waveGlobal = blockIdx.x * wavesPerBlock + waveInBlock;
    unsigned char* lds = smem + (size_t)waveInBlock * L.ldsPerWave;
    uint32_t* gfilt = L.gfilt ? L.gfilt + ((size_t)waveGlobal << L.gbitsMax) : nullptr;
    constexpr bool VG = FAST < 0;
    uint32_t* gvk = VG ? L.gvkey + (size_t)waveGlobal * L.capV : nullptr;
    unsigned char* gvc = VG ? L.gvcnt + (size_t)waveGlobal * L.capV : nullptr;

    Prof P;
    P.start();
    // queue mode ends on kQueueEmpty, strided mode past the last query
    const uint32_t qEnd = L.queue ? kQueueEmpty : L.nq;
    uint32_t q = L.queue ? takeQuery(L.queue, L.order, L.nq, lane) : waveGlobal;
    while (q < qEnd) {
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
        makeUniform(c);
        const int fb = flatFilterBits(c.Ls);
        if constexpr (FAST >= 21) {
            if (c.Ls < 1 || c.Ls > L.LsMax || c.k < 0 || (uint32_t)c.k > L.outStride || fb > L.gbitsMax) {
                rejectQuery(L, q, lane);
            } else {
                searchFast2<((FAST - 21) & 1) != 0, ((FAST - 21) & 2) != 0, true>(L, lds, q, c, gfilt, P);
            }
        } else if constexpr (FAST >= 17) {
            if (c.Ls < 1 || c.Ls > L.LsMax || c.k < 0 || (uint32_t)c.k > L.outStride) {
                rejectQuery(L, q, lane);
            } else {
                searchFast<((FAST - 17) & 1) != 0, 0, ((FAST - 17) & 2) != 0, true>(L, lds, q, c, P);
            }
        } else if constexpr (FAST >= 13) {
            if (c.Ls < 1 || c.Ls > L.LsMax || c.k < 0 || (uint32_t)c.k > L.outStride) {
                rejectQuery(L, q, lane);
            } else {
                searchFast<((FAST - 13) & 1) != 0, 2, ((FAST - 13) & 2) != 0>(L, lds, q, c, P);
            }
        } else if constexpr (FAST >= 9) {
            if (c.Ls < 1 || c.Ls > L.LsMax || c.k < 0 || (uint32_t)c.k > L.outStride || fb > L.gbitsMax) {
                rejectQuery(L, q, lane);
            } else {
                searchFast2<((FAST - 9) & 1) != 0, ((FAST - 9) & 2) != 0>(L, lds, q, c, gfilt, P);
            }
        } else if constexpr (FAST > 0) {
            if (c.Ls < 1 || c.Ls > L.LsMax || c.k < 0 || (uint32_t)c.k > L.outStride) {
                rejectQuery(L, q, lane);
            } else {
                searchFast<((FAST - 1) & 1) != 0, ((FAST - 1) & 2) != 0 ? 1 : 0, ((FAST - 1) & 4) != 0>(L, lds, q, c, P);
            }
        } else {
            if (c.Ls < 1 || c.Ls > L.LsMax || c.k < 0 || (uint32_t)c.k > L.outStride ||
                (fb > L.fbitsLds && (L.gfilt == nullptr || fb > L.gbitsMax))) {
                rejectQuery(L, q, lane);
            } else if (!VG && MAXCH == 1 && c.Ls <= kFastLsMax && L.fbitsLds >= kFastFilterBits) {
                if (L.normsInRow) searchFast<false, 0, true>(L, lds, q, c, P);
                else searchFast<false, 0, false>(L, lds, q, c, P);
            } else if (fb <= L.fbitsLds) {
                searchOne<MAXCH, false, VG>(L, lds, q, c, nullptr, gvk, gvc, P);
            } else {
                searchOne<MAXCH, true, VG>(L, lds, q, c, gfilt, gvk, gvc, P);
            }
        }
        q = L.queue ? takeQuery(L.queue, L.order, L.nq, lane) : q + L.totalWaves;
    }
    P.flush(lane);
    if (L.queue) finishQueue(L.queue, L.totalWaves, lane);
}

template <int MAXCH>
__global__ __launch_bounds__(256) void beamFlatKernel(FlatLaunch L) {
    drainQueries<MAXCH, 0>(L);
}

// Long beams: with the visited table in HBM a wavefront needs only its frontier and a few
// small buffers in LDS, so up to four wavefronts share a SIMD (128 VGPRs; rows of more
// than 64 entries need more and get three); each hop is a chain of dependent HBM round
// trips and the wavefronts hide one another's.  On rows of at most 64 entries the chain
// is shortened where it can be: the row of the guessed next node is prefetched, the seen
// filter's slot reads overlap the visit, and the vector loads go out before the filter's
// shared-slot resolution (five wavefronts per SIMD at 96 VGPRs spill and run 2% slower).
template <int MAXCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MAXCH == 1 ? 4 : 3, 8))) void beamFlatWideKernel(FlatLaunch L) {
    drainQueries<MAXCH, -1>(L);
}

// The register-resident kernel is held to 96 VGPRs so five wavefronts share a SIMD (its
// LDS, 7.7 KB per wavefront at Ls <= 25, allows as many): the hop is a chain of
// dependent gathers and LDS round trips, and the extra wavefront hides them.
template <int FAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void beamFlatFastKernel(FlatLaunch L) {
    drainQueries<(FAST >= 17 ? 2 : 1), FAST>(L);
}

// The two-register form: its LDS (6-11 KB per wavefront) allows about four wavefronts
// per SIMD, so it is held to 128 VGPRs.
template <int FAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void beamFlatFast2Kernel(FlatLaunch L) {
    drainQueries<(FAST >= 21 ? 2 : 1), FAST>(L);
}

// The cooperative general form: blocks of WPQ wavefronts, one query per block at a time
// (the queue ticket goes to the block through LDS), HBM scratch per block.
template <int MAXCH, int WPQ, bool VG>
__global__ __launch_bounds__(64 * WPQ) void beamFlatCoopKernel(FlatLaunch L) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t ticket;
    const int lane = laneId();
    const int w = threadIdx.x >> 6;
    const uint32_t slot = blockIdx.x;
    uint32_t* gfilt = L.gfilt ? L.gfilt + ((size_t)slot << L.gbitsMax) : nullptr;
    uint32_t* gvk = VG ? L.gvkey + (size_t)slot * L.capV : nullptr;
    unsigned char* gvc = VG ? L.gvcnt + (size_t)slot * L.capV : nullptr;
    const uint32_t qEnd = L.queue ? kQueueEmpty : L.nq;
    auto next = [&](uint32_t prev, bool first) -> uint32_t {
        if (!L.queue) return first ? slot : prev + L.totalWaves;
        if (w == 0) {
            const uint32_t t = takeQuery(L.queue, L.order, L.nq, lane);
            if (lane == 0) ticket = t;
        }
        __syncthreads();
        const uint32_t t = ticket;
        __syncthreads();
        return t;
    };
    uint32_t q = next(0, true);
    while (q < qEnd) {
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
        makeUniform(c);
        const int fb = flatFilterBits(c.Ls);
        if (c.Ls < 1 || c.Ls > L.LsMax || c.k < 0 || (uint32_t)c.k > L.outStride ||
            (fb > L.fbitsLds && (L.gfilt == nullptr || fb > L.gbitsMax))) {
            if (w == 0) rejectQuery(L, q, lane);
        } else if (fb <= L.fbitsLds) {
            searchCoop<MAXCH, false, VG, WPQ>(L, smem, q, c, nullptr, gvk, gvc, w);
        } else {
            searchCoop<MAXCH, true, VG, WPQ>(L, smem, q, c, gfilt, gvk, gvc, w);
        }
        q = next(q, false);
    }
    if (L.queue && w == 0) finishQueue(L.queue, L.totalWaves, lane);
}

template <int WPQ, bool VG>
hipError_t launchCoop(const FlatLaunch& L, int maxch, int blocks, hipStream_t stream) {
    const dim3 grid(blocks), block(64 * WPQ);
    switch (maxch) {
        case 1: hipLaunchKernelGGL((beamFlatCoopKernel<1, WPQ, VG>), grid, block, L.ldsPerWave, stream, L); break;
        case 2: hipLaunchKernelGGL((beamFlatCoopKernel<2, WPQ, VG>), grid, block, L.ldsPerWave, stream, L); break;
        case 4: hipLaunchKernelGGL((beamFlatCoopKernel<4, WPQ, VG>), grid, block, L.ldsPerWave, stream, L); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int WPQ, bool VG>
hipError_t occCoop(int maxch, size_t lds, int* bpc) {
    switch (maxch) {
        case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, beamFlatCoopKernel<1, WPQ, VG>, 64 * WPQ, lds);
        case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, beamFlatCoopKernel<2, WPQ, VG>, 64 * WPQ, lds);
        case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, beamFlatCoopKernel<4, WPQ, VG>, 64 * WPQ, lds);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

#ifdef PANN_PROFILE
namespace dev {
__device__ unsigned long long g_profile[kProfSections];
}

hipError_t profileCounters(uint64_t* out, int n, bool reset) {
    unsigned long long h[kProfSections] = {};
    hipError_t e = hipMemcpyFromSymbol(h, HIP_SYMBOL(dev::g_profile), sizeof(h));
    for (int i = 0; i < n && i < kProfSections; i++) out[i] = h[i];
    if (e == hipSuccess && reset) {
        unsigned long long z[kProfSections] = {};
        e = hipMemcpyToSymbol(HIP_SYMBOL(dev::g_profile), z, sizeof(z));
    }
    return e;
}
#else
hipError_t profileCounters(uint64_t*, int, bool) { return hipErrorNotSupported; }
#endif

hipError_t launchBeamFlat(const FlatLaunch& L, int maxch, int wavesPerBlock, int blocks,
                          hipStream_t stream) {
    if (L.wpq > 1) {
        if (L.fast > 0 || wavesPerBlock != L.wpq) return hipErrorInvalidValue;
        if (L.fast < 0 && (L.gvkey == nullptr || L.gvcnt == nullptr)) return hipErrorInvalidValue;
        switch (L.wpq * 2 + (L.fast < 0 ? 1 : 0)) {
            case 4: return launchCoop<2, false>(L, maxch, blocks, stream);
            case 5: return launchCoop<2, true>(L, maxch, blocks, stream);
            case 8: return launchCoop<4, false>(L, maxch, blocks, stream);
            case 9: return launchCoop<4, true>(L, maxch, blocks, stream);
            default: return hipErrorInvalidValue;
        }
    }
    const size_t lds = (size_t)L.ldsPerWave * wavesPerBlock;
    const dim3 grid(blocks), block(64 * wavesPerBlock);
    if (L.fast < 0) {
        if (L.gvkey == nullptr || L.gvcnt == nullptr) return hipErrorInvalidValue;
        switch (maxch) {
            case 1: hipLaunchKernelGGL((beamFlatWideKernel<1>), grid, block, lds, stream, L); break;
            case 2: hipLaunchKernelGGL((beamFlatWideKernel<2>), grid, block, lds, stream, L); break;
            case 4: hipLaunchKernelGGL((beamFlatWideKernel<4>), grid, block, lds, stream, L); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (L.fast) {
        if (L.fast >= 21) {
            // two registers per lane, rows of up to 128 entries in two halves
            if (maxch != 2 || L.LsMax > kFast2LsMax || L.gfilt == nullptr) return hipErrorInvalidValue;
            switch (L.fast + (L.normsInRow ? 2 : 0)) {
                case 21: hipLaunchKernelGGL((beamFlatFast2Kernel<21>), grid, block, lds, stream, L); break;
                case 22: hipLaunchKernelGGL((beamFlatFast2Kernel<22>), grid, block, lds, stream, L); break;
                case 23: hipLaunchKernelGGL((beamFlatFast2Kernel<23>), grid, block, lds, stream, L); break;
                case 24: hipLaunchKernelGGL((beamFlatFast2Kernel<24>), grid, block, lds, stream, L); break;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        }
        if (L.fast >= 17) {
            // rows of up to 128 entries in two halves
            if (maxch != 2 || L.LsMax > kFastLsMax) return hipErrorInvalidValue;
            switch (L.fast + (L.normsInRow ? 2 : 0)) {
                case 17: hipLaunchKernelGGL((beamFlatFastKernel<17>), grid, block, lds, stream, L); break;
                case 18: hipLaunchKernelGGL((beamFlatFastKernel<18>), grid, block, lds, stream, L); break;
                case 19: hipLaunchKernelGGL((beamFlatFastKernel<19>), grid, block, lds, stream, L); break;
                case 20: hipLaunchKernelGGL((beamFlatFastKernel<20>), grid, block, lds, stream, L); break;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        }
        if (maxch != 1) return hipErrorInvalidValue;
        if (L.fast >= 13) {
            if (L.LsMax > kFastLsMax) return hipErrorInvalidValue;
            switch (L.fast + (L.normsInRow ? 2 : 0)) {
                case 13: hipLaunchKernelGGL((beamFlatFastKernel<13>), grid, block, lds, stream, L); break;
                case 14: hipLaunchKernelGGL((beamFlatFastKernel<14>), grid, block, lds, stream, L); break;
                case 15: hipLaunchKernelGGL((beamFlatFastKernel<15>), grid, block, lds, stream, L); break;
                case 16: hipLaunchKernelGGL((beamFlatFastKernel<16>), grid, block, lds, stream, L); break;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        }
        if (L.fast >= 9) {
            if (L.LsMax > kFast2LsMax || L.gfilt == nullptr) return hipErrorInvalidValue;
            switch (L.fast + (L.normsInRow ? 2 : 0)) {
                case 9: hipLaunchKernelGGL((beamFlatFast2Kernel<9>), grid, block, lds, stream, L); break;
                case 10: hipLaunchKernelGGL((beamFlatFast2Kernel<10>), grid, block, lds, stream, L); break;
                case 11: hipLaunchKernelGGL((beamFlatFast2Kernel<11>), grid, block, lds, stream, L); break;
                case 12: hipLaunchKernelGGL((beamFlatFast2Kernel<12>), grid, block, lds, stream, L); break;
                default: return hipErrorInvalidValue;
            }
            return hipGetLastError();
        }
        if (L.LsMax > kFastLsMax) return hipErrorInvalidValue;
        const int mode = L.fast + (L.normsInRow ? 4 : 0);
        switch (mode) {
            case 1: hipLaunchKernelGGL((beamFlatFastKernel<1>), grid, block, lds, stream, L); break;
            case 2: hipLaunchKernelGGL((beamFlatFastKernel<2>), grid, block, lds, stream, L); break;
            case 3: hipLaunchKernelGGL((beamFlatFastKernel<3>), grid, block, lds, stream, L); break;
            case 4: hipLaunchKernelGGL((beamFlatFastKernel<4>), grid, block, lds, stream, L); break;
            case 5: hipLaunchKernelGGL((beamFlatFastKernel<5>), grid, block, lds, stream, L); break;
            case 6: hipLaunchKernelGGL((beamFlatFastKernel<6>), grid, block, lds, stream, L); break;
            case 7: hipLaunchKernelGGL((beamFlatFastKernel<7>), grid, block, lds, stream, L); break;
            case 8: hipLaunchKernelGGL((beamFlatFastKernel<8>), grid, block, lds, stream, L); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (maxch) {
        case 1: hipLaunchKernelGGL((beamFlatKernel<1>), grid, block, lds, stream, L); break;
        case 2: hipLaunchKernelGGL((beamFlatKernel<2>), grid, block, lds, stream, L); break;
        case 4: hipLaunchKernelGGL((beamFlatKernel<4>), grid, block, lds, stream, L); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t beamFlatCoopOccupancy(int maxch, int wpq, bool hbmVisited, size_t ldsBytes, int* blocksPerCU) {
    switch (wpq * 2 + (hbmVisited ? 1 : 0)) {
        case 4: return occCoop<2, false>(maxch, ldsBytes, blocksPerCU);
        case 5: return occCoop<2, true>(maxch, ldsBytes, blocksPerCU);
        case 8: return occCoop<4, false>(maxch, ldsBytes, blocksPerCU);
        case 9: return occCoop<4, true>(maxch, ldsBytes, blocksPerCU);
        default: return hipErrorInvalidValue;
    }
}

hipError_t beamFlatOccupancy(int maxch, int fast, int wavesPerBlock, size_t ldsBytes,
                             int* blocksPerCU) {
    const int threads = 64 * wavesPerBlock;
    if (fast < 0) {
        switch (maxch) {
            case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatWideKernel<1>, threads, ldsBytes);
            case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatWideKernel<2>, threads, ldsBytes);
            case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatWideKernel<4>, threads, ldsBytes);
            default: return hipErrorInvalidValue;
        }
    }
    if (fast >= 21) {
        if (maxch != 2) return hipErrorInvalidValue;
        return fast == 21 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFast2Kernel<21>, threads, ldsBytes)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFast2Kernel<22>, threads, ldsBytes);
    }
    if (fast >= 17) {
        if (maxch != 2) return hipErrorInvalidValue;
        return fast == 17 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFastKernel<17>, threads, ldsBytes)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFastKernel<18>, threads, ldsBytes);
    }
    if (fast >= 13) {
        if (maxch != 1) return hipErrorInvalidValue;
        return fast == 13 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFastKernel<13>, threads, ldsBytes)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFastKernel<14>, threads, ldsBytes);
    }
    if (fast >= 9) {
        if (maxch != 1) return hipErrorInvalidValue;
        return fast == 9 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFast2Kernel<9>, threads, ldsBytes)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFast2Kernel<10>, threads, ldsBytes);
    }
    if (fast) {
        if (maxch != 1) return hipErrorInvalidValue;
        switch (fast) {
            case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFastKernel<1>, threads, ldsBytes);
            case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFastKernel<2>, threads, ldsBytes);
            case 3: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFastKernel<3>, threads, ldsBytes);
            case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatFastKernel<4>, threads, ldsBytes);
            default: return hipErrorInvalidValue;
        }
    }
    switch (maxch) {
        case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatKernel<1>, threads, ldsBytes);
        case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatKernel<2>, threads, ldsBytes);
        case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCU, beamFlatKernel<4>, threads, ldsBytes);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pann
