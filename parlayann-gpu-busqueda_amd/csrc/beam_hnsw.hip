// beam_hnsw.hip -- ParlayANN HNSW::search for gfx950: one 64-lane wavefront per query,
// upper-layer greedy descent (ef = 1) and the layer-0 beam (ef = Ls) in one kernel.
//
// Semantics follow the reference's restatement (cuda/busqueda_hnsw.cu:179-457):
//   * per layer: fresh lossy visited filter (hash64 marks entry points, hash64_2 the
//     neighbours; 2 bits above layer 0, 2*ceil(log2 ef)-2 on layer 0), fresh is_inw set
//     (the ids that entered W on the layer);
//   * cand ordered and unique by distance, W (the workset) ordered by (distance, id);
//   * loop: stop when cand's nearest is farther than W's farthest or the visit limit is
//     hit; expand the nearest candidate; neighbours are accepted in row order when
//     |W| < ef or d < W's farthest and they never entered W before in this layer;
//   * the next layer starts from W's farthest element; output W[0:k] as base point ids.
// Layout choices of this file: W and cand live in registers as wave-wide sorted lists
// (element e = r*64 + lane), so an insertion position is one ballot popcount and an
// insertion a whole-wave DPP lane shift; is_inw is not stored but read off W (see the
// acceptance in expandNode); the sequential acceptance only visits neighbours that can
// still pass the threshold, found by ballot.  Rows of at most 64 entries keep entry e on
// lane e and issue the vector loads of every entry that may pass the seen filter before
// the filter's shared-slot resolution, so the filter and the gathers overlap.
#include "device_common.h"

namespace pann {
namespace {

using namespace dev;

// Sorted u64 keys (d << 32 | id) spread over the wave: element e = r*64 + lane in v[r].
// SKIP: operations touch only the registers that hold live elements (see below).
template <int WR, bool SKIP>
struct WaveList {
    uint64_t v[WR];
    int size;

    __device__ void clear() {
#pragma unroll
        for (int r = 0; r < WR; r++) v[r] = ~0ull;
        size = 0;
    }
    __device__ uint64_t get(int e) const {  // e uniform
        uint64_t x = ~0ull;
#pragma unroll
        for (int r = 0; r < WR; r++)
            if (r == (e >> 6)) x = readlane64(v[r], e & 63);
        return x;
    }
    // Elements at and past `size` are never read, so with SKIP every operation touches
    // only the registers that hold live elements (size and pos are wave-uniform: the
    // skips are scalar branches).  Only lists of seven and eight registers gain from the
    // branches (1M HNSW ef 256 / 350 ran 1-2% slower with them on five and six registers,
    // ef 150 / 200 6% slower on three and four), and the kernels for rows wider than 64
    // entries run out of registers with them, so the others shift and compare every
    // register.
    __device__ int countLess(uint64_t key, int lane) const {
        int c = 0;
        const int sz = __builtin_amdgcn_readfirstlane(size);
#pragma unroll
        for (int r = 0; r < WR; r++) {
            if (SKIP && r * 64 >= sz) break;
            c += __popcll(ballot(r * 64 + lane < size && v[r] < key));
        }
        return c;
    }
    __device__ bool hasDist(uint32_t d, int lane) const {
        uint64_t m = 0;
        const int sz = __builtin_amdgcn_readfirstlane(size);
#pragma unroll
        for (int r = 0; r < WR; r++) {
            if (SKIP && r * 64 >= sz) break;
            m |= ballot(r * 64 + lane < size && (uint32_t)(v[r] >> 32) == d);
        }
        return m != 0;
    }
    // Lane shifts are whole-wave DPP moves (no LDS round trip).  With SKIP, registers
    // wholly below pos keep their elements and those past the new last element (index
    // size) hold none, so neither is shifted.
    __device__ void insertAt(int pos, uint64_t key, int lane) {
        const int sz = __builtin_amdgcn_readfirstlane(size);
        const int ps = __builtin_amdgcn_readfirstlane(pos);
#pragma unroll
        for (int r = WR - 1; r >= 0; r--) {
            if (SKIP && r * 64 > sz) continue;
            if (SKIP && r * 64 + 63 < ps) break;
            const uint64_t first = r > 0 ? readlane64(v[r > 0 ? r - 1 : 0], 63) : ~0ull;
            const uint64_t prev = prevLane64(v[r], first, lane);
            const int e = r * 64 + lane;
            v[r] = e > pos ? prev : (e == pos ? key : v[r]);
        }
        size++;
    }
    __device__ void popFront(int lane) {
        const int sz = __builtin_amdgcn_readfirstlane(size);
#pragma unroll
        for (int r = 0; r < WR; r++) {
            if (SKIP && r * 64 >= sz) break;
            const uint64_t last = r + 1 < WR ? readlane64(v[r + 1 < WR ? r + 1 : r], 0) : ~0ull;
            v[r] = nextLane64(v[r], last, lane);
        }
        size--;
    }
};

template <int WR, int MAXCH>
constexpr bool kListSkips = MAXCH == 1 && WR >= 7;

struct LayerState {
    int comps, edges, hops, flags;
};

// Expands node u on `layer`: row gather, seen filter, distances, in-order acceptance.
template <int WR, int MAXCH, bool G>
__device__ void expandNode(const HnswLaunch& L, unsigned char* lds, uint32_t* filt, uint32_t fmask,
                           uint32_t u, int layer, int ef, const int4& qv, int qq,
                           WaveList<WR, kListSkips<WR, MAXCH>>& W, WaveList<WR, kListSkips<WR, MAXCH>>& cand, LayerState& st, int lane,
                           uint64_t lanesBelow) {
    const int grp = lane >> 3, sub = lane & 7;
    uint32_t* pend = reinterpret_cast<uint32_t*>(lds + L.lds.pend);
    unsigned char* tag = lds + L.lds.tag;
    const uint32_t width = layer > 0 ? L.GA : L.G0;
    const uint32_t* row = layer > 0 ? L.adjUp + ((size_t)L.upStart[u] + layer - 1) * L.GA
                                    : L.adj0 + (size_t)u * L.G0;
    const uint32_t nch = (width + 63) / 64;
    uint32_t a[MAXCH];
#pragma unroll
    for (int ch = 0; ch < MAXCH; ch++) {
        const uint32_t j = ch * 64 + lane;
        a[ch] = (ch < (int)nch && j < width) ? row[j] : kEmpty;
    }
    int dist[MAXCH];
    uint32_t id[MAXCH];
    bool valid[MAXCH];
    if constexpr (MAXCH == 1) {
        // Rows of at most 64 entries: entry e stays on lane e.  Every entry that may pass
        // the seen filter has its vector load issued once the slots are read (group g loads
        // entry 8g + r in round r), and the distance of entry e lands on lane e; the entries
        // kept although not speculated (an earlier lane took their slot) load afterwards.
        const uint32_t a0 = a[0];
        const bool part = a0 != kEmpty;
        st.edges += __popcll(ballot(part));
        const SeenProbe sp = seenFilterBegin<G>(filt, tag, fmask, a0, part, lane);
        const bool spec = part && (sp.coll || sp.prev != a0);
        const uint32_t gmask = (uint32_t)(ballot(spec) >> (8 * grp)) & 0xFFu;
        // rounds of entries that are not loaded keep whatever the registers hold: their
        // distances are computed and thrown away.  The ids are broadcast again for the late
        // loads rather than held (eight registers more would spill at the occupancy bound).
        int4 v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            asm("" : "=v"(v[r].x), "=v"(v[r].y), "=v"(v[r].z), "=v"(v[r].w));
            if ((gmask >> r) & 1u)
                v[r] = *reinterpret_cast<const int4*>(L.vec + (size_t)groupBroadcast8(a0, r) * kVecBytes + sub * 16);
        }
        const bool keep = seenFilterEnd<G>(filt, sp, a0, part, lane, lanesBelow);
        const uint64_t keepM = ballot(keep);
        st.comps += __popcll(keepM);
        const uint64_t late = keepM & ~ballot(spec);
        if (late) {
            const uint32_t lmask = (uint32_t)(late >> (8 * grp)) & 0xFFu;
#pragma unroll
            for (int r = 0; r < 8; r++)
                if ((lmask >> r) & 1u)
                    v[r] = *reinterpret_cast<const int4*>(L.vec + (size_t)groupBroadcast8(a0, r) * kVecBytes + sub * 16);
        }
        int part8[8];
        const int normMask = (L.normsInRow && sub == 7) ? -1 : 0;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            v[r].w &= ~normMask;
            part8[r] = slicePartial(v[r], qv);
        }
        valid[0] = keep;
        dist[0] = qq + transposeReduce8(part8, lane);  // distance of entry `lane`
        id[0] = a0;
    } else {
        uint32_t nP = 0;
#pragma unroll
        for (int ch = 0; ch < MAXCH; ch++) {
            if (ch >= (int)nch) break;
            const bool part = a[ch] != kEmpty;
            st.edges += __popcll(ballot(part));
            const bool keep = seenFilterChunk<G>(filt, tag, fmask, a[ch], part, lane, lanesBelow);
            const uint64_t km = ballot(keep);
            if (keep) pend[nP + mbcnt(km)] = a[ch];
            nP += (uint32_t)__popcll(km);
            __builtin_amdgcn_wave_barrier();
        }
        st.comps += (int)nP;


        // distances, then move candidate j to lane j of block j/64
#pragma unroll
        for (int b = 0; b < MAXCH; b++) {
            valid[b] = false;
            dist[b] = 0;
            id[b] = kEmpty;
            if ((uint32_t)b * 64 >= nP) continue;
            const uint32_t nb = min(64u, nP - (uint32_t)b * 64);
            const uint32_t rounds = (nb + 7) / 8;
            int4 v[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                v[r] = make_int4(0, 0, 0, 0);
                const uint32_t i = (uint32_t)r * 8 + grp;
                if ((uint32_t)r < rounds && i < nb)
                    v[r] = *reinterpret_cast<const int4*>(L.vec + (size_t)pend[b * 64 + i] * kVecBytes + sub * 16);
            }
            int part8[8];
            const int normMask = (L.normsInRow && sub == 7) ? -1 : 0;
#pragma unroll
            for (int r = 0; r < 8; r++) {
                v[r].w &= ~normMask;
                part8[r] = slicePartial(v[r], qv);
            }
            const int tot = transposeReduce8(part8, lane);
            // lane l holds candidate (l & 7) * 8 + (l >> 3); the map is its own inverse
            const int mine = __shfl(tot, (lane & 7) * 8 + (lane >> 3), 64);
            if ((uint32_t)lane < nb) {
                valid[b] = true;
                dist[b] = qq + mine;
                id[b] = pend[b * 64 + lane];
            }
        }

    }

    // In-order acceptance; only neighbours that can still pass are visited.  is_inw (the
    // ids that ever entered W on this layer) is not kept: a neighbour that passes the
    // threshold is in is_inw exactly when it is in W now.  W only grows until it holds ef
    // entries and then keeps ef, each insertion dropping the farthest, so W's farthest
    // distance never rises after the first drop; an id dropped from W lay at or beyond it,
    // cannot pass `d < farthest` again, and every other id of is_inw is still in W, where
    // its key (d, id) is found at its insertion position.
    int pos = -1;
    while (true) {
        const bool full = W.size >= ef;
        const uint32_t wl = (uint32_t)(W.get(W.size - 1) >> 32);
        int j = -1;
#pragma unroll
        for (int b = 0; b < MAXCH; b++) {
            if (j >= 0) break;
            const int e = b * 64 + lane;
            const uint64_t m = ballot(valid[b] && e > pos && (!full || (uint32_t)dist[b] < wl));
            if (m) j = b * 64 + __ffsll((unsigned long long)m) - 1;
        }
        if (j < 0) break;
        pos = j;
        int dj = 0;
        uint32_t idj = 0;
#pragma unroll
        for (int b = 0; b < MAXCH; b++)
            if (b == (j >> 6)) {
                dj = readlane(dist[b], j & 63);
                idj = (uint32_t)readlane((int)id[b], j & 63);
            }
        const uint64_t key = ((uint64_t)(uint32_t)dj << 32) | idj;
        const int at = W.countLess(key, lane);
        if (at < W.size && W.get(at) == key) continue;  // in W: already in is_inw
        W.insertAt(at, key, lane);
        if (!cand.hasDist((uint32_t)dj, lane))
            cand.insertAt(cand.countLess((uint64_t)(uint32_t)dj << 32, lane), key, lane);
        if (W.size > ef) W.size--;
        if (cand.size > ef) cand.size--;
    }
    __builtin_amdgcn_wave_barrier();
}

// FGLOBAL: layer 0's seen filter in this wave's HBM region (gfilt) rather than in LDS.
template <int WR, int MAXCH, bool FGLOBAL>
__device__ void hnswOne(const HnswLaunch& L, unsigned char* lds, uint32_t* gfilt, uint32_t q,
                        int Ls, int k, int64_t visit) {
    const int lane = laneId();
    const int sub = lane & 7;
    const uint64_t lanesBelow = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t* ldsFilt = reinterpret_cast<uint32_t*>(lds + L.lds.filt);
    // rows may hold |p|^2 in bytes 124..127 (lane 7's .w): masked out of the products
    const int normMask = (L.normsInRow && sub == 7) ? -1 : 0;
    int4 qv = *reinterpret_cast<const int4*>(L.queries + (size_t)q * kVecBytes + sub * 16);
    qv.w &= ~normMask;
    const int qq = groupSum8(sliceNorm(qv));
    const int64_t lim0 = (visit > 0 && visit < (int64_t)L.n) ? visit : (int64_t)L.n;
    const int bits0 = hnswFilterBits(Ls);

    uint32_t eps[8];
    int numEps = L.numEntries;
#pragma unroll
    for (int e = 0; e < 8; e++) eps[e] = L.entries[e];
    LayerState st{0, 0, 0, 0};
    WaveList<WR, kListSkips<WR, MAXCH>> W, cand;
    W.clear();
    cand.clear();

    for (int layer = L.top; layer >= 0; layer--) {
        const int ef = layer > 0 ? 1 : Ls;
        const int bits = layer > 0 ? 2 : bits0;
        const uint32_t fmask = (1u << bits) - 1;
        const int64_t limit = layer > 0 ? INT64_MAX : lim0;
        const bool globalFilter = FGLOBAL && layer == 0;
        uint32_t* filt = globalFilter ? gfilt : ldsFilt;
        if (globalFilter) filtClear<true>(filt, bits, lane);
        else filtClear<false>(filt, bits, lane);
        W.clear();
        cand.clear();
        __builtin_amdgcn_wave_barrier();

        for (int e = 0; e < numEps; e++) {
            const uint32_t pe = eps[e];
            if (lane == 0) {
                const uint32_t slot = (uint32_t)hash64(pe) & fmask;
                if (globalFilter) filtStore<true>(filt, slot, pe);
                else filtStore<false>(filt, slot, pe);
            }
            if (globalFilter) __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
            const int d = readlane(groupDistance(L.vec, pe, sub, qv, qq, normMask), 0);
            const uint64_t key = ((uint64_t)(uint32_t)d << 32) | pe;
            if (!cand.hasDist((uint32_t)d, lane))
                cand.insertAt(cand.countLess((uint64_t)(uint32_t)d << 32, lane), key, lane);
            W.insertAt(W.countLess(key, lane), key, lane);
            st.comps++;
        }

        int64_t cnt = 0;
        while (cand.size > 0) {
            const uint64_t c0 = cand.get(0);
            const uint64_t wl = W.get(W.size - 1);
            if ((uint32_t)(c0 >> 32) > (uint32_t)(wl >> 32)) break;
            if (++cnt > limit) break;
            cand.popFront(lane);
            st.hops++;
            if (globalFilter)
                expandNode<WR, MAXCH, true>(L, lds, filt, fmask, (uint32_t)c0, layer, ef, qv, qq, W, cand,
                                            st, lane, lanesBelow);
            else
                expandNode<WR, MAXCH, false>(L, lds, filt, fmask, (uint32_t)c0, layer, ef, qv, qq, W, cand,
                                             st, lane, lanesBelow);
        }
        if (layer > 0) {
            eps[0] = (uint32_t)W.get(W.size - 1);
            numEps = 1;
        }
    }

    if (W.size < k) st.flags |= kFlagShort;
    uint32_t* oi = L.outIds + (size_t)q * L.outStride;
    int32_t* od = L.outDists + (size_t)q * L.outStride;
#pragma unroll
    for (int r = 0; r < WR; r++) {
        const int e = r * 64 + lane;
        if (e < (int)L.outStride) {
            if (e < k && e < W.size) {
                oi[e] = L.nodePoint[(uint32_t)W.v[r]];
                od[e] = (int32_t)(W.v[r] >> 32);
            } else {
                oi[e] = kEmpty;
                od[e] = 0x7FFFFFFF;
            }
        }
    }
    for (int e = WR * 64 + lane; e < (int)L.outStride; e += 64) {
        oi[e] = kEmpty;
        od[e] = 0x7FFFFFFF;
    }
    if (lane == 0) {
        if (L.outComps) L.outComps[q] = st.comps;
        if (L.outFlags) L.outFlags[q] = st.flags;
        if (L.outEdges) L.outEdges[q] = st.edges;
        if (L.outHops) L.outHops[q] = st.hops;
    }
    __builtin_amdgcn_wave_barrier();
}

// On rows of at most 64 entries, one-register lists are held to 80 VGPRs (six wavefronts
// per SIMD, 16 bytes of spills; 3-7% faster at ef 10-50 than at five), lists of two to four
// registers to 96 (five) and longer lists to 128 (four): the hop is a chain of dependent
// memory round trips that more wavefronts hide.  With the query parameters in scalar
// registers the kernels need 87 to 126 VGPRs up to six registers; held to 96, lists of two
// to four registers spill 8 / 24 / 48 bytes beyond the entry-point array and still run
// the 1M HNSW 8-14% faster at ef 64-200 than at four wavefronts (five and six registers at
// five wavefronts: 4-13% slower).  At 128 the seven- and eight-register lists spill 24 and
// 64 bytes and run ef=400 in 20.5 ms instead of 24.5 at 168 VGPRs (three wavefronts),
// ef=500 in 29.9 instead of 30.3.
template <int WR, int MAXCH>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MAXCH == 1 ? (WR == 1 ? 6 : WR <= 4 ? 5 : 4) : 1, 8)))
void beamHnswKernel(HnswLaunch L) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = laneId();
    const int waveInBlock = threadIdx.x >> 6;
    const uint32_t wavesPerBlock = blockDim.x >> 6;
    const uint32_t waveGlobal = blockIdx.x * wavesPerBlock + waveInBlock;
    unsigned char* lds = smem + (size_t)waveInBlock * L.ldsPerWave;
    uint32_t* gfilt = L.gfilt ? L.gfilt + ((size_t)waveGlobal << L.gbitsMax) : nullptr;

    // queue mode ends on kQueueEmpty, strided mode past the last query
    const uint32_t qEnd = L.queue ? kQueueEmpty : L.nq;
    uint32_t q = L.queue ? takeQuery(L.queue, L.order, L.nq, lane) : waveGlobal;
    while (q < qEnd) {
        int Ls = L.Ls, k = L.k;
        int64_t visit = L.visit;
        if (L.qp) {
            Ls = L.qp[q].Ls;
            k = L.qp[q].k;
            visit = L.qp[q].visit_limit;
        }
        // q and the parameters are the same in every lane.  Read through generic pointers
        // they would count as per-lane values, and with them ef, the list sizes and every
        // branch on those: saying so keeps them in scalar registers and the branches scalar.
        q = (uint32_t)__builtin_amdgcn_readfirstlane((int)q);
        Ls = __builtin_amdgcn_readfirstlane(Ls);
        k = __builtin_amdgcn_readfirstlane(k);
        {
            const uint64_t v = (uint64_t)visit;
            visit = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32)) << 32) |
                              (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v));
        }
        const int fb = hnswFilterBits(Ls);
        if (Ls < 1 || Ls > L.LsMax || k < 0 || (uint32_t)k > L.outStride ||
            (fb > L.fbitsLds && (L.gfilt == nullptr || fb > L.gbitsMax))) {
            rejectQuery(L, q, lane);
        } else if (fb <= L.fbitsLds) {
            hnswOne<WR, MAXCH, false>(L, lds, nullptr, q, Ls, k, visit);
        } else {
            hnswOne<WR, MAXCH, true>(L, lds, gfilt, q, Ls, k, visit);
        }
        q = L.queue ? takeQuery(L.queue, L.order, L.nq, lane) : q + L.totalWaves;
    }
    if (L.queue) finishQueue(L.queue, L.totalWaves, lane);
}

template <int WR, int MAXCH>
hipError_t launchT(const HnswLaunch& L, int wpb, int blocks, hipStream_t s) {
    const size_t lds = (size_t)L.ldsPerWave * wpb;
    hipLaunchKernelGGL((beamHnswKernel<WR, MAXCH>), dim3(blocks), dim3(64 * wpb), lds, s, L);
    return hipGetLastError();
}

template <int WR, int MAXCH>
hipError_t occT(int wpb, size_t lds, int* bpc) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, beamHnswKernel<WR, MAXCH>, 64 * wpb, lds);
}

}  // namespace

#define PANN_HNSW_DISPATCH(FN, ...)                                           \
    switch (wr * 10 + maxch) {                                                \
        case 11: return FN<1, 1>(__VA_ARGS__);                                \
        case 12: return FN<1, 2>(__VA_ARGS__);                                \
        case 14: return FN<1, 4>(__VA_ARGS__);                                \
        case 21: return FN<2, 1>(__VA_ARGS__);                                \
        case 22: return FN<2, 2>(__VA_ARGS__);                                \
        case 24: return FN<2, 4>(__VA_ARGS__);                                \
        case 31: return FN<3, 1>(__VA_ARGS__);                                \
        case 32: return FN<3, 2>(__VA_ARGS__);                                \
        case 34: return FN<3, 4>(__VA_ARGS__);                                \
        case 41: return FN<4, 1>(__VA_ARGS__);                                \
        case 42: return FN<4, 2>(__VA_ARGS__);                                \
        case 44: return FN<4, 4>(__VA_ARGS__);                                \
        case 51: return FN<5, 1>(__VA_ARGS__);                                \
        case 52: return FN<5, 2>(__VA_ARGS__);                                \
        case 54: return FN<5, 4>(__VA_ARGS__);                                \
        case 61: return FN<6, 1>(__VA_ARGS__);                                \
        case 62: return FN<6, 2>(__VA_ARGS__);                                \
        case 64: return FN<6, 4>(__VA_ARGS__);                                \
        case 71: return FN<7, 1>(__VA_ARGS__);                                \
        case 72: return FN<7, 2>(__VA_ARGS__);                                \
        case 74: return FN<7, 4>(__VA_ARGS__);                                \
        case 81: return FN<8, 1>(__VA_ARGS__);                                \
        case 82: return FN<8, 2>(__VA_ARGS__);                                \
        case 84: return FN<8, 4>(__VA_ARGS__);                                \
        default: return hipErrorInvalidValue;                                 \
    }

hipError_t launchBeamHnsw(const HnswLaunch& L, int wr, int maxch, int wavesPerBlock, int blocks,
                          hipStream_t stream) {
    PANN_HNSW_DISPATCH(launchT, L, wavesPerBlock, blocks, stream)
}

hipError_t beamHnswOccupancy(int wr, int maxch, int wavesPerBlock, size_t ldsBytes, int* bpc) {
    PANN_HNSW_DISPATCH(occT, wavesPerBlock, ldsBytes, bpc)
}

}  // namespace pann
