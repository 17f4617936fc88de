// pann_batch.cpp -- the batched persistent form: resident wavefronts drain a device queue
// of queries that each carry their own qid, Ls, k, visit limit and cut; the launch can be
// captured into a HIP graph and replayed; an autotuner picks the launch shape per index.
// A batch is split by Ls class -- flat: <= 64, <= 128, the rest; HNSW: the widths of its
// register-resident lists, Ls <= 55, <= 119, <= 183, <= 247, <= 311, <= 375, <= 439, the rest -- and each class drains in a
// launch of its own, sized for the longest beam of that class, so short beams run at the
// occupancy and list width of their own class rather than max_Ls's.  The class launches
// run side by side (the narrower classes on the batch's own streams, forked from and
// joined back to the caller's), each with its own HBM scratch.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "pann_objects.h"

using namespace pann;

// One launch of a batch: the queries with LsLo < Ls <= LsHi, a contiguous range of the
// Ls-descending order, drained by a kernel sized for LsHi.
struct BatchPart {
    int LsLo = 0, LsHi = 0;
    int fastKind = 0;  // flat: FlatLaunch::fast (0: general); HNSW: 0
    int wr = 1;        // HNSW: list width in wavefront registers
    int wavesPerBlock = 1, blocks = 0;
    int wpq = 1;       // flat general form: wavefronts per query (blocks hold one query)
    uint32_t begin = 0, count = 0;
    size_t gfiltOff = 0, gvisOff = 0;  // this launch's HBM scratch within the batch's regions
    FlatLaunch L{};
    HnswLaunch H{};
};

struct pann_batch {
    int kind = 0;  // 0 flat, 1 hnsw
    const pann_vectors* base = nullptr;
    const pann_graph* graph = nullptr;
    const pann_hnsw* hnsw = nullptr;
    pann_batch_config cfg{};
    uint32_t maxQueries = 0, nq = 0, dim = 0;
    int maxch = 1, wr = 1, wavesPerBlock = 1, blocks = 0;
    uint32_t totalWaves = 0;
    std::vector<BatchPart> parts;  // one per Ls class, ascending LsHi
    int8_t* queries = nullptr;
    QueryParams* params = nullptr;
    uint32_t* order = nullptr;
    uint32_t* queue = nullptr;
    uint32_t* gfilt = nullptr;
    uint32_t* gvis = nullptr;  // HBM visited tables of the flat launches that use them
    pann_search_outputs out{};
    std::vector<uint32_t> hostOrder;
    hipGraph_t graphDef = nullptr;
    hipGraphExec_t graphExec = nullptr;
    uint32_t capturedNq = 0;
    std::vector<uint32_t> capturedCounts;
    std::vector<hipStream_t> side;  // parts[i] (but the widest) launch on side[i]
    hipEvent_t forkEv = nullptr;
    std::vector<hipEvent_t> joinEv;
};

namespace {

void releaseGraph(pann_batch* b) {
    if (b->graphExec) hipGraphExecDestroy(b->graphExec);
    if (b->graphDef) hipGraphDestroy(b->graphDef);
    b->graphExec = nullptr;
    b->graphDef = nullptr;
    b->capturedNq = 0;
    b->capturedCounts.clear();
}

int checkBatchConfig(const pann_batch_config* cfg) {
    if (!cfg) return setError(PANN_ERR_ARG, "config is NULL");
    if (cfg->max_Ls < 1) return setError(PANN_ERR_ARG, "max_Ls must be >= 1");
    if (cfg->max_k < 1) return setError(PANN_ERR_ARG, "max_k must be >= 1");
    if (cfg->waves_per_block < 0 || cfg->waves_per_block > 4 || cfg->waves_per_block == 3)
        return setError(PANN_ERR_ARG, "waves_per_block must be 0, 1, 2 or 4");
    if (cfg->lds_filter_bits != 0 && (cfg->lds_filter_bits < 2 || cfg->lds_filter_bits > 12))
        return setError(PANN_ERR_ARG, "lds_filter_bits must be 0 or 2..12");
    if (cfg->visited_slots &&
        ((cfg->visited_slots & (cfg->visited_slots - 1)) || cfg->visited_slots < 64))
        return setError(PANN_ERR_ARG, "visited_slots must be 0 or a power of two >= 64");
    if (cfg->waves_per_query > 4 || cfg->waves_per_query == 3)
        return setError(PANN_ERR_ARG, "waves_per_query must be 0, 1, 2 or 4");
    if (cfg->prefetch > 2) return setError(PANN_ERR_ARG, "prefetch must be 0, 1 or 2");
    return PANN_OK;
}

// common allocation of query/params/order/queue/output buffers
int allocCommon(pann_batch* b) {
    const size_t mq = std::max<uint32_t>(1, b->maxQueries);
    HIP_TRY(hipMalloc(&b->queries, mq * kVecBytes));
    HIP_TRY(hipMemset(b->queries, 0, mq * kVecBytes));
    HIP_TRY(hipMalloc(&b->params, mq * sizeof(QueryParams)));
    HIP_TRY(hipMalloc(&b->order, mq * 4));
    const size_t queues = std::max<size_t>(1, b->parts.size());  // one counter pair per launch
    HIP_TRY(hipMalloc(&b->queue, queues * 2 * 4));
    HIP_TRY(hipMemset(b->queue, 0, queues * 2 * 4));
    int r = allocOutputs(b->out, b->maxQueries, (uint32_t)b->cfg.max_k);
    if (r != PANN_OK) return r;
    return PANN_OK;
}

std::vector<uint32_t> partCounts(const pann_batch* b) {
    std::vector<uint32_t> c;
    for (const BatchPart& P : b->parts) c.push_back(P.count);
    return c;
}

int allocFilters(pann_batch* b, size_t bytes) {
    if (bytes) HIP_TRY(hipMalloc(&b->gfilt, bytes));
    return PANN_OK;
}

int allocVisited(pann_batch* b, size_t bytes) {
    if (bytes) HIP_TRY(hipMalloc(&b->gvis, bytes));
    return PANN_OK;
}

// Wavefronts per block and resident blocks for one kernel of the batch.
int sizeLaunch(pann_batch* b, int fastKind, int wr, uint32_t ldsPerWave, int* wpbOut, int* blocksOut) {
    int cus = 0;
    int r = deviceCUs(&cus);
    if (r != PANN_OK) return r;
    auto occ = [&](int w, size_t lds, int* bpc) {
        return b->kind == 0 ? beamFlatOccupancy(b->maxch, fastKind, w, lds, bpc)
                            : beamHnswOccupancy(wr, b->maxch, w, lds, bpc);
    };
    int wpb = b->cfg.waves_per_block;
    int bpc = 0;
    if (wpb == 0) {
        r = pickWavesPerBlock(ldsPerWave, occ, &wpb, &bpc);
        if (r != PANN_OK) return r;
    } else {
        if ((size_t)wpb * ldsPerWave > kMaxLdsPerBlock)
            return setError(PANN_ERR_CAPACITY, "%d wavefronts x %u bytes of LDS exceed 64 KB", wpb, ldsPerWave);
        const hipError_t e = occ(wpb, (size_t)ldsPerWave * wpb, &bpc);
        if (e != hipSuccess) return setError(PANN_ERR_HIP, "occupancy query: %s", hipGetErrorString(e));
        bpc = std::min(bpc, ldsBlocksPerCU((size_t)ldsPerWave * wpb));
    }
    *wpbOut = wpb;
    *blocksOut = std::max(1, bpc) * cus;
    return PANN_OK;
}

// Sizes one flat launch for beams up to LsHi under the batch's configuration.  The visited
// hash and the LDS filter never exceed what LsHi itself derives.  Records the launch's
// global filter bits (0: none) in *gbits.
int planFlatPart(pann_batch* b, int LsLo, int LsHi, bool top, BatchPart* P, int* gbits) {
    const pann_batch_config& c = b->cfg;
    const int fbMax = flatFilterBits(LsHi);
    const int fbLds = std::min(c.lds_filter_bits ? c.lds_filter_bits : kFilterBitsLds, fbMax);
    const uint32_t derived = flatVisitedCap(LsHi);
    const uint32_t capV = !c.visited_slots ? derived : top ? c.visited_slots : std::min(c.visited_slots, derived);
    // one frontier key per lane (filter in LDS), or two (64 < Ls <= 128, filter in HBM);
    // rows of 65 to 128 entries are read in two halves
    const bool fast = b->maxch <= 2 && LsHi <= kFastLsMax && fbLds >= kFastFilterBits;
    const bool row2 = b->maxch == 2;
    const bool fast2 = b->maxch <= 2 && !fast && LsHi <= kFast2LsMax;
    const bool hbmVisited = !fast && !fast2 && capV >= kVisitedHbmMinCap;
    // a general launch whose shortest beam already needs a larger filter keeps none in LDS
    const int fbLaunch = (!fast && !fast2 && flatFilterBits(LsLo + 1) > fbLds) ? 0 : fbLds;
    const int mode = fastMode(b->graph->n, b->graph->G);
    P->LsLo = LsLo;
    P->LsHi = LsHi;
    P->fastKind = fast && row2 ? 16 + mode
                : fast ? (c.prefetch == 2 ? 12 + mode : mode + (c.prefetch ? 2 : 0))
                       : fast2 ? (row2 ? 20 : 8) + mode : hbmVisited ? -1 : 0;
    // several wavefronts per query: the general form only
    const int wpq = (!fast && !fast2 && c.waves_per_query > 1) ? (int)c.waves_per_query : 1;
    FlatLds lay = fast ? flatFastLdsLayout(capV)
                : fast2 ? flatFast2LdsLayout(capV)
                        : flatLdsLayout(LsHi, b->maxch, hbmVisited ? 0 : capV, fbLaunch);
    if (wpq > 1) lay = flatCoopLayout(lay, b->maxch);
    if (lay.total > kMaxLdsPerBlock)
        return setError(PANN_ERR_CAPACITY, "max_Ls=%d needs %u bytes of LDS per query", LsHi, lay.total);
    int r = PANN_OK;
    if (wpq > 1) {
        int cus = 0, bpc = 0;
        r = deviceCUs(&cus);
        if (r != PANN_OK) return r;
        const hipError_t e = beamFlatCoopOccupancy(b->maxch, wpq, hbmVisited, lay.total, &bpc);
        if (e != hipSuccess) return setError(PANN_ERR_HIP, "occupancy query: %s", hipGetErrorString(e));
        bpc = std::min(bpc, ldsBlocksPerCU(lay.total));
        P->wavesPerBlock = wpq;
        P->blocks = std::max(1, bpc) * cus;
    } else {
        r = sizeLaunch(b, P->fastKind, 1, lay.total, &P->wavesPerBlock, &P->blocks);
        if (r != PANN_OK) return r;
    }
    P->wpq = wpq;
    *gbits = (fbMax > fbLds || fast2) ? fbMax : 0;
    FlatLaunch& L = P->L;
    L.n = b->graph->n;
    L.G = b->graph->G;
    L.LsMax = LsHi;
    L.fbitsLds = fast || fast2 ? fbLds : fbLaunch;
    L.gbitsMax = *gbits;
    L.capV = capV;
    L.ldsPerWave = lay.total;
    L.lds = lay;
    L.fast = P->fastKind;
    L.wpq = wpq;
    if (top) {
        b->cfg.lds_filter_bits = fbLds;
        b->cfg.visited_slots = capV;
        if (!fast || row2) b->cfg.prefetch = 0;
    }
    return PANN_OK;
}

// Sizes one HNSW launch for beams up to LsHi (as planFlatPart).  Records the launch's
// global filter bits in *gbits.  HNSW launches keep no visited table (visited_slots is
// not used and reads back 0).
int planHnswPart(pann_batch* b, int LsLo, int LsHi, bool top, BatchPart* P, int* gbits) {
    const pann_batch_config& c = b->cfg;
    const pann_hnsw* h = b->hnsw;
    const int wr = hnswListWidth(LsHi, h->numEntries);
    if (!wr) return setError(PANN_ERR_CAPACITY, "HNSW max_Ls=%d above the supported 503", LsHi);
    const int fbMax = hnswFilterBits(LsHi);
    const int fbLds = std::max(2, std::min(c.lds_filter_bits ? c.lds_filter_bits : kFilterBitsLds, fbMax));
    // a launch whose shortest beam already needs a larger layer-0 filter keeps only the
    // 4-slot upper-layer filter in LDS
    const int fbLaunch = hnswFilterBits(LsLo + 1) > fbLds ? 2 : fbLds;
    const HnswLds lay = hnswLdsLayout(b->maxch, fbLaunch);
    if (lay.total > kMaxLdsPerBlock)
        return setError(PANN_ERR_CAPACITY, "max_Ls=%d needs %u bytes of LDS per query", LsHi, lay.total);
    P->LsLo = LsLo;
    P->LsHi = LsHi;
    P->wr = wr;
    P->fastKind = 0;
    int r = sizeLaunch(b, P->fastKind, wr, lay.total, &P->wavesPerBlock, &P->blocks);
    if (r != PANN_OK) return r;
    *gbits = fbMax > fbLds ? fbMax : 0;
    HnswLaunch& H = P->H;
    H.vec = h->nodeVec;
    H.normsInRow = h->dim <= (uint32_t)kNormInRowMaxDim;
    H.adj0 = h->adj0;
    H.adjUp = h->adjUp;
    H.upStart = h->upStart;
    H.nodePoint = h->nodePoint;
    H.n = h->n;
    H.G0 = h->G0;
    H.GA = h->GA;
    H.top = (int32_t)h->top;
    H.numEntries = (int32_t)h->numEntries;
    for (int e = 0; e < 8; e++) H.entries[e] = e < (int)h->numEntries ? h->entries[e] : 0;
    H.LsMax = LsHi;
    H.fbitsLds = fbLaunch;
    H.gbitsMax = *gbits;
    H.ldsPerWave = lay.total;
    H.lds = lay;
    if (top) {
        b->cfg.lds_filter_bits = fbLds;
        b->cfg.visited_slots = 0;
    }
    return PANN_OK;
}

// Allocates the shared scratch of the parts and the common buffers, then points every
// part's launch at them.
int finishParts(pann_batch* b, size_t gfiltBytes, size_t gvisBytes) {
    int r = allocFilters(b, gfiltBytes);
    if (r == PANN_OK) r = allocVisited(b, gvisBytes);
    if (r == PANN_OK) r = allocCommon(b);
    if (r != PANN_OK) return r;
    if (b->parts.size() > 1) {
        // default-priority streams: the class launches are independent and run side by side
        b->side.assign(b->parts.size() - 1, nullptr);
        b->joinEv.assign(b->parts.size() - 1, nullptr);
        for (size_t i = 0; i + 1 < b->parts.size(); i++) {
            HIP_TRY(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&b->joinEv[i], hipEventDisableTiming));
        }
        HIP_TRY(hipEventCreateWithFlags(&b->forkEv, hipEventDisableTiming));
    }
    const BatchPart& top = b->parts.back();
    b->wavesPerBlock = top.wavesPerBlock;
    b->blocks = top.blocks;
    b->cfg.waves_per_block = top.wavesPerBlock;
    return PANN_OK;
}

}  // namespace

extern "C" {

int pann_batch_prepare_flat(const pann_vectors* base, const pann_graph* g, uint32_t max_queries,
                            const pann_batch_config* cfg, pann_batch** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!base || !g) return setError(PANN_ERR_ARG, "NULL index");
    if (base->n != g->n) return setError(PANN_ERR_ARG, "base has %u rows, graph %u nodes", base->n, g->n);
    int r = checkBatchConfig(cfg);
    if (r != PANN_OK) return r;
    pann_batch* b = new pann_batch();
    b->kind = 0;
    b->base = base;
    b->graph = g;
    b->cfg = *cfg;
    b->maxQueries = max_queries;
    b->dim = base->dim;
    b->maxch = g->G <= 64 ? 1 : g->G <= 128 ? 2 : 4;
    const int LsMax = cfg->max_Ls;
    // only the one-key form over rows of at most 64 entries prefetches
    if (LsMax > kFastLsMax || b->maxch > 1) b->cfg.prefetch = 0;
    std::vector<int> his;
    for (int hi : {kFastLsMax, kFast2LsMax})
        if (hi < LsMax) his.push_back(hi);
    his.push_back(LsMax);
    b->parts.resize(his.size());
    size_t gfiltBytes = 0, gvisBytes = 0;
    for (size_t i = 0; i < his.size(); i++) {
        int gbits = 0;
        r = planFlatPart(b, i ? his[i - 1] : 0, his[i], i + 1 == his.size(), &b->parts[i], &gbits);
        if (r != PANN_OK) {
            pann_batch_free(b);
            return r;
        }
        BatchPart& P = b->parts[i];
        const uint32_t waves = (uint32_t)P.blocks * P.wavesPerBlock;
        b->totalWaves = std::max(b->totalWaves, waves);
        P.gfiltOff = gfiltBytes;
        P.gvisOff = gvisBytes;
        if (gbits) gfiltBytes += ((size_t)waves << gbits) * 4;
        if (P.fastKind < 0) gvisBytes += (size_t)waves * P.L.capV * 5;
    }
    r = finishParts(b, gfiltBytes, gvisBytes);
    if (r != PANN_OK) {
        pann_batch_free(b);
        return r;
    }
    for (BatchPart& P : b->parts) {
        FlatLaunch& L = P.L;
        L.gfilt = L.gbitsMax ? reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(b->gfilt) + P.gfiltOff) : nullptr;
        if (P.fastKind < 0) {
            const size_t slots = (size_t)P.blocks * P.wavesPerBlock * L.capV;
            L.gvkey = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(b->gvis) + P.gvisOff);
            L.gvcnt = reinterpret_cast<unsigned char*>(L.gvkey + slots);
        }
        L.qp = b->params;
        L.qids = nullptr;
        L.queries = b->queries;
        L.base = base->dev;
        L.normsInRow = base->dim <= (uint32_t)kNormInRowMaxDim;
        L.adj = g->adj;
        L.outIds = b->out.ids;
        L.outDists = b->out.dists;
        L.outStride = b->out.stride;
        L.outComps = b->out.comps;
        L.outFlags = b->out.flags;
        L.outEdges = b->out.edges;
        L.outHops = b->out.hops;
    }
    *out = b;
    return PANN_OK;
}

int pann_batch_prepare_hnsw(const pann_hnsw* h, uint32_t max_queries, const pann_batch_config* cfg,
                            pann_batch** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!h) return setError(PANN_ERR_ARG, "NULL index");
    int r = checkBatchConfig(cfg);
    if (r != PANN_OK) return r;
    const int LsMax = cfg->max_Ls;
    if (!hnswListWidth(LsMax, h->numEntries))
        return setError(PANN_ERR_CAPACITY, "HNSW max_Ls=%d above the supported 503", LsMax);
    pann_batch* b = new pann_batch();
    b->kind = 1;
    b->hnsw = h;
    b->cfg = *cfg;
    b->cfg.prefetch = 0;  // only the flat register-resident form prefetches
    b->maxQueries = max_queries;
    b->dim = h->dim;
    const uint32_t G = std::max(h->G0, h->GA);
    b->maxch = G <= 64 ? 1 : G <= 128 ? 2 : 4;
    // classes by list width: Ls + 9 within 64, 128, ..., 448 entries
    std::vector<int> his;
    for (int hi : {55, 119, 183, 247, 311, 375, 439})
        if (hi < LsMax) his.push_back(hi);
    his.push_back(LsMax);
    b->parts.resize(his.size());
    size_t gfiltBytes = 0;
    for (size_t i = 0; i < his.size(); i++) {
        int gbits = 0;
        r = planHnswPart(b, i ? his[i - 1] : 0, his[i], i + 1 == his.size(), &b->parts[i], &gbits);
        if (r != PANN_OK) {
            pann_batch_free(b);
            return r;
        }
        BatchPart& P = b->parts[i];
        const uint32_t waves = (uint32_t)P.blocks * P.wavesPerBlock;
        b->totalWaves = std::max(b->totalWaves, waves);
        P.gfiltOff = gfiltBytes;
        if (gbits) gfiltBytes += ((size_t)waves << gbits) * 4;
    }
    r = finishParts(b, gfiltBytes, 0);
    if (r != PANN_OK) {
        pann_batch_free(b);
        return r;
    }
    b->wr = b->parts.back().wr;
    for (BatchPart& P : b->parts) {
        HnswLaunch& H = P.H;
        H.queries = b->queries;
        H.qp = b->params;
        H.gfilt = H.gbitsMax ? reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(b->gfilt) + P.gfiltOff) : nullptr;
        H.outIds = b->out.ids;
        H.outDists = b->out.dists;
        H.outStride = b->out.stride;
        H.outComps = b->out.comps;
        H.outFlags = b->out.flags;
        H.outEdges = b->out.edges;
        H.outHops = b->out.hops;
    }
    *out = b;
    return PANN_OK;
}

int pann_batch_set_queries(pann_batch* b, const int8_t* host_rows, const pann_query_params* params,
                           uint32_t nq, void* stream) {
    if (!b || (nq && (!host_rows || !params))) return setError(PANN_ERR_ARG, "NULL argument");
    if (nq > b->maxQueries) return setError(PANN_ERR_CAPACITY, "%u queries, batch holds %u", nq, b->maxQueries);
    for (uint32_t i = 0; i < nq; i++) {
        if (params[i].Ls < 1 || params[i].Ls > b->cfg.max_Ls)
            return setError(PANN_ERR_ARG, "query %u: Ls=%d outside 1..%d", i, params[i].Ls, b->cfg.max_Ls);
        if (params[i].k < 0 || params[i].k > b->cfg.max_k)
            return setError(PANN_ERR_ARG, "query %u: k=%d outside 0..%d", i, params[i].k, b->cfg.max_k);
        if (!(params[i].cut > 0)) return setError(PANN_ERR_ARG, "query %u: cut must be > 0", i);
    }
    const hipStream_t st = (hipStream_t)stream;
    b->nq = nq;
    if (nq == 0) return PANN_OK;
    // longest beams first: the persistent wavefronts then finish close together
    b->hostOrder.resize(nq);
    std::iota(b->hostOrder.begin(), b->hostOrder.end(), 0u);
    std::stable_sort(b->hostOrder.begin(), b->hostOrder.end(),
                     [&](uint32_t x, uint32_t y) { return params[x].Ls > params[y].Ls; });
    // Each launch takes the range of that order whose Ls falls in its class.  A batch that
    // fits one round of the widest class's resident wavefronts runs whole in that launch:
    // every query then starts at once, and the batch takes as long as its slowest query
    // rather than the sum of each class's slowest.
    BatchPart& top = b->parts.back();
    if (nq <= (uint32_t)top.blocks * (top.wpq > 1 ? 1u : (uint32_t)top.wavesPerBlock)) {
        for (BatchPart& P : b->parts) P.begin = P.count = 0;
        top.count = nq;
    } else {
        for (BatchPart& P : b->parts) {
            P.begin = P.count = 0;
            for (uint32_t i = 0; i < nq; i++) {
                const int Ls = params[i].Ls;
                P.begin += Ls > P.LsHi;
                P.count += Ls > P.LsLo && Ls <= P.LsHi;
            }
        }
    }
    HIP_TRY(hipMemcpy2DAsync(b->queries, kVecBytes, host_rows, b->dim, b->dim, nq, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->params, params, (size_t)nq * sizeof(QueryParams), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->order, b->hostOrder.data(), (size_t)nq * 4, hipMemcpyHostToDevice, st));
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return PANN_OK;
}

int pann_batch_launch(pann_batch* b, void* stream) {
    if (!b) return setError(PANN_ERR_ARG, "batch is NULL");
    if (b->nq == 0) return PANN_OK;
    hipError_t e = hipSuccess;
    const hipStream_t st = (hipStream_t)stream;
    int used = 0;
    for (const BatchPart& P : b->parts) used += P.count > 0;
    // more than one class: the narrower ones fork onto the batch's streams
    if (used > 1) HIP_TRY(hipEventRecord(b->forkEv, st));
    // widest class first; a small class launches only as many wavefronts as it has queries
    for (size_t i = b->parts.size(); i-- > 0 && e == hipSuccess;) {
        const BatchPart& P = b->parts[i];
        if (P.count == 0) continue;
        const bool fork = used > 1 && i + 1 < b->parts.size();
        const hipStream_t ps = fork ? b->side[i] : st;
        if (fork) HIP_TRY(hipStreamWaitEvent(ps, b->forkEv, 0));
        const uint32_t qpb = P.wpq > 1 ? 1u : (uint32_t)P.wavesPerBlock;  // queries per block
        const int blocks = (int)std::min<uint32_t>((uint32_t)P.blocks, (P.count + qpb - 1) / qpb);
        const uint32_t waves = (uint32_t)blocks * qpb;  // query slots: the queue's finishers
        if (b->kind == 0) {
            FlatLaunch L = P.L;
            L.nq = P.count;
            L.order = b->order + P.begin;
            L.queue = b->queue + 2 * i;
            L.totalWaves = waves;
            e = launchBeamFlat(L, b->maxch, P.wavesPerBlock, blocks, ps);
        } else {
            HnswLaunch H = P.H;
            H.nq = P.count;
            H.order = b->order + P.begin;
            H.queue = b->queue + 2 * i;
            H.totalWaves = waves;
            e = launchBeamHnsw(H, P.wr, b->maxch, P.wavesPerBlock, blocks, ps);
        }
        if (e == hipSuccess && fork) e = hipEventRecord(b->joinEv[i], ps);
    }
    // join: the caller's stream waits for every forked launch
    for (size_t i = 0; e == hipSuccess && used > 1 && i + 1 < b->parts.size(); i++)
        if (b->parts[i].count) e = hipStreamWaitEvent(st, b->joinEv[i], 0);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "batched launch: %s", hipGetErrorString(e));
    return PANN_OK;
}

int pann_batch_capture(pann_batch* b, void* stream) {
    if (!b) return setError(PANN_ERR_ARG, "batch is NULL");
    if (!stream) return setError(PANN_ERR_ARG, "capture needs a non-default stream");
    releaseGraph(b);
    const hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int r = pann_batch_launch(b, stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (r != PANN_OK) {
        if (g) hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "end capture: %s", hipGetErrorString(e));
    b->graphDef = g;
    HIP_TRY(hipGraphInstantiate(&b->graphExec, g, nullptr, nullptr, 0));
    b->capturedNq = b->nq;
    b->capturedCounts = partCounts(b);
    return PANN_OK;
}

int pann_batch_replay(pann_batch* b, void* stream) {
    if (!b) return setError(PANN_ERR_ARG, "batch is NULL");
    if (!b->graphExec) return setError(PANN_ERR_ARG, "no captured launch");
    if (b->capturedNq != b->nq || b->capturedCounts != partCounts(b))
        return setError(PANN_ERR_ARG, "queries changed since the capture");
    HIP_TRY(hipGraphLaunch(b->graphExec, (hipStream_t)stream));
    return PANN_OK;
}

int pann_batch_outputs(const pann_batch* b, pann_search_outputs* out) {
    if (!b || !out) return setError(PANN_ERR_ARG, "NULL argument");
    *out = b->out;
    out->nq = b->nq;
    return PANN_OK;
}

int pann_batch_download(const pann_batch* b, uint32_t* ids, int32_t* dists, int32_t* comps,
                        int32_t* flags, int32_t* edges, int32_t* hops, void* stream) {
    if (!b) return setError(PANN_ERR_ARG, "batch is NULL");
    const hipStream_t st = (hipStream_t)stream;
    const size_t nk = (size_t)b->nq * b->out.stride * 4, nn = (size_t)b->nq * 4;
    if (b->nq == 0) return PANN_OK;
    if (ids) HIP_TRY(hipMemcpyAsync(ids, b->out.ids, nk, hipMemcpyDeviceToHost, st));
    if (dists) HIP_TRY(hipMemcpyAsync(dists, b->out.dists, nk, hipMemcpyDeviceToHost, st));
    if (comps) HIP_TRY(hipMemcpyAsync(comps, b->out.comps, nn, hipMemcpyDeviceToHost, st));
    if (flags) HIP_TRY(hipMemcpyAsync(flags, b->out.flags, nn, hipMemcpyDeviceToHost, st));
    if (edges) HIP_TRY(hipMemcpyAsync(edges, b->out.edges, nn, hipMemcpyDeviceToHost, st));
    if (hops) HIP_TRY(hipMemcpyAsync(hops, b->out.hops, nn, hipMemcpyDeviceToHost, st));
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return PANN_OK;
}

int pann_batch_info(const pann_batch* b, pann_batch_config* cfg, uint32_t* resident_waves) {
    if (!b) return setError(PANN_ERR_ARG, "batch is NULL");
    if (cfg) *cfg = b->cfg;
    if (resident_waves) *resident_waves = b->totalWaves;
    return PANN_OK;
}

void pann_batch_free(pann_batch* b) {
    if (!b) return;
    releaseGraph(b);
    hipFree(b->queries);
    hipFree(b->params);
    hipFree(b->order);
    hipFree(b->queue);
    hipFree(b->gfilt);
    hipFree(b->gvis);
    for (hipStream_t x : b->side)
        if (x) hipStreamDestroy(x);
    for (hipEvent_t x : b->joinEv)
        if (x) hipEventDestroy(x);
    if (b->forkEv) hipEventDestroy(b->forkEv);
    freeOutputs(b->out);
    delete b;
}

int pann_batch_autotune(const pann_vectors* base, const pann_graph* g, const pann_hnsw* h,
                        const int8_t* host_rows, const pann_query_params* params, uint32_t nq,
                        int32_t max_Ls, int32_t max_k, pann_batch_config* best, float* best_ms) {
    if (!best || (!g && !h) || (g && !base) || !host_rows || !params || nq == 0)
        return setError(PANN_ERR_ARG, "invalid autotune arguments");
    // visited-table slots: HNSW searches keep none (one value, 0)
    const uint32_t derived = h ? 0 : flatVisitedCap(max_Ls);
    const std::vector<uint32_t> slotChoices =
        h ? std::vector<uint32_t>{0} : std::vector<uint32_t>{derived, derived / 2, derived * 2};
    const int fbMax = h ? hnswFilterBits(max_Ls) : flatFilterBits(max_Ls);
    std::vector<pann_batch_config> cands;
    for (int wpb : {1, 2, 4})
        for (int fb : {10, 11, 12}) {
            if (fb > std::max(fbMax, 10)) continue;
            for (uint32_t vs : slotChoices) {
                if (!h && (vs < 256u || vs > 16384 || vs < 2u * (uint32_t)max_Ls + 64)) continue;
                for (uint32_t pf : {0u, 1u, 2u}) {
                    if (pf && (h || max_Ls > kFastLsMax || g->G > 64)) continue;
                    // several wavefronts per query: flat graphs whose top class is the
                    // general form (beams above 128 or rows above 128)
                    const bool general = !h && (max_Ls > kFast2LsMax || g->G > 128);
                    for (uint32_t wq : {1u, 2u, 4u}) {
                        if (wq > 1 && (!general || wpb != 1)) continue;
                        cands.push_back({max_Ls, max_k, wpb, fb, vs, pf, wq});
                    }
                }
            }
        }
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float bestMs = 1e30f;
    bool found = false;
    std::vector<int32_t> flags(nq);
    for (const auto& c : cands) {
        pann_batch* b = nullptr;
        const int r = h ? pann_batch_prepare_hnsw(h, nq, &c, &b) : pann_batch_prepare_flat(base, g, nq, &c, &b);
        if (r != PANN_OK) continue;  // does not fit: skip this shape
        float ms = 1e30f;
        bool ok = pann_batch_set_queries(b, host_rows, params, nq, st) == PANN_OK &&
                  pann_batch_launch(b, st) == PANN_OK;
        for (int rep = 0; ok && rep < 3; rep++) {
            ok = hipEventRecord(e0, st) == hipSuccess && pann_batch_launch(b, st) == PANN_OK &&
                 hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            float t = 0;
            if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess) ms = std::min(ms, t);
        }
        ok = ok && pann_batch_download(b, nullptr, nullptr, nullptr, flags.data(), nullptr, nullptr, st) == PANN_OK &&
             hipStreamSynchronize(st) == hipSuccess;
        bool clean = ok;
        for (uint32_t i = 0; clean && i < nq; i++) clean = (flags[i] & (kFlagOverflow | kFlagConfig)) == 0;
        if (clean && ms < bestMs) {
            bestMs = ms;
            pann_batch_info(b, best, nullptr);
            found = true;
        }
        pann_batch_free(b);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    if (!found) return setError(PANN_ERR_CAPACITY, "no candidate configuration ran cleanly");
    if (best_ms) *best_ms = bestMs;
    return PANN_OK;
}

}  // extern "C"
