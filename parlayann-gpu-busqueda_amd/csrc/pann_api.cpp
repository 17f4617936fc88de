// pann_api.cpp -- the C ABI of include/pann.h: device-resident vectors and graphs, and
// searchers that size LDS/scratch and launch the gfx950 kernels.
#include "../../include/pann.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "formats.h"
#include "hnsw_plan.h"
#include "pann_objects.h"
#include "tools_common.h"

#include <thread>

using namespace pann;

namespace pann {

thread_local std::string g_lastError;

int setError(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_lastError = buf;
    return code;
}

int fromStatus(const Status& s) {
    if (s.ok()) return PANN_OK;
    g_lastError = s.msg;
    return s.code;
}

int deviceCUs(int* cus) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
    return PANN_OK;
}

// rows of `dim` bytes -> device rows with kVecBytes pitch (padding stays zero)
int uploadRows(int8_t* dev, const int8_t* host, uint32_t n, uint32_t dim, hipStream_t s,
               bool async) {
    if (n == 0) return PANN_OK;
    if (async) {
        HIP_TRY(hipMemcpy2DAsync(dev, kVecBytes, host, dim, dim, n, hipMemcpyHostToDevice, s));
    } else {
        HIP_TRY(hipMemcpy2D(dev, kVecBytes, host, dim, dim, n, hipMemcpyHostToDevice));
    }
    return PANN_OK;
}

void freeOutputs(pann_search_outputs& o) {
    hipFree(o.ids);
    hipFree(o.dists);
    hipFree(o.comps);
    hipFree(o.flags);
    hipFree(o.edges);
    hipFree(o.hops);
    o = pann_search_outputs{};
}

int allocOutputs(pann_search_outputs& o, uint32_t nq, uint32_t stride) {
    o.nq = nq;
    o.stride = stride;
    const size_t nk = std::max<size_t>(1, (size_t)nq * stride), nn = std::max<uint32_t>(1, nq);
    HIP_TRY(hipMalloc(&o.ids, nk * 4));
    HIP_TRY(hipMalloc(&o.dists, nk * 4));
    HIP_TRY(hipMalloc(&o.comps, nn * 4));
    HIP_TRY(hipMalloc(&o.flags, nn * 4));
    HIP_TRY(hipMalloc(&o.edges, nn * 4));
    HIP_TRY(hipMalloc(&o.hops, nn * 4));
    HIP_TRY(hipMemset(o.flags, 0, nn * 4));
    return PANN_OK;
}

int checkConfig(const pann_search_config* cfg) {
    if (!cfg) return setError(PANN_ERR_ARG, "config is NULL");
    if (cfg->Ls < 1) return setError(PANN_ERR_ARG, "Ls must be >= 1 (got %d)", cfg->Ls);
    if (cfg->k < 0) return setError(PANN_ERR_ARG, "k must be >= 0 (got %d)", cfg->k);
    if (!(cfg->cut > 0)) return setError(PANN_ERR_ARG, "cut must be > 0");
    return PANN_OK;
}

}  // namespace pann

extern "C" {

const char* pann_last_error(void) { return g_lastError.c_str(); }

int pann_version(void) { return 100; }

int pann_device_count(int* count) {
    if (!count) return setError(PANN_ERR_ARG, "count is NULL");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess || *count == 0) {
        *count = 0;
        return setError(PANN_ERR_HIP, "no HIP device: %s", hipGetErrorString(e));
    }
    return PANN_OK;
}

int pann_synchronize(void* stream) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return PANN_OK;
}

int pann_device_synchronize(void) {
    HIP_TRY(hipDeviceSynchronize());
    return PANN_OK;
}

int pann_device_alloc(size_t bytes, void** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    HIP_TRY(hipMalloc(out, std::max<size_t>(bytes, 16)));
    return PANN_OK;
}

void pann_device_free(void* p) { hipFree(p); }

int pann_host_alloc(size_t bytes, void** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    HIP_TRY(hipHostMalloc(out, std::max<size_t>(bytes, 16), hipHostMallocDefault));
    return PANN_OK;
}

void pann_host_free(void* p) { hipHostFree(p); }

int pann_memcpy(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return PANN_OK;
    if (!dst || !src) return setError(PANN_ERR_ARG, "NULL pointer");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream));
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return PANN_OK;
}

int pann_stream_create(void** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    HIP_TRY(hipStreamCreateWithFlags((hipStream_t*)out, hipStreamNonBlocking));
    return PANN_OK;
}

void pann_stream_destroy(void* s) {
    if (s) hipStreamDestroy((hipStream_t)s);
}

int pann_event_create(void** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    HIP_TRY(hipEventCreate((hipEvent_t*)out));
    return PANN_OK;
}

int pann_event_record(void* ev, void* stream) {
    HIP_TRY(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return PANN_OK;
}

int pann_event_elapsed_ms(void* start, void* end, float* ms) {
    if (!ms) return setError(PANN_ERR_ARG, "ms is NULL");
    HIP_TRY(hipEventSynchronize((hipEvent_t)end));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return PANN_OK;
}

void pann_event_destroy(void* ev) {
    if (ev) hipEventDestroy((hipEvent_t)ev);
}

int pann_profile_counters(uint64_t* out, int n, int reset) {
    if (!out || n < 0) return setError(PANN_ERR_ARG, "NULL argument");
    const hipError_t e = profileCounters(out, n, reset != 0);
    if (e == hipErrorNotSupported) return setError(PANN_ERR_CAPACITY, "not a profiling build");
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "profile counters: %s", hipGetErrorString(e));
    return PANN_OK;
}

int pann_copy_kernel(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src || bytes % 16) return setError(PANN_ERR_ARG, "copy needs 16-byte multiples");
    HIP_TRY(launchCopy(dst, src, bytes, (hipStream_t)stream));
    return PANN_OK;
}

int pann_gather_bench(const pann_vectors* rows, const pann_vectors* queries, uint64_t count,
                      uint64_t seed, int reps, float* ms, uint64_t* checksum) {
    if (!rows || !queries || !ms || !checksum || count == 0 || reps < 1 || rows->n == 0 ||
        queries->n == 0)
        return setError(PANN_ERR_ARG, "invalid gather bench arguments");
    uint32_t nq = 1;
    while (nq * 2 <= queries->n && nq < (1u << 30)) nq *= 2;
    int cus = 0;
    int r = deviceCUs(&cus);
    if (r != PANN_OK) return r;
    unsigned long long* sum = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&sum, sizeof(*sum));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // one launch to warm up, then `reps` timed launches; the checksum is the last one's
    for (int i = 0; e == hipSuccess && i <= reps; i++) {
        if (i == 1) e = hipEventRecord(e0, nullptr);
        if (e == hipSuccess) e = hipMemsetAsync(sum, 0, sizeof(*sum), nullptr);
        if (e == hipSuccess)
            e = launchGatherBench(rows->dev, rows->n, queries->dev, nq, count, seed, sum, 8, cus, nullptr);
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float total = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&total, e0, e1);
    unsigned long long host = 0;
    if (e == hipSuccess) e = hipMemcpy(&host, sum, sizeof(host), hipMemcpyDeviceToHost);
    hipFree(sum);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "gather bench: %s", hipGetErrorString(e));
    *ms = total / reps;
    *checksum = host;
    return PANN_OK;
}

// ---------------------------------------------------------------- vectors
int pann_vectors_create(const int8_t* host_rows, uint32_t n, uint32_t dim, pann_vectors** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (dim == 0 || dim > (uint32_t)kMaxDim)
        return setError(PANN_ERR_CAPACITY, "dim %u not supported (1..%d)", dim, kMaxDim);
    if (n > 0 && !host_rows) return setError(PANN_ERR_ARG, "host_rows is NULL");
    pann_vectors* v = new pann_vectors();
    v->n = n;
    v->dim = dim;
    const size_t bytes = std::max<size_t>(kVecBytes, (size_t)n * kVecBytes);
    hipError_t e = hipMalloc(&v->dev, bytes);
    if (e == hipSuccess) e = hipMemset(v->dev, 0, bytes);
    if (e == hipSuccess) e = hipMalloc(&v->norms, 4 * std::max<size_t>(1, n));
    if (e == hipSuccess) e = hipMemset(v->norms, 0, 4 * std::max<size_t>(1, n));
    if (e != hipSuccess) {
        pann_vectors_free(v);
        return setError(PANN_ERR_HIP, "allocating %zu bytes of vectors: %s", bytes, hipGetErrorString(e));
    }
    int r = uploadRows(v->dev, host_rows, n, dim, nullptr, false);
    if (r == PANN_OK) {
        e = launchRowNorms(v->dev, n, v->norms, dim <= (uint32_t)kNormInRowMaxDim, nullptr);
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) r = setError(PANN_ERR_HIP, "row norms: %s", hipGetErrorString(e));
    }
    if (r != PANN_OK) {
        pann_vectors_free(v);
        return r;
    }
    *out = v;
    return PANN_OK;
}

int pann_vectors_load(const char* path, pann_vectors** out) {
    if (!path || !out) return setError(PANN_ERR_ARG, "NULL argument");
    HostVectors hv;
    int r = fromStatus(readI8bin(path, hv));
    if (r != PANN_OK) return r;
    return pann_vectors_create(hv.data.data(), hv.n, hv.dim, out);
}

int pann_vectors_copy_async(pann_vectors* v, const int8_t* host_rows, uint32_t n, void* stream) {
    if (!v || (!host_rows && n)) return setError(PANN_ERR_ARG, "NULL argument");
    if (n > v->n) return setError(PANN_ERR_ARG, "%u rows do not fit in %u", n, v->n);
    const int r = uploadRows(v->dev, host_rows, n, v->dim, (hipStream_t)stream, true);
    if (r != PANN_OK) return r;
    HIP_TRY(launchRowNorms(v->dev, n, v->norms, v->dim <= (uint32_t)kNormInRowMaxDim, (hipStream_t)stream));
    return PANN_OK;
}

int pann_vectors_info(const pann_vectors* v, uint32_t* n, uint32_t* dim, const int8_t** rows,
                      uint32_t* pitch) {
    if (!v) return setError(PANN_ERR_ARG, "vectors is NULL");
    if (n) *n = v->n;
    if (dim) *dim = v->dim;
    if (rows) *rows = v->dev;
    if (pitch) *pitch = kVecBytes;
    return PANN_OK;
}

void pann_vectors_free(pann_vectors* v) {
    if (!v) return;
    hipFree(v->dev);
    hipFree(v->norms);
    delete v;
}

// ---------------------------------------------------------------- flat graph
int pann_graph_create(uint32_t n, uint32_t max_degree, const uint32_t* degrees,
                      const uint32_t* edges, pann_graph** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n == 0 || !degrees) return setError(PANN_ERR_ARG, "empty graph");
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) total += degrees[i];
    if (total > 0 && !edges) return setError(PANN_ERR_ARG, "edges is NULL");
    int r = fromStatus(checkGraph(n, max_degree, degrees, edges, &total));
    if (r != PANN_OK) return r;
    if (max_degree == 0) max_degree = 1;
    if (max_degree > 256)
        return setError(PANN_ERR_CAPACITY, "max degree %u above the supported 256", max_degree);
    std::vector<uint32_t> dense((size_t)n * max_degree, kEmpty);
    uint64_t off = 0;
    for (uint32_t i = 0; i < n; i++) {
        memcpy(dense.data() + (size_t)i * max_degree, edges + off, 4ull * degrees[i]);
        off += degrees[i];
    }
    pann_graph* g = new pann_graph();
    g->n = n;
    g->G = max_degree;
    hipError_t e = hipMalloc(&g->adj, dense.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(g->adj, dense.data(), dense.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(g->adj);
        delete g;
        return setError(PANN_ERR_HIP, "uploading graph: %s", hipGetErrorString(e));
    }
    *out = g;
    return PANN_OK;
}

int pann_graph_load(const char* path, pann_graph** out) {
    if (!path || !out) return setError(PANN_ERR_ARG, "NULL argument");
    HostGraph hg;
    int r = fromStatus(readParlayGraph(path, hg));
    if (r != PANN_OK) return r;
    return pann_graph_create(hg.n, hg.maxDeg, hg.degrees.data(), hg.edges.data(), out);
}

int pann_graph_info(const pann_graph* g, uint32_t* n, uint32_t* max_degree) {
    if (!g) return setError(PANN_ERR_ARG, "graph is NULL");
    if (n) *n = g->n;
    if (max_degree) *max_degree = g->G;
    return PANN_OK;
}

void pann_graph_free(pann_graph* g) {
    if (!g) return;
    hipFree(g->adj);
    delete g;
}

// ---------------------------------------------------------------- HNSW
int pann_hnsw_create(uint32_t n, const uint32_t* level, const uint32_t* node_point,
                     const uint64_t* node_first_list, const uint64_t* list_offsets,
                     const uint32_t* list_sizes, const uint32_t* list_ids,
                     const uint32_t* entries, uint32_t num_entries, const pann_vectors* base,
                     pann_hnsw** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!n || !level || !node_point || !node_first_list || !list_offsets || !list_sizes ||
        !entries || !base)
        return setError(PANN_ERR_ARG, "NULL argument");
    if (num_entries < 1 || num_entries > 8)
        return setError(PANN_ERR_FORMAT, "%u entry points (1..8 supported)", num_entries);
    uint32_t G0 = 1, GA = 1;
    uint64_t blocks = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (node_point[i] >= base->n)
            return setError(PANN_ERR_FORMAT, "node %u holds point %u beyond the base (%u)", i,
                            node_point[i], base->n);
        for (uint32_t l = 0; l <= level[i]; l++) {
            const uint32_t sz = list_sizes[node_first_list[i] + l];
            if (l == 0) G0 = std::max(G0, sz);
            else GA = std::max(GA, sz);
            const uint32_t* ids = list_ids + list_offsets[node_first_list[i] + l];
            for (uint32_t j = 0; j < sz; j++) {
                if (ids[j] >= n) return setError(PANN_ERR_FORMAT, "neighbour %u out of range", ids[j]);
                // a layer-l list may only name nodes present on layer l
                if (level[ids[j]] < l)
                    return setError(PANN_ERR_FORMAT, "node %u lists node %u on layer %u above its level %u",
                                    i, ids[j], l, level[ids[j]]);
            }
        }
        blocks += level[i];
    }
    for (uint32_t e = 0; e < num_entries; e++) {
        if (entries[e] >= n) return setError(PANN_ERR_FORMAT, "entry point out of range");
        // every entry point starts the top-layer search, so it must live on the top layer
        if (level[entries[e]] != level[entries[0]])
            return setError(PANN_ERR_FORMAT, "entry point %u is on level %u, not the top level %u",
                            entries[e], level[entries[e]], level[entries[0]]);
    }
    if (G0 > 256 || GA > 256)
        return setError(PANN_ERR_CAPACITY, "degree %u/%u above the supported 256", G0, GA);

    std::vector<uint32_t> adj0((size_t)n * G0, kEmpty);
    std::vector<uint32_t> up(std::max<uint64_t>(1, blocks) * GA, kEmpty);
    std::vector<uint32_t> start(n, kEmpty);
    uint64_t b = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t f = node_first_list[i];
        memcpy(adj0.data() + (size_t)i * G0, list_ids + list_offsets[f], 4ull * list_sizes[f]);
        if (level[i] >= 1) {
            start[i] = (uint32_t)b;
            for (uint32_t l = 1; l <= level[i]; l++)
                memcpy(up.data() + (b + l - 1) * GA, list_ids + list_offsets[f + l],
                       4ull * list_sizes[f + l]);
            b += level[i];
        }
    }
    pann_hnsw* h = new pann_hnsw();
    h->n = n;
    h->G0 = G0;
    h->GA = GA;
    h->numEntries = num_entries;
    for (uint32_t e = 0; e < num_entries; e++) h->entries[e] = entries[e];
    h->top = level[entries[0]];
    h->dim = base->dim;
    bool identity = base->n == n;
    for (uint32_t i = 0; i < n && identity; i++) identity = node_point[i] == i;
    hipError_t e = hipMalloc(&h->adj0, adj0.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(h->adj0, adj0.data(), adj0.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&h->adjUp, up.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(h->adjUp, up.data(), up.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&h->upStart, 4ull * n);
    if (e == hipSuccess) e = hipMemcpy(h->upStart, start.data(), 4ull * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&h->nodePoint, 4ull * n);
    if (e == hipSuccess) e = hipMemcpy(h->nodePoint, node_point, 4ull * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (identity) {
            h->nodeVec = base->dev;
            h->ownVec = false;
        } else {
            // vectors in node order: one gather launch over the node -> point map
            e = hipMalloc(&h->nodeVec, (size_t)n * kVecBytes);
            h->ownVec = e == hipSuccess;
            if (e == hipSuccess)
                e = launchGatherRows(h->nodeVec, base->dev, h->nodePoint, n, base->n, nullptr);
            if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        }
    }
    if (e != hipSuccess) {
        pann_hnsw_free(h);
        return setError(PANN_ERR_HIP, "uploading HNSW: %s", hipGetErrorString(e));
    }
    *out = h;
    return PANN_OK;
}

int pann_hnsw_load(const char* path, const pann_vectors* base, pann_hnsw** out) {
    if (!path || !out) return setError(PANN_ERR_ARG, "NULL argument");
    HostHnsw h;
    int r = fromStatus(readHnsw(path, h));
    if (r != PANN_OK) return r;
    if (!base) return setError(PANN_ERR_ARG, "base vectors are NULL");
    return pann_hnsw_create(h.n, h.level.data(), h.nodePoint.data(), h.firstList.data(),
                            h.listOffset.data(), h.listSize.data(), h.ids.data(),
                            h.entries.data(), (uint32_t)h.entries.size(), base, out);
}

int pann_hnsw_info(const pann_hnsw* h, uint32_t* n, uint32_t* g0, uint32_t* ga, uint32_t* top,
                   uint32_t* ne) {
    if (!h) return setError(PANN_ERR_ARG, "hnsw is NULL");
    if (n) *n = h->n;
    if (g0) *g0 = h->G0;
    if (ga) *ga = h->GA;
    if (top) *top = h->top;
    if (ne) *ne = h->numEntries;
    return PANN_OK;
}

void pann_hnsw_free(pann_hnsw* h) {
    if (!h) return;
    hipFree(h->adj0);
    hipFree(h->adjUp);
    hipFree(h->upStart);
    hipFree(h->nodePoint);
    if (h->ownVec) hipFree(h->nodeVec);
    delete h;
}

}  // extern "C"

namespace pann {

// The launch of one flat search configuration over nq queries on graph g (the sizing of
// BuscadorGPU::preparar, cuda/busqueda.cu:525-568): kernel form, LDS layout, wavefronts per
// block, resident blocks, and the HBM scratch each wavefront needs (a seen filter above
// kFilterBitsLds bits, a visited table from kVisitedHbmMinCap slots).  Scratch pointers
// and outputs are left for the caller.
int flatShape(const pann_graph* g, uint32_t nq, const pann_search_config* cfg, FlatShape* out) {
    if (!g) return setError(PANN_ERR_ARG, "graph is NULL");
    int r = checkConfig(cfg);
    if (r != PANN_OK) return r;
    FlatShape sh;
    sh.maxch = g->G <= 64 ? 1 : g->G <= 128 ? 2 : 4;
    const int Ls = cfg->Ls;
    sh.fbits = flatFilterBits(Ls);
    const uint32_t capV = flatVisitedCap(Ls);
    // one frontier key per lane, or two (64 < Ls <= 128, filter in HBM); rows of 65 to 128
    // entries are read in two halves
    const bool fast = sh.maxch <= 2 && Ls <= kFastLsMax;
    const bool fast2 = sh.maxch <= 2 && !fast && Ls <= kFast2LsMax;
    sh.ldsFilter = sh.fbits <= kFilterBitsLds && !fast2;
    sh.hbmVisited = !fast && !fast2 && capV >= kVisitedHbmMinCap;
    const int fastKind = fast ? (sh.maxch == 2 ? 16 : 0) + fastMode(g->n, g->G)
                       : fast2 ? (sh.maxch == 2 ? 20 : 8) + fastMode(g->n, g->G) : sh.hbmVisited ? -1 : 0;
    const FlatLds lay = fast ? flatFastLdsLayout(capV)
                      : fast2 ? flatFast2LdsLayout(capV)
                              : flatLdsLayout(Ls, sh.maxch, sh.hbmVisited ? 0 : capV, sh.ldsFilter ? sh.fbits : 0);
    if (lay.total > kMaxLdsPerBlock)
        return setError(PANN_ERR_CAPACITY, "Ls=%d needs %u bytes of LDS per query", Ls, lay.total);
    int bpc = 0, cus = 0;
    r = pickWavesPerBlock(lay.total, [&](int wpb, size_t lds, int* b) {
            return beamFlatOccupancy(sh.maxch, fastKind, wpb, lds, b); }, &sh.wavesPerBlock, &bpc);
    if (r == PANN_OK) r = deviceCUs(&cus);
    if (r != PANN_OK) return r;
    const uint32_t want = (nq + sh.wavesPerBlock - 1) / sh.wavesPerBlock;
    sh.blocks = (int)std::max<uint32_t>(1, std::min<uint32_t>(want, (uint32_t)std::max(1, bpc) * cus));
    sh.scratchPerWave = (sh.ldsFilter ? 0 : ((size_t)4 << sh.fbits)) + (sh.hbmVisited ? (size_t)capV * 5 : 0);

    FlatLaunch& L = sh.L;
    L = FlatLaunch{};
    L.n = g->n;
    L.G = g->G;
    L.nq = nq;
    L.Ls = Ls;
    L.k = cfg->k;
    L.limit = cfg->visit_limit > 0 ? cfg->visit_limit : (int64_t)g->n;
    L.cut = cfg->cut;
    L.LsMax = Ls;
    L.fbitsLds = sh.ldsFilter ? sh.fbits : 0;
    L.gbitsMax = sh.ldsFilter ? 0 : sh.fbits;
    L.capV = capV;
    L.ldsPerWave = lay.total;
    L.lds = lay;
    L.fast = fastKind;
    L.totalWaves = (uint32_t)sh.blocks * sh.wavesPerBlock;
    L.queue = nullptr;
    L.outStride = (uint32_t)std::max(1, cfg->k);
    *out = sh;
    return PANN_OK;
}

// Points the launch at `totalWaves` wavefronts' scratch starting at `scratch`: the seen
// filters, then the visited keys, then their counts.
void flatBindScratch(FlatShape& sh, unsigned char* scratch) {
    FlatLaunch& L = sh.L;
    const size_t waves = L.totalWaves;
    size_t off = 0;
    L.gfilt = nullptr;
    L.gvkey = nullptr;
    L.gvcnt = nullptr;
    if (!sh.ldsFilter) {
        L.gfilt = reinterpret_cast<uint32_t*>(scratch);
        off += (waves << sh.fbits) * 4;
    }
    if (sh.hbmVisited) {
        L.gvkey = reinterpret_cast<uint32_t*>(scratch + off);
        L.gvcnt = scratch + off + waves * L.capV * 4;
    }
}

// Long beams (Ls or ef of 65 and more) take their queries from a work queue (two
// counters that the launch's last wavefront resets) instead of in turn: their search
// lengths vary by hundreds of hops, and a wavefront that finishes early takes the next
// query at once.  Short beams stay strided, which measured faster there.
constexpr int kQueueMinLs = 65;

int allocQueue(pann_searcher* s) {
    hipError_t e = hipMalloc(&s->queue, 2 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(s->queue, 0, 2 * sizeof(uint32_t));
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "allocating the work queue: %s", hipGetErrorString(e));
    return PANN_OK;
}

}  // namespace pann

extern "C" {

// ---------------------------------------------------------------- flat searcher
int pann_searcher_prepare(const pann_graph* g, uint32_t nq, const pann_search_config* cfg,
                          pann_searcher** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    FlatShape sh;
    int r = flatShape(g, nq, cfg, &sh);
    if (r != PANN_OK) return r;
    pann_searcher* s = new pann_searcher();
    s->kind = 0;
    s->cfg = *cfg;
    s->nq = nq;
    s->maxch = sh.maxch;
    s->wavesPerBlock = sh.wavesPerBlock;
    s->blocks = sh.blocks;
    const size_t scratch = sh.scratchPerWave * sh.L.totalWaves;
    if (scratch) {
        const hipError_t he = hipMalloc(&s->gvis, scratch);
        if (he != hipSuccess) {
            delete s;
            return setError(PANN_ERR_HIP, "allocating %zu bytes of search scratch: %s", scratch,
                            hipGetErrorString(he));
        }
    }
    flatBindScratch(sh, reinterpret_cast<unsigned char*>(s->gvis));
    s->L = sh.L;
    FlatLaunch& L = s->L;
    r = cfg->Ls >= kQueueMinLs ? allocQueue(s) : PANN_OK;
    L.queue = s->queue;
    if (r == PANN_OK) r = allocOutputs(s->out, nq, (uint32_t)std::max(1, cfg->k));
    if (r != PANN_OK) {
        pann_searcher_free(s);
        return r;
    }
    L.outIds = s->out.ids;
    L.outDists = s->out.dists;
    L.outStride = s->out.stride;
    L.outComps = s->out.comps;
    L.outFlags = s->out.flags;
    L.outEdges = s->out.edges;
    L.outHops = s->out.hops;
    *out = s;
    return PANN_OK;
}

int pann_searcher_launch(pann_searcher* s, const pann_vectors* base, const pann_vectors* queries,
                         const pann_graph* g, const uint32_t* qids, void* stream) {
    if (!s || !base || !queries || !g) return setError(PANN_ERR_ARG, "NULL argument");
    if (s->kind != 0) return setError(PANN_ERR_ARG, "searcher was prepared for HNSW");
    if (base->n != g->n) return setError(PANN_ERR_ARG, "base has %u rows, graph %u nodes", base->n, g->n);
    // the prepared launch derives its addressing form and default visit limit from n and G
    if (g->G != s->L.G || g->n != s->L.n)
        return setError(PANN_ERR_ARG, "graph (%u nodes, degree %u) differs from the prepared one "
                        "(%u, %u)", g->n, g->G, s->L.n, s->L.G);
    if (base->dim != queries->dim)
        return setError(PANN_ERR_ARG, "dim mismatch base %u queries %u", base->dim, queries->dim);
    if (queries->n < s->nq) return setError(PANN_ERR_ARG, "%u queries, searcher prepared for %u", queries->n, s->nq);
    if (s->nq == 0) return PANN_OK;
    FlatLaunch L = s->L;
    L.base = base->dev;
    L.normsInRow = base->dim <= (uint32_t)kNormInRowMaxDim;
    L.adj = g->adj;
    L.queries = queries->dev;
    L.qids = qids;
    L.qp = nullptr;
    hipError_t e = launchBeamFlat(L, s->maxch, s->wavesPerBlock, s->blocks, (hipStream_t)stream);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "beam search launch: %s", hipGetErrorString(e));
    return PANN_OK;
}

int pann_hnsw_searcher_prepare(const pann_hnsw* h, uint32_t nq, const pann_search_config* cfg,
                               pann_searcher** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!h) return setError(PANN_ERR_ARG, "hnsw is NULL");
    int r = checkConfig(cfg);
    if (r != PANN_OK) return r;
    const int Ls = cfg->Ls;
    const int wr = hnswListWidth(Ls, h->numEntries);
    if (!wr) return setError(PANN_ERR_CAPACITY, "HNSW Ls=%d above the supported 503", Ls);
    const uint32_t G = std::max(h->G0, h->GA);
    const int maxch = G <= 64 ? 1 : G <= 128 ? 2 : 4;
    const int bits0 = hnswFilterBits(Ls);
    const bool ldsFilter = bits0 <= kFilterBitsLds;
    const HnswLds lay = hnswLdsLayout(maxch, ldsFilter ? bits0 : 0);
    if (lay.total > kMaxLdsPerBlock)
        return setError(PANN_ERR_CAPACITY, "Ls=%d needs %u bytes of LDS per query", Ls, lay.total);

    pann_searcher* s = new pann_searcher();
    s->kind = 1;
    s->cfg = *cfg;
    s->nq = nq;
    s->maxch = maxch;
    s->wr = wr;
    int bpc = 0, cus = 0;
    r = pickWavesPerBlock(lay.total, [&](int wpb, size_t lds, int* b) {
            return beamHnswOccupancy(wr, maxch, wpb, lds, b); }, &s->wavesPerBlock, &bpc);
    if (r == PANN_OK) r = deviceCUs(&cus);
    if (r != PANN_OK) {
        delete s;
        return r;
    }
    hipError_t he = hipSuccess;
    const uint32_t want = (nq + s->wavesPerBlock - 1) / s->wavesPerBlock;
    s->blocks = (int)std::max<uint32_t>(1, std::min<uint32_t>(want, (uint32_t)std::max(1, bpc) * cus));
    const uint32_t totalWaves = (uint32_t)s->blocks * s->wavesPerBlock;
    HnswLaunch& H = s->H;
    H.n = h->n;
    H.G0 = h->G0;
    H.GA = h->GA;
    H.top = (int32_t)h->top;
    H.numEntries = (int32_t)h->numEntries;
    for (int e = 0; e < 8; e++) H.entries[e] = e < (int)h->numEntries ? h->entries[e] : 0;
    H.nq = nq;
    H.qp = nullptr;
    H.Ls = Ls;
    H.k = cfg->k;
    H.visit = cfg->visit_limit;
    H.LsMax = Ls;
    H.fbitsLds = ldsFilter ? bits0 : 2;
    H.gbitsMax = ldsFilter ? 0 : bits0;
    H.ldsPerWave = lay.total;
    H.lds = lay;
    H.queue = nullptr;
    H.totalWaves = totalWaves;
    if (!ldsFilter) {
        he = hipMalloc(&s->gfilt, ((size_t)totalWaves << bits0) * 4);
        if (he != hipSuccess) {
            delete s;
            return setError(PANN_ERR_HIP, "allocating filter scratch: %s", hipGetErrorString(he));
        }
    }
    H.gfilt = s->gfilt;
    r = Ls >= kQueueMinLs ? allocQueue(s) : PANN_OK;
    H.queue = s->queue;
    if (r == PANN_OK) r = allocOutputs(s->out, nq, (uint32_t)std::max(1, cfg->k));
    if (r != PANN_OK) {
        pann_searcher_free(s);
        return r;
    }
    H.outIds = s->out.ids;
    H.outDists = s->out.dists;
    H.outStride = s->out.stride;
    H.outComps = s->out.comps;
    H.outFlags = s->out.flags;
    H.outEdges = s->out.edges;
    H.outHops = s->out.hops;
    *out = s;
    return PANN_OK;
}

int pann_hnsw_searcher_launch(pann_searcher* s, const pann_vectors* queries, const pann_hnsw* h,
                              void* stream) {
    if (!s || !queries || !h) return setError(PANN_ERR_ARG, "NULL argument");
    if (s->kind != 1) return setError(PANN_ERR_ARG, "searcher was prepared for a flat graph");
    if (h->n != s->H.n || h->G0 != s->H.G0 || h->GA != s->H.GA)
        return setError(PANN_ERR_ARG, "index differs from the prepared one");
    if (queries->dim != h->dim)
        return setError(PANN_ERR_ARG, "dim mismatch index %u queries %u", h->dim, queries->dim);
    if (queries->n < s->nq) return setError(PANN_ERR_ARG, "%u queries, searcher prepared for %u", queries->n, s->nq);
    // list width and LDS were sized for the prepared index's entry count
    if (hnswListWidth(s->cfg.Ls, h->numEntries) != s->wr)
        return setError(PANN_ERR_ARG, "index has %u entry points; the searcher was sized for "
                        "lists of %d registers", h->numEntries, s->wr);
    if (s->nq == 0) return PANN_OK;
    HnswLaunch H = s->H;
    // the top level and entry points come from the index of this launch
    // (BuscadorHnsw::lanzar, cuda/busqueda_hnsw.cu:588-596)
    H.top = (int32_t)h->top;
    H.numEntries = (int32_t)h->numEntries;
    for (int e = 0; e < 8; e++) H.entries[e] = e < (int)h->numEntries ? h->entries[e] : 0;
    H.vec = h->nodeVec;
    H.normsInRow = h->dim <= (uint32_t)kNormInRowMaxDim;
    H.adj0 = h->adj0;
    H.adjUp = h->adjUp;
    H.upStart = h->upStart;
    H.nodePoint = h->nodePoint;
    H.queries = queries->dev;
    hipError_t e = launchBeamHnsw(H, s->wr, s->maxch, s->wavesPerBlock, s->blocks, (hipStream_t)stream);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "HNSW search launch: %s", hipGetErrorString(e));
    return PANN_OK;
}

int pann_searcher_outputs(const pann_searcher* s, pann_search_outputs* out) {
    if (!s || !out) return setError(PANN_ERR_ARG, "NULL argument");
    *out = s->out;
    return PANN_OK;
}

int pann_searcher_download(const pann_searcher* s, uint32_t* ids, int32_t* dists, int32_t* comps,
                           int32_t* flags, int32_t* edges, int32_t* hops, void* stream) {
    if (!s) return setError(PANN_ERR_ARG, "searcher is NULL");
    const hipStream_t st = (hipStream_t)stream;
    const size_t nk = (size_t)s->out.nq * s->out.stride * 4, nn = (size_t)s->out.nq * 4;
    if (s->out.nq == 0) return PANN_OK;
    if (ids) HIP_TRY(hipMemcpyAsync(ids, s->out.ids, nk, hipMemcpyDeviceToHost, st));
    if (dists) HIP_TRY(hipMemcpyAsync(dists, s->out.dists, nk, hipMemcpyDeviceToHost, st));
    if (comps) HIP_TRY(hipMemcpyAsync(comps, s->out.comps, nn, hipMemcpyDeviceToHost, st));
    if (flags) HIP_TRY(hipMemcpyAsync(flags, s->out.flags, nn, hipMemcpyDeviceToHost, st));
    if (edges) HIP_TRY(hipMemcpyAsync(edges, s->out.edges, nn, hipMemcpyDeviceToHost, st));
    if (hops) HIP_TRY(hipMemcpyAsync(hops, s->out.hops, nn, hipMemcpyDeviceToHost, st));
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return PANN_OK;
}

int pann_bruteforce_knn(const pann_vectors* base, const pann_vectors* queries, uint32_t nq,
                        uint32_t k, uint32_t* ids_dev, int32_t* dists_dev, void* stream) {
    if (!base || !queries || (nq && (!ids_dev || !dists_dev)))
        return setError(PANN_ERR_ARG, "NULL argument");
    if (base->dim != queries->dim) return setError(PANN_ERR_ARG, "dim mismatch");
    if (nq > queries->n) return setError(PANN_ERR_ARG, "nq %u > %u query rows", nq, queries->n);
    if (k == 0 || k > 128) return setError(PANN_ERR_CAPACITY, "k=%u outside 1..128", k);
    if (nq == 0) return PANN_OK;
    const hipStream_t st = (hipStream_t)stream;
    hipError_t e = launchBruteForceKnn(base->dev, base->norms, base->n, queries->dev, queries->norms,
                                       nq, k, base->dim <= (uint32_t)kNormInRowMaxDim, ids_dev,
                                       dists_dev, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "brute-force k-NN: %s", hipGetErrorString(e));
    return PANN_OK;
}

// Joins what node 0 cannot reach (as the host builders do, tools::bridgeFromStart: the
// base rows are copied back only when a bridge is needed) and writes the graph file.
static int bridgeAndWrite(const pann_vectors* base, std::vector<std::vector<uint32_t>>& adj,
                          const char* out_path) {
    if (!tools::allReachableFromStart(adj)) {
        const uint32_t n = base->n, dim = base->dim;
        std::vector<int8_t> rows((size_t)n * kVecBytes), compact((size_t)n * dim);
        const hipError_t e = hipMemcpy(rows.data(), base->dev, rows.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return setError(PANN_ERR_HIP, "copy of the base rows: %s", hipGetErrorString(e));
        for (uint32_t i = 0; i < n; i++)
            std::copy_n(rows.data() + (size_t)i * kVecBytes, dim, compact.data() + (size_t)i * dim);
        tools::bridgeFromStart(compact.data(), n, dim, adj, (int)std::max(1u, std::thread::hardware_concurrency()));
    }
    const int w = tools::writeGraphListsFile(out_path, adj);
    if (w == 1) return setError(PANN_ERR_IO, "cannot write %s", out_path);
    if (w != 0) return setError(PANN_ERR_IO, "short write to %s", out_path);
    return PANN_OK;
}

int pann_build_vamana_gpu(const pann_vectors* base, uint32_t R, uint32_t L, float alpha,
                          int two_pass, uint64_t seed, uint32_t* out_degrees,
                          uint32_t* out_edges, void* stream) {
    if (!base || !out_degrees || !out_edges) return setError(PANN_ERR_ARG, "NULL argument");
    if (base->n == 0) return setError(PANN_ERR_ARG, "no points");
    if (R == 0 || R > 128) return setError(PANN_ERR_CAPACITY, "R=%u outside 1..128", R);
    if (L == 0 || L > 256) return setError(PANN_ERR_CAPACITY, "L=%u outside 1..256", L);
    if (alpha != alpha) return setError(PANN_ERR_ARG, "alpha is NaN");
    int cus = 0;
    int r = deviceCUs(&cus);
    if (r != PANN_OK) return r;
    uint32_t status = 0;
    const hipError_t e = launchBuildVamana(base->dev, base->dim <= (uint32_t)kNormInRowMaxDim, base->n,
                                           R, L, alpha, two_pass != 0, seed, cus, out_degrees,
                                           out_edges, &status, (hipStream_t)stream);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "Vamana build: %s", hipGetErrorString(e));
    if (status & 1u) return setError(PANN_ERR_CAPACITY, "a build search expanded more than 1024 nodes");
    if (status & 2u) return setError(PANN_ERR_CAPACITY, "a robust prune met more than 2048 candidates");
    return PANN_OK;
}

int pann_build_hcnng_gpu(const pann_vectors* base, uint32_t num_trees, uint32_t leaf_size,
                         uint32_t mst_deg, uint64_t seed, const char* out_path, void* stream) {
    if (!base || !out_path) return setError(PANN_ERR_ARG, "NULL argument");
    if (base->n == 0) return setError(PANN_ERR_ARG, "no points");
    if (num_trees == 0) return setError(PANN_ERR_ARG, "num_trees must be positive");
    if (leaf_size < 2 || leaf_size > 1024)
        return setError(PANN_ERR_CAPACITY, "leaf_size=%u outside 2..1024", leaf_size);
    if (mst_deg == 0 || mst_deg > 4) return setError(PANN_ERR_CAPACITY, "mst_deg=%u outside 1..4", mst_deg);
    std::vector<std::vector<uint32_t>> adj;
    const hipError_t e = launchBuildHcnng(base->dev, base->dim <= (uint32_t)kNormInRowMaxDim, base->n,
                                          num_trees, leaf_size, mst_deg, seed, adj, (hipStream_t)stream);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "HCNNG build: %s", hipGetErrorString(e));
    const int r = bridgeAndWrite(base, adj, out_path);
    return r;
}

int pann_build_pynndescent_gpu(const pann_vectors* base, uint32_t K, uint32_t num_trees,
                               uint32_t leaf_size, float alpha, float delta, uint32_t max_degree,
                               uint64_t seed, const char* out_path, void* stream) {
    if (!base || !out_path) return setError(PANN_ERR_ARG, "NULL argument");
    if (base->n < 2) return setError(PANN_ERR_ARG, "fewer than 2 points");
    if (K == 0 || K > 64) return setError(PANN_ERR_CAPACITY, "K=%u outside 1..64", K);
    if (num_trees == 0) return setError(PANN_ERR_ARG, "num_trees must be positive");
    if (leaf_size < 2 || leaf_size > 1024)
        return setError(PANN_ERR_CAPACITY, "leaf_size=%u outside 2..1024", leaf_size);
    if (max_degree == 0 || max_degree > 128)
        return setError(PANN_ERR_CAPACITY, "max_degree=%u outside 1..128", max_degree);
    if (alpha != alpha || delta != delta) return setError(PANN_ERR_ARG, "alpha or delta is NaN");
    int cus = 0;
    int r = deviceCUs(&cus);
    if (r != PANN_OK) return r;
    std::vector<std::vector<uint32_t>> adj;
    uint32_t status = 0;
    const hipError_t e = launchBuildPynndescent(base->dev, base->dim <= (uint32_t)kNormInRowMaxDim, base->n, K,
                                                num_trees, leaf_size, alpha, delta, max_degree, seed, cus, adj,
                                                &status, (hipStream_t)stream);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "PyNNDescent build: %s", hipGetErrorString(e));
    if (status) return setError(PANN_ERR_CAPACITY, "a robust prune met more than 2048 candidates");
    return bridgeAndWrite(base, adj, out_path);
}

int pann_build_hnsw_gpu(const pann_vectors* base, uint32_t M, uint32_t efc, float ml, float alpha,
                        uint64_t seed, const char* out_path, void* stream) {
    if (!base || !out_path) return setError(PANN_ERR_ARG, "NULL argument");
    if (base->n == 0) return setError(PANN_ERR_ARG, "no points");
    if (M == 0 || M > 64) return setError(PANN_ERR_CAPACITY, "M=%u outside 1..64", M);
    if (efc == 0 || efc > 256) return setError(PANN_ERR_CAPACITY, "efc=%u outside 1..256", efc);
    if (!(ml > 0)) return setError(PANN_ERR_ARG, "ml must be positive");
    const float al = alpha > 0 ? alpha : 1.0f;
    int cus = 0;
    int r = deviceCUs(&cus);
    if (r != PANN_OK) return r;
    HnswBuildOut h;
    uint32_t status = 0;
    const hipError_t e = launchBuildHnsw(base->dev, base->dim <= (uint32_t)kNormInRowMaxDim, base->n, M, efc, ml,
                                         al, seed, cus, &h, &status, (hipStream_t)stream);
    if (e == hipErrorInvalidValue) return setError(PANN_ERR_CAPACITY, "too many upper-layer lists");
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "HNSW build: %s", hipGetErrorString(e));
    if (status & 1u) return setError(PANN_ERR_CAPACITY, "a build search expanded more than 1024 nodes");
    if (status & 2u) return setError(PANN_ERR_CAPACITY, "a neighbour selection met more than 2048 candidates");
    const int w = tools::writeHnswV3(
        out_path, base->n, base->dim, ml, M, efc, al, h.level,
        [&](uint32_t i, uint32_t l) { return l == 0 ? h.deg0[i] : h.degUp[h.upStart[i] + l - 1]; },
        [&](uint32_t i, uint32_t l) {
            return l == 0 ? h.adj0.data() + (size_t)i * h.M0 : h.adjUp.data() + (h.upStart[i] + l - 1) * M;
        },
        h.entry);
    if (w == 1) return setError(PANN_ERR_IO, "cannot write %s", out_path);
    if (w != 0) return setError(PANN_ERR_IO, "short write to %s", out_path);
    return PANN_OK;
}

void pann_searcher_free(pann_searcher* s) {
    if (!s) return;
    freeOutputs(s->out);
    hipFree(s->gfilt);
    hipFree(s->gvis);
    hipFree(s->queue);
    delete s;
}

}  // extern "C"
