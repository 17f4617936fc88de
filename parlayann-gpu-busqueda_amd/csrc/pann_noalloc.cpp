// pann_noalloc.cpp -- the flat search of pann_searcher_prepare + pann_searcher_launch in
// one call that allocates no device memory, for callers that cannot call hipMalloc (code
// under HIP graph capture, tight serving loops, harnesses that forbid allocation).
//
// The per-wavefront HBM scratch of long beams (seen filters above 2^10 slots, visited
// tables from 2,048 slots) comes from `scratch`, a static device array of kScratchBytes
// that every GPU the library runs on holds until the process ends.  A launch whose
// resident wavefronts would need more uses fewer wavefronts (each drains more queries);
// one whose single block needs more is refused.  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU (other streams, other threads) race, and must be
// serialised by the caller.  The kernels clear each query's tables before use, so the
// contents a previous call left do not matter.
#include <mutex>

#include "pann_objects.h"

using namespace pann;

namespace {

constexpr size_t kScratchBytes = (size_t)PANN_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
unsigned char* g_scratchPtr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch`, looked up once per device
// with that device current.  On the host the array's name is only a stand-in for the
// symbol, never an address to hand to a kernel or a copy.
int scratchOfDevice(unsigned char** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return setError(PANN_ERR_CAPACITY, "device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return setError(PANN_ERR_HIP, "address of the static search scratch on device %d: %s", dev,
                            hipGetErrorString(e));
        g_scratchPtr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratchPtr[dev];
    return PANN_OK;
}

}  // namespace

extern "C" int pann_search_flat_noalloc(const pann_vectors* base, const pann_graph* g,
                                        const pann_vectors* queries, const uint32_t* qids,
                                        uint32_t nq, const pann_search_config* cfg,
                                        uint32_t* out_ids, int32_t* out_dists, int32_t* out_comps,
                                        int32_t* out_flags, int32_t* out_edges, int32_t* out_hops,
                                        void* stream) {
    if (!base || !g || !queries || !cfg) return setError(PANN_ERR_ARG, "NULL argument");
    if (nq && (!out_ids || !out_dists)) return setError(PANN_ERR_ARG, "output ids/dists are NULL");
    if (base->n != g->n) return setError(PANN_ERR_ARG, "base has %u rows, graph %u nodes", base->n, g->n);
    if (base->dim != queries->dim)
        return setError(PANN_ERR_ARG, "dim mismatch base %u queries %u", base->dim, queries->dim);
    if (queries->n < nq) return setError(PANN_ERR_ARG, "%u queries asked, %u rows given", nq, queries->n);
    FlatShape sh;
    int r = flatShape(g, nq, cfg, &sh);
    if (r != PANN_OK || nq == 0) return r;
    if (sh.scratchPerWave) {
        const size_t maxWaves = kScratchBytes / sh.scratchPerWave;
        if (maxWaves < (size_t)sh.wavesPerBlock)
            return setError(PANN_ERR_CAPACITY,
                            "Ls=%d needs %zu bytes of scratch per block; the allocation-free "
                            "search holds %zu", cfg->Ls, sh.scratchPerWave * sh.wavesPerBlock,
                            kScratchBytes);
        const size_t maxBlocks = maxWaves / (size_t)sh.wavesPerBlock;
        if ((size_t)sh.blocks > maxBlocks) {
            sh.blocks = (int)maxBlocks;
            sh.L.totalWaves = (uint32_t)sh.blocks * sh.wavesPerBlock;
        }
        unsigned char* p = nullptr;
        r = scratchOfDevice(&p);
        if (r != PANN_OK) return r;
        flatBindScratch(sh, p);
    }
    FlatLaunch& L = sh.L;
    L.base = base->dev;
    L.normsInRow = base->dim <= (uint32_t)kNormInRowMaxDim;
    L.adj = g->adj;
    L.queries = queries->dev;
    L.qids = qids;
    L.qp = nullptr;
    L.outIds = out_ids;
    L.outDists = out_dists;
    L.outComps = out_comps;
    L.outFlags = out_flags;
    L.outEdges = out_edges;
    L.outHops = out_hops;
    const hipError_t e = launchBeamFlat(L, sh.maxch, sh.wavesPerBlock, sh.blocks, (hipStream_t)stream);
    if (e != hipSuccess) return setError(PANN_ERR_HIP, "beam search launch: %s", hipGetErrorString(e));
    return PANN_OK;
}
