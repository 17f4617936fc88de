// pann_objects.h -- the opaque handles of include/pann.h and helpers shared by the host
// translation units of libpann.so.
#pragma once

#include <algorithm>
#include <string>

#include "../../include/pann.h"
#include "pann_internal.h"

using pann::FlatLaunch;
using pann::HnswLaunch;

struct pann_vectors {
    uint32_t n = 0, dim = 0;
    int8_t* dev = nullptr;     // n x kVecBytes
    int32_t* norms = nullptr;  // n squared norms (p . p), kept in step with the rows
};

struct pann_graph {
    uint32_t n = 0, G = 0;
    uint32_t* adj = nullptr;  // n x G
};

struct pann_hnsw {
    uint32_t n = 0;
    uint32_t G0 = 0, GA = 0;
    uint32_t top = 0, numEntries = 0;
    uint32_t entries[8] = {};
    uint32_t* adj0 = nullptr;       // n x G0
    uint32_t* adjUp = nullptr;      // blocks x GA
    uint32_t* upStart = nullptr;    // n
    uint32_t* nodePoint = nullptr;  // n
    int8_t* nodeVec = nullptr;      // n x kVecBytes, vectors in node order
    bool ownVec = false;
    uint32_t dim = 0;
};

struct pann_searcher {
    int kind = 0;  // 0 flat, 1 hnsw
    pann_search_config cfg{};
    uint32_t nq = 0;
    int maxch = 1;
    int wr = 1;
    int wavesPerBlock = 4;
    int blocks = 0;
    FlatLaunch L{};
    HnswLaunch H{};
    uint32_t* gfilt = nullptr;
    uint32_t* gvis = nullptr;  // HBM visited tables (FlatLaunch::fast -1)
    uint32_t* queue = nullptr;  // work-queue counters of long-beam launches (else strided)
    pann_search_outputs out{};
};

namespace pann {

int setError(int code, const char* fmt, ...);
int deviceCUs(int* cus);

// Launch shape of one flat search configuration (see flatShape in pann_api.cpp).
struct FlatShape {
    FlatLaunch L{};             // filled but for scratch pointers, inputs and outputs
    int maxch = 1;              // 64-entry chunks of an adjacency row
    int wavesPerBlock = 1;
    int blocks = 1;             // resident blocks (L.totalWaves = blocks * wavesPerBlock)
    int fbits = 10;             // seen filter bits
    bool ldsFilter = true;      // the filter lives in LDS (else per-wavefront HBM scratch)
    bool hbmVisited = false;    // the visited table lives in per-wavefront HBM scratch
    size_t scratchPerWave = 0;  // HBM scratch bytes per wavefront
};
int flatShape(const pann_graph* g, uint32_t nq, const pann_search_config* cfg, FlatShape* out);
void flatBindScratch(FlatShape& sh, unsigned char* scratch);
int allocOutputs(pann_search_outputs& o, uint32_t nq, uint32_t stride);
void freeOutputs(pann_search_outputs& o);

#define HIP_TRY(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::pann::setError(PANN_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Largest seen filter kept in LDS by default (2^10 u32 = 4 KB, every Ls <= 64); larger ones
// use a per-wavefront HBM region.  An 8-16 KB LDS filter costs more in occupancy than its
// L2 round trips: Ls=100 on the 1M Vamana graph runs in 8.2 ms with the filter in HBM and
// 10.9 ms with it in LDS.
constexpr int kFilterBitsLds = 10;
constexpr size_t kMaxLdsPerBlock = 64 * 1024;

// Blocks of `bytes` of LDS that are resident on one CU at once.  A gfx950 CU holds 160 KB
// of LDS, handed to workgroups in 2 KB granules: three 54,272-B blocks do not fit (a launch
// of three per CU takes two rounds) while three 53,248-B and four 40,960-B blocks do.  The
// runtime's occupancy query counts bytes, so its answer is capped by this.
inline int ldsBlocksPerCU(size_t bytes) {
    if (bytes == 0) return 1 << 30;
    const size_t granted = (bytes + 2047) & ~(size_t)2047;
    return (int)((size_t)160 * 1024 / granted);
}

// Wavefronts per block (4, 2 or 1, within 64 KB of LDS) that leaves the most wavefronts
// resident on a CU, ties to the larger block; *blocksPerCU is the resident block count.
// occ(wpb, ldsBytes, &blocksPerCU) is the kernel's occupancy query.
template <class Occ>
int pickWavesPerBlock(uint32_t ldsPerWave, Occ occ, int* wpbOut, int* blocksPerCU) {
    int best = -1;
    for (int wpb : {4, 2, 1}) {
        if ((size_t)wpb * ldsPerWave > kMaxLdsPerBlock) continue;
        int bpc = 0;
        const hipError_t e = occ(wpb, (size_t)ldsPerWave * wpb, &bpc);
        if (e != hipSuccess) return setError(PANN_ERR_HIP, "occupancy query: %s", hipGetErrorString(e));
        bpc = std::min(bpc, ldsBlocksPerCU((size_t)ldsPerWave * wpb));
        if (bpc * wpb > best) {
            best = bpc * wpb;
            *wpbOut = wpb;
            *blocksPerCU = bpc;
        }
    }
    if (best < 0) return setError(PANN_ERR_CAPACITY, "%u bytes of LDS per wavefront exceed 64 KB", ldsPerWave);
    return PANN_OK;
}

}  // namespace pann
