// pann_internal.h -- structures shared by the HIP kernels and the host library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include <vector>

namespace pann {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // padding of adjacency rows / missing neighbour
constexpr int kVecBytes = 128;            // device row pitch of base and query vectors
constexpr int kMaxDim = 128;              // widest vector the kernels accept
// Rows of at most this many bytes carry their squared norm (int32) in bytes 124..127.
constexpr int kNormInRowMaxDim = 124;
constexpr int kWave = 64;
constexpr int kTagSlots = 1024;           // per-wave byte table used to detect slot collisions
constexpr int kFlagOverflow = 1;          // cuda/busqueda.cuh:14
constexpr int kFlagShort = 2;             // cuda/busqueda.cuh:15
constexpr int kFlagConfig = 4;            // query parameters outside what the launch was sized for

// Per-query parameters of the batched form (one record per query).
struct QueryParams {
    uint32_t qid;         // id used by the "a == query index" exclusion
    int32_t Ls;
    int32_t k;
    int32_t pad_;
    int64_t visit_limit;  // <= 0: unlimited (n)
    double cut;
};
static_assert(sizeof(QueryParams) == 32, "QueryParams layout is part of the C ABI");

// Byte offsets of the per-wave LDS regions of the flat search.
struct FlatLds {
    uint32_t fkey;     // frontier keys (d << 32 | id), LsMax
    uint32_t fvis;     // frontier visited flags (u8), LsMax
    uint32_t occ;      // u8 per merged position
    uint32_t kkey;     // entering candidate keys (64*maxch)
    uint32_t kvis;     // u8
    uint32_t pend;     // pending ids after the seen filter (64*maxch u32)
    uint32_t tag;      // kTagSlots u8
    uint32_t vkey;     // visited hash keys (capV u32)
    uint32_t vcnt;     // visited hash counts (capV u8)
    uint32_t filt;     // seen filter (1 << fbitsLds u32), when held in LDS
    uint32_t ckey;     // cooperative form: candidate keys of the hop (64*maxch u64)
    uint32_t cscal;    // cooperative form: the leader's published scalars (8 x i32)
    uint32_t total;    // bytes per wave (per query in the cooperative form)
};

__host__ __device__ inline uint32_t alignUp(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

__host__ __device__ inline FlatLds flatLdsLayout(int LsMax, int maxch, uint32_t capV,
                                                 int fbitsLds) {
    FlatLds L;
    uint32_t o = 0;
    const uint32_t cmax = 64u * (uint32_t)maxch;
    L.fkey = o;    o += 8u * LsMax;
    L.kkey = o;    o += 8u * cmax;
    L.pend = o;    o += 4u * cmax;
    L.vkey = o;    o += 4u * capV;
    L.filt = o;    o += (fbitsLds > 0) ? (4u << fbitsLds) : 0u;
    L.fvis = o;    o += alignUp(LsMax, 8);
    L.occ = o;     o += alignUp(LsMax, 8);
    L.kvis = o;    o += alignUp(cmax, 8);
    L.tag = o;     o += kTagSlots;
    L.vcnt = o;    o += alignUp(capV, 8);
    L.ckey = 0;
    L.cscal = 0;
    L.total = alignUp(o, 16);
    return L;
}

// The cooperative general form (2 or 4 wavefronts per query) appends the candidate-key
// buffer its wavefronts fill and the scalars its leader publishes.
__host__ __device__ inline FlatLds flatCoopLayout(FlatLds L, int maxch) {
    uint32_t o = L.total;
    L.ckey = o;  o += 8u * 64u * (uint32_t)maxch;
    L.cscal = o; o += 32u;
    L.total = alignUp(o, 16);
    return L;
}

// The register-resident form of the flat search (frontier of at most 64 entries, one
// register per lane; adjacency rows of at most 64 entries) needs only the seen filter,
// its byte tags, the expansion counts (visited hash) and a 64-entry compaction buffer,
// which shares the tag bytes (tags live only inside the filter step).
constexpr int kFastLsMax = 64;
constexpr int kFastFilterBits = 10;  // flatFilterBits(Ls) for every Ls <= 64

// Value of FlatLaunch::fast for a register-resident launch over n rows of G entries.
// PANN_WIDE_ADDRESSING=1 in the environment selects 64-bit addressing for any size (so
// that form can be tested on small indexes).
inline int fastMode(uint32_t n, uint32_t G) {
    const char* w = getenv("PANN_WIDE_ADDRESSING");
    if (w && w[0] == '1') return 1;
    const uint64_t vecBytes = (uint64_t)n * kVecBytes, adjBytes = (uint64_t)n * G * 4;
    return (vecBytes <= 0xFFFFFFFFull && adjBytes <= 0xFFFFFFFFull) ? 2 : 1;
}

__host__ __device__ inline FlatLds flatFastLdsLayout(uint32_t capV) {
    static_assert(8 * kFastLsMax <= kTagSlots, "compaction buffer must fit in the tag bytes");
    FlatLds L = {};
    uint32_t o = 0;
    L.vkey = o; o += 4u * capV;
    L.filt = o; o += 4u << kFastFilterBits;
    L.tag = o;  o += kTagSlots;
    L.kkey = L.tag;
    L.vcnt = o; o += alignUp(capV, 8);
    L.total = alignUp(o, 16);
    return L;
}

// The two-register form (64 < Ls <= 128) keeps its seen filter in HBM; LDS holds the
// expansion counts and the byte tags, which double as a 128-key compaction buffer.
constexpr int kFast2LsMax = 128;
__host__ __device__ inline FlatLds flatFast2LdsLayout(uint32_t capV) {
    static_assert(8 * kFast2LsMax <= kTagSlots, "compaction buffer must fit in the tag bytes");
    FlatLds L = {};
    uint32_t o = 0;
    L.vkey = o; o += 4u * capV;
    L.tag = o;  o += kTagSlots;
    L.kkey = L.tag;
    L.vcnt = o; o += alignUp(capV, 8);
    L.total = alignUp(o, 16);
    return L;
}

// bits of ParlayANN's seen filter: max(10, ceil(log2(Ls*Ls)) - 2)   cuda/busqueda.cu:536-538
__host__ __device__ inline int ceilLog2(uint64_t x) {
    int b = 0;
    while (((uint64_t)1 << b) < x) b++;
    return b;
}
__host__ __device__ inline int flatFilterBits(int Ls) {
    int b = ceilLog2((uint64_t)Ls * (uint64_t)Ls) - 2;
    return b < 10 ? 10 : b;
}
// visited table capacity: nextpow2(8*Ls + 128)                       cuda/busqueda.cu:540
__host__ __device__ inline uint32_t flatVisitedCap(int Ls) {
    return 1u << ceilLog2((uint64_t)8 * Ls + 128);
}
// General-form launches whose visited table has at least this many slots keep it in HBM
// (FlatLaunch::fast -1): 5 bytes a slot would otherwise hold LDS (and so occupancy) down.
constexpr uint32_t kVisitedHbmMinCap = 2048;

struct FlatLaunch {
    const int8_t* base;        // n rows, kVecBytes pitch, zero padded
    int32_t normsInRow;        // rows (and query rows) hold |p|^2 in bytes 124..127
    const uint32_t* adj;       // n x G, kEmpty padded
    const int8_t* queries;     // nq rows, kVecBytes pitch
    const uint32_t* qids;      // nullable: query index = row
    const QueryParams* qp;     // nullable: uniform parameters below
    uint32_t n, nq, G;
    int32_t Ls, k;
    int64_t limit;
    double cut;
    // sizing of this launch
    int32_t LsMax;             // frontier capacity in LDS
    int32_t fbitsLds;          // filters up to this many bits live in LDS
    int32_t gbitsMax;          // global filter region bits per wave (0: none)
    uint32_t capV;             // visited hash capacity in LDS (power of two)
    uint32_t ldsPerWave;
    FlatLds lds;
    int32_t fast;              // register-resident kernel (layout flatFastLdsLayout): 1, or
                               // 2 when base and adjacency are below 4 GiB; +2 with next-row
                               // prefetch (the launcher adds 4 for normsInRow); 13 or 14 with
                               // the two-row prefetch (the launcher adds 2); 17 or 18 over
                               // rows of 65 to 128 entries read in two halves (the launcher
                               // adds 2).  Two-register
                               // form (flatFast2LdsLayout, filter in gfilt): 9, or 10 below
                               // 4 GiB (the launcher adds 2 for normsInRow); 21 or 22 over
                               // rows read in two halves.  0: general;
                               // -1: general with the visited table in HBM (gvkey/gvcnt)
    int32_t wpq;               // general form: wavefronts per query (0 or 1: one; 2 or 4:
                               // the cooperative kernel, one query per block, scratch per
                               // block, layout flatCoopLayout)
    uint32_t* gfilt;           // per-wave global filter regions
    uint32_t* gvkey;           // fast -1: capV visited keys per wave (then the counts:
    unsigned char* gvcnt;      //          capV bytes per wave)
    uint32_t* queue;           // persistent mode: {next ticket, finished waves}; null: static
    const uint32_t* order;     // persistent mode: ticket -> query (null: identity)
    uint32_t totalWaves;
    // outputs
    uint32_t* outIds;
    int32_t* outDists;
    uint32_t outStride;
    int32_t* outComps;
    int32_t* outFlags;
    int32_t* outEdges;
    int32_t* outHops;
};

// ---- HNSW ----
// filter bits of HNSW::search at ef: ef > 2 ? 2*ceil(log2 ef) - 2 : 2   cuda/busqueda_hnsw.cu:556-557
__host__ __device__ inline int hnswFilterBits(int ef) { return ef > 2 ? 2 * ceilLog2(ef) - 2 : 2; }
// Width in wavefront registers (64 entries each) of the HNSW kernel's W and cand lists for
// beams up to Ls: Ls + 9 entries (the entry points and a margin) must fit; 0 = too wide.
inline int hnswListWidth(int Ls, uint32_t numEntries) {
    const int need = (Ls > (int)numEntries ? Ls : (int)numEntries) + 9;
    return need <= 64 ? 1 : need <= 128 ? 2 : need <= 192 ? 3 : need <= 256 ? 4 : need <= 320 ? 5
         : need <= 384 ? 6 : need <= 448 ? 7 : need <= 512 ? 8 : 0;
}

// The HNSW kernel keeps no is_inw table (cuda/busqueda_hnsw.cu:39-60): a neighbour that
// passes the workset threshold entered the workset before exactly when it is in the
// workset now, which the kernel reads off its register-resident list.
struct HnswLds {
    uint32_t pend;   // pending ids (64*maxch u32)
    uint32_t filt;   // seen filter when in LDS (max(4, 1 << fbitsLds) u32)
    uint32_t tag;    // kTagSlots u8
    uint32_t total;
};

__host__ __device__ inline HnswLds hnswLdsLayout(int maxch, int fbitsLds) {
    HnswLds L;
    uint32_t o = 0;
    L.pend = o; o += 4u * 64u * (uint32_t)maxch;
    L.filt = o; o += 4u * (fbitsLds > 2 ? (1u << fbitsLds) : 4u);
    L.tag = o;  o += kTagSlots;
    L.total = alignUp(o, 16);
    return L;
}

struct HnswLaunch {
    const int8_t* vec;         // vectors in node order, kVecBytes pitch
    int32_t normsInRow;        // rows hold |p|^2 in bytes 124..127
    const uint32_t* adj0;      // n x G0
    const uint32_t* adjUp;     // blocks x GA
    const uint32_t* upStart;   // n: first upper block of a node
    const uint32_t* nodePoint; // n: base point held by a node
    uint32_t n, G0, GA;
    int32_t top, numEntries;
    uint32_t entries[8];
    const int8_t* queries;
    uint32_t nq;
    const QueryParams* qp;     // nullable: uniform parameters below
    int32_t Ls, k;
    int64_t visit;             // <= 0: none
    int32_t LsMax, fbitsLds, gbitsMax;
    uint32_t ldsPerWave;
    HnswLds lds;
    uint32_t* gfilt;
    uint32_t* queue;
    const uint32_t* order;
    uint32_t totalWaves;
    uint32_t* outIds;
    int32_t* outDists;
    uint32_t outStride;
    int32_t* outComps;
    int32_t* outFlags;
    int32_t* outEdges;
    int32_t* outHops;
};

hipError_t launchBeamHnsw(const HnswLaunch& L, int wr, int maxch, int wavesPerBlock, int blocks,
                          hipStream_t stream);
hipError_t beamHnswOccupancy(int wr, int maxch, int wavesPerBlock, size_t ldsBytes, int* blocksPerCU);

// Launchers implemented in the .hip translation units.
hipError_t launchBeamFlat(const FlatLaunch& L, int maxch, int wavesPerBlock, int blocks,
                          hipStream_t stream);
// Blocks of the cooperative general form (wpq wavefronts, one query, ldsBytes of LDS)
// resident per CU.
hipError_t beamFlatCoopOccupancy(int maxch, int wpq, bool hbmVisited, size_t ldsBytes, int* blocksPerCU);
hipError_t beamFlatOccupancy(int maxch, int fast, int wavesPerBlock, size_t ldsBytes,
                             int* blocksPerCU);
// Squared norms of n rows into out; with embed, over bytes 0..123 only, also written
// into bytes 124..127 of each row.
hipError_t launchRowNorms(int8_t* rows, uint32_t n, int32_t* out, bool embed, hipStream_t s);
// Per-section cycle counters of the flat search (profiling builds only).
hipError_t profileCounters(uint64_t* out, int n, bool reset);
hipError_t launchCopy(void* dst, const void* src, size_t bytes, hipStream_t s);
// Random-gather microbenchmark over kVecBytes rows (see rows.hip); adds the sum of the
// count dot products into *sum.
hipError_t launchGatherBench(const int8_t* rows, uint32_t nrows, const int8_t* queries,
                             uint32_t nq, uint64_t count, uint64_t seed,
                             unsigned long long* sum, int blocksPerCU, int cus, hipStream_t s);
// dst row i = src row idx[i] (kVecBytes rows); indices >= nsrc leave the row untouched.
hipError_t launchGatherRows(int8_t* dst, const int8_t* src, const uint32_t* idx, uint32_t n,
                            uint32_t nsrc, hipStream_t s);
// HNSW graph construction (build_gpu.hip): node levels, layer-0 lists (n x M0) and upper
// lists (one M-wide block per node and layer above 0, from upStart) on the host.
struct HnswBuildOut {
    std::vector<uint32_t> level;
    std::vector<uint64_t> upStart;
    uint32_t entry = 0, M0 = 0;
    std::vector<uint32_t> deg0, adj0, degUp, adjUp;
};
hipError_t launchBuildHnsw(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t M, uint32_t efc, float ml,
                           float alpha, uint64_t seed, int cus, HnswBuildOut* out, uint32_t* status,
                           hipStream_t s);
// Vamana graph construction (build_gpu.hip) into host arrays; *status 1: a search
// expanded too many nodes, 2: a prune met too many candidates.
hipError_t launchBuildVamana(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t R, uint32_t L,
                             float alpha, bool twoPass, uint64_t seed, int cus, uint32_t* outDeg,
                             uint32_t* outAdj, uint32_t* status, hipStream_t s);
// HCNNG over the n device rows of vec: adj[i] = node i's list, before any bridging.
hipError_t launchBuildHcnng(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t numTrees,
                            uint32_t leafSize, uint32_t mstDeg, uint64_t seed,
                            std::vector<std::vector<uint32_t>>& adj, hipStream_t s);
// PyNNDescent over the n device rows of vec: adj[i] = node i's list, before the final
// bridging from node 0.  *status: 0, or 2 when a prune met more than 2048 candidates.
hipError_t launchBuildPynndescent(const int8_t* vec, bool normsInRow, uint32_t n, uint32_t K,
                                  uint32_t numTrees, uint32_t leafSize, float alpha, float delta,
                                  uint32_t maxDeg, uint64_t seed, int cus,
                                  std::vector<std::vector<uint32_t>>& adj, uint32_t* status, hipStream_t s);
hipError_t launchBruteForceKnn(const int8_t* base, const int32_t* bnorm, uint32_t n,
                               const int8_t* queries, const int32_t* qnorm, uint32_t nq,
                               uint32_t k, bool normsInRow, uint32_t* outIds, int32_t* outDists,
                               hipStream_t s);

}  // namespace pann
