// formats.h -- on-disk formats of the search path (host only, no HIP).
//   .i8bin        [n u32][dim u32][n*dim int8]                     cuda/lectores.cpp:70-95
//   ParlayANN     [n u32][maxDeg u32][deg[n] u32][edges u32...]    cuda/lectores.cpp:97-144
//   ground truth  [nq u32][k u32][ids nq*k u32][dists nq*k f32]    cuda/lectores.cpp:146-170
//   HNSW v3       header, (level, point) per node, lists, entries   cuda/hnsw_lector.cpp:28-111
// Readers check sizes exactly as the reference does but return an error instead of exiting.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace pann {

struct Status {
    int code = 0;  // PANN_OK or PANN_ERR_*
    std::string msg;
    bool ok() const { return code == 0; }
};

struct HostVectors {
    uint32_t n = 0, dim = 0;
    std::vector<int8_t> data;  // n * dim
};

struct HostGraph {
    uint32_t n = 0, maxDeg = 0;
    std::vector<uint32_t> degrees;
    std::vector<uint64_t> offsets;  // n + 1
    std::vector<uint32_t> edges;
};

struct HostHnsw {
    uint32_t n = 0, dim = 0, m = 0, efc = 0;
    float ml = 0, alpha = 0;
    std::vector<uint32_t> level, nodePoint;
    std::vector<uint64_t> firstList, listOffset;
    std::vector<uint32_t> listSize, ids, entries;
    uint32_t maxDeg0 = 0, maxDegUpper = 0;
};

Status readI8bin(const std::string& path, HostVectors& out);
Status readParlayGraph(const std::string& path, HostGraph& out);
Status readHnsw(const std::string& path, HostHnsw& out);
// Checks a graph held in memory (degrees <= maxDeg, edges < n).
Status checkGraph(uint32_t n, uint32_t maxDeg, const uint32_t* degrees, const uint32_t* edges,
                  uint64_t* totalEdges);

}  // namespace pann
