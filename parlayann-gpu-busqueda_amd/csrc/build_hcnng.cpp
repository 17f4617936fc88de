// build_hcnng.cpp -- HCNNG graph builder (host, multi-threaded), writes a ParlayANN graph.
//
// Algorithm as published (Munoz et al., "Hierarchical Clustering-Based Graphs for Large
// Scale Approximate Nearest Neighbor Search", and ParlayANN algorithms/HCNNG): num_trees
// random hierarchical clusterings (two random pivots per split, leaves of at most
// leaf_size points); in every leaf a minimum spanning forest with node degree at most
// mst_deg (Kruskal with a degree bound); a node's out-list is the union over trees of its
// forest edges, duplicates included (the reference measures 27% duplicate edges in
// ParlayANN's HCNNG graphs, paper/paper.tex:172-174).  Kruskal runs over each point's 16
// nearest leaf mates rather than the complete leaf graph.
#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/pann_tools.h"
#include "build_trees.h"
#include "tools_common.h"

namespace pann {
namespace tools {
namespace {

struct Dsu {
    std::vector<uint32_t> p;
    explicit Dsu(uint32_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
    uint32_t find(uint32_t x) {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
    }
};

// degree-bounded minimum spanning forest of one leaf, appended to per-node lists
void leafForest(const int8_t* base, uint32_t dim, const uint32_t* ids, uint32_t m, uint32_t deg,
                std::vector<std::vector<uint32_t>>& out) {
    if (m < 2) return;
    const uint32_t K = std::min<uint32_t>(16, m - 1);
    struct E {
        int32_t d;
        uint32_t a, b;
        bool operator<(const E& o) const {
            return d != o.d ? d < o.d : a != o.a ? a < o.a : b < o.b;
        }
    };
    std::vector<E> edges;
    edges.reserve((size_t)m * K);
    std::vector<int32_t> D((size_t)m * m);
    for (uint32_t i = 0; i < m; i++) {
        const int8_t* x = base + (size_t)ids[i] * dim;
        D[(size_t)i * m + i] = INT32_MAX;
        for (uint32_t j = i + 1; j < m; j++)
            D[(size_t)i * m + j] = D[(size_t)j * m + i] = distI8(x, base + (size_t)ids[j] * dim, dim);
    }
    std::vector<Nbr> row(m);
    for (uint32_t i = 0; i < m; i++) {
        for (uint32_t j = 0; j < m; j++) row[j] = {D[(size_t)i * m + j], j};
        std::partial_sort(row.begin(), row.begin() + K, row.end());
        for (uint32_t t = 0; t < K; t++) {
            const uint32_t j = row[t].id;
            edges.push_back({row[t].d, std::min(i, j), std::max(i, j)});
        }
    }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end(),
                            [](const E& x, const E& y) { return x.a == y.a && x.b == y.b; }),
                edges.end());
    Dsu dsu(m);
    std::vector<uint32_t> degree(m, 0);
    for (const E& e : edges) {
        if (degree[e.a] >= deg || degree[e.b] >= deg) continue;
        const uint32_t ra = dsu.find(e.a), rb = dsu.find(e.b);
        if (ra == rb) continue;
        dsu.p[ra] = rb;
        degree[e.a]++;
        degree[e.b]++;
        out[ids[e.a]].push_back(ids[e.b]);
        out[ids[e.b]].push_back(ids[e.a]);
    }
}

}  // namespace
}  // namespace tools
}  // namespace pann

extern "C" int pann_build_hcnng(const int8_t* base, uint32_t n, uint32_t dim, uint32_t num_trees,
                                uint32_t leaf_size, uint32_t mst_deg, uint64_t seed, int nthreads,
                                const char* out_path) {
    using namespace pann::tools;
    if (!base || !out_path || n == 0 || dim == 0 || num_trees == 0 || leaf_size < 2 || mst_deg == 0)
        return setError(1, "invalid argument");
    nthreads = std::max(1, nthreads);
    std::vector<std::vector<uint32_t>> adj(n);
    for (uint32_t t = 0; t < num_trees; t++) {
        std::vector<uint32_t> ids(n);
        std::iota(ids.begin(), ids.end(), 0u);
        const auto leaves = clusterTreeLeaves(base, dim, ids, leaf_size, mixSeed(seed, 100 + t), nthreads);
        std::vector<std::vector<uint32_t>> treeAdj(n);
        parallelFor(leaves.size(), nthreads, [&](size_t li, int) {
            leafForest(base, dim, ids.data() + leaves[li].first, leaves[li].second - leaves[li].first,
                       mst_deg, treeAdj);
        }, 1);
        for (uint32_t i = 0; i < n; i++) adj[i].insert(adj[i].end(), treeAdj[i].begin(), treeAdj[i].end());
    }
    bridgeFromStart(base, n, dim, adj, nthreads);
    return writeGraphLists(out_path, adj);
}
