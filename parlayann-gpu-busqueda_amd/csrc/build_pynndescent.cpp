// build_pynndescent.cpp -- PyNNDescent graph builder (host, multi-threaded), writes a
// ParlayANN graph.
//
// Algorithm as published (Dong et al., NN-descent; McInnes' PyNNDescent; ParlayANN
// algorithms/pyNNDescent): a K-NN graph initialised from num_trees random cluster trees
// (exact K nearest inside each leaf of at most leaf_size points), refined by NN-descent
// local joins (new x new and new x old neighbour pairs, reverse lists included) until an
// iteration changes fewer than delta * n * K entries, then every node's K-NN list plus its
// reverse edges is robust-pruned with alpha down to max_degree (as Vamana prunes).  The
// pruned graph's components are then joined inside the cluster-tree leaves, each point
// proposing its nearest leaf mate of another component (step 4), which keeps the graph
// searchable from node 0 on clustered data.
//
// Each K-NN list keeps the K best (distance, id) pairs of everything proposed to it, so
// the lists after an iteration do not depend on the order in which threads propose; the
// change count is measured on the final lists for the same reason.
#include <algorithm>
#include <atomic>
#include <numeric>
#include <vector>

#include "../../include/pann_tools.h"
#include "build_trees.h"
#include "tools_common.h"

namespace pann {
namespace tools {
namespace {

struct Knn {
    uint32_t K;
    std::vector<Nbr> nb;          // n x K, sorted, unused slots {INT32_MAX, ~0}
    std::vector<uint8_t> isNew;   // n x K
    std::vector<std::atomic_flag> lock;

    Knn(uint32_t n, uint32_t k) : K(k), nb((size_t)n * k, Nbr{INT32_MAX, 0xFFFFFFFFu}),
                                  isNew((size_t)n * k, 0), lock(n) {
        for (auto& l : lock) l.clear();
    }
    // insert (d, id) into list i if better than its worst and absent
    void propose(uint32_t i, int32_t d, uint32_t id) {
        Nbr* L = nb.data() + (size_t)i * K;
        if (!(Nbr{d, id} < L[K - 1])) return;
        while (lock[i].test_and_set(std::memory_order_acquire)) {
        }
        if (Nbr{d, id} < L[K - 1]) {
            bool present = false;
            for (uint32_t j = 0; j < K && !present; j++) present = L[j].id == id;
            if (!present) {
                uint32_t p = K - 1;
                while (p > 0 && Nbr{d, id} < L[p - 1]) {
                    L[p] = L[p - 1];
                    isNew[(size_t)i * K + p] = isNew[(size_t)i * K + p - 1];
                    p--;
                }
                L[p] = {d, id};
                isNew[(size_t)i * K + p] = 1;
            }
        }
        lock[i].clear(std::memory_order_release);
    }
};

void robustPrune(const int8_t* base, uint32_t dim, uint32_t p, std::vector<Nbr>& cand, float alpha,
                 uint32_t R, std::vector<uint32_t>& out) {
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end(),
                           [](const Nbr& a, const Nbr& b) { return a.id == b.id; }),
               cand.end());
    out.clear();
    std::vector<char> pruned(cand.size(), 0);
    for (size_t i = 0; i < cand.size() && out.size() < R; i++) {
        if (pruned[i] || cand[i].id == p) continue;
        out.push_back(cand[i].id);
        const int8_t* s = base + (size_t)cand[i].id * dim;
        for (size_t j = i + 1; j < cand.size(); j++) {
            if (pruned[j] || cand[j].id == p) continue;
            if (alpha * (float)distI8(s, base + (size_t)cand[j].id * dim, dim) <= (float)cand[j].d)
                pruned[j] = 1;
        }
    }
}

}  // namespace
}  // namespace tools
}  // namespace pann

constexpr uint32_t kBridgeCap = 24;  // bridges per node at most (step 4)

extern "C" int pann_build_pynndescent(const int8_t* base, uint32_t n, uint32_t dim, uint32_t K,
                                      uint32_t num_trees, uint32_t leaf_size, float alpha,
                                      float delta, uint32_t max_degree, uint64_t seed, int nthreads,
                                      const char* out_path) {
    using namespace pann::tools;
    if (!base || !out_path || n < 2 || dim == 0 || K == 0 || num_trees == 0 || leaf_size < 2 ||
        leaf_size > 1024 || max_degree == 0)
        return setError(1, "invalid argument");
    nthreads = std::max(1, nthreads);
    K = std::min(K, n - 1);
    Knn knn(n, K);
    auto vec = [&](uint32_t i) { return base + (size_t)i * dim; };

    // 1. initial lists from random cluster trees
    for (uint32_t t = 0; t < num_trees; t++) {
        std::vector<uint32_t> ids(n);
        std::iota(ids.begin(), ids.end(), 0u);
        const auto leaves = clusterTreeLeaves(base, dim, ids, leaf_size, mixSeed(seed, 200 + t), nthreads);
        parallelFor(leaves.size(), nthreads, [&](size_t li, int) {
            const uint32_t b = leaves[li].first, m = leaves[li].second - b;
            for (uint32_t i = 0; i < m; i++)
                for (uint32_t j = i + 1; j < m; j++) {
                    const uint32_t x = ids[b + i], y = ids[b + j];
                    const int32_t d = distI8(vec(x), vec(y), dim);
                    knn.propose(x, d, y);
                    knn.propose(y, d, x);
                }
        }, 1);
    }

    // 2. NN-descent
    const uint32_t S = std::max<uint32_t>(1, K / 2);  // sampled new / reverse lists
    for (int it = 0; it < 12; it++) {
        std::vector<std::vector<uint32_t>> nw(n), old(n), rnw(n), rold(n);
        for (uint32_t i = 0; i < n; i++) {
            for (uint32_t j = 0; j < K; j++) {
                const Nbr& x = knn.nb[(size_t)i * K + j];
                if (x.id == 0xFFFFFFFFu) continue;
                if (knn.isNew[(size_t)i * K + j] && nw[i].size() < S) {
                    nw[i].push_back(x.id);
                    knn.isNew[(size_t)i * K + j] = 0;
                } else if (!knn.isNew[(size_t)i * K + j]) {
                    old[i].push_back(x.id);
                }
            }
        }
        for (uint32_t i = 0; i < n; i++) {
            for (uint32_t x : nw[i]) if (rnw[x].size() < S) rnw[x].push_back(i);
            for (uint32_t x : old[i]) if (rold[x].size() < S) rold[x].push_back(i);
        }
        std::vector<Nbr> before(knn.nb);
        parallelFor(n, nthreads, [&](size_t u, int) {
            std::vector<uint32_t> A(nw[u]);
            A.insert(A.end(), rnw[u].begin(), rnw[u].end());
            std::vector<uint32_t> B(old[u]);
            B.insert(B.end(), rold[u].begin(), rold[u].end());
            for (size_t a = 0; a < A.size(); a++) {
                for (size_t b = a + 1; b < A.size(); b++) {
                    if (A[a] == A[b]) continue;
                    const int32_t d = distI8(vec(A[a]), vec(A[b]), dim);
                    knn.propose(A[a], d, A[b]);
                    knn.propose(A[b], d, A[a]);
                }
                for (uint32_t y : B) {
                    if (y == A[a]) continue;
                    const int32_t d = distI8(vec(A[a]), vec(y), dim);
                    knn.propose(A[a], d, y);
                    knn.propose(y, d, A[a]);
                }
            }
        }, 64);
        size_t changed = 0;
        for (size_t e = 0; e < knn.nb.size(); e++) changed += !(knn.nb[e] == before[e]);
        if ((double)changed < (double)delta * n * K) break;
    }

    // 3. reverse edges + robust prune to max_degree
    std::vector<std::vector<uint32_t>> rev(n);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < K; j++) {
            const uint32_t x = knn.nb[(size_t)i * K + j].id;
            if (x != 0xFFFFFFFFu) rev[x].push_back(i);
        }
    std::vector<std::vector<uint32_t>> adj(n);
    parallelFor(n, nthreads, [&](size_t i, int) {
        std::vector<Nbr> cand;
        for (uint32_t j = 0; j < K; j++) {
            const Nbr& x = knn.nb[i * K + j];
            if (x.id != 0xFFFFFFFFu) cand.push_back(x);
        }
        for (uint32_t r : rev[i]) cand.push_back({distI8(vec((uint32_t)i), vec(r), dim), r});
        robustPrune(base, dim, (uint32_t)i, cand, alpha, max_degree, adj[i]);
    }, 64);
    // 4. leaf bridges.  A K-NN graph of well-separated clusters falls apart into one
    // component per cluster, and a search that starts at node 0 never leaves its own.
    // The cluster trees of step 1 put pieces of neighbouring clusters into one leaf; in
    // every such leaf each point proposes its nearest leaf mate of another component, and
    // the proposals are taken by Kruskal over the leaf's components in (distance, leaf
    // position) order -- the way HCNNG's per-leaf spanning trees join a leaf, over each
    // point's nearest leaf mates -- each accepted pair added in both directions.  Each
    // leaf's Kruskal is its own, so a cluster keeps the bridges of every tree.  A node
    // takes part in at most kBridgeCap bridges: only nodes with room left when the tree
    // starts propose or are proposed, and a pair whose endpoint filled up in this leaf is
    // skipped, so rows stay within max_degree + kBridgeCap entries.
    std::vector<uint32_t> comp(n), bridges(n, 0);
    {
        std::vector<uint32_t> p(n);
        std::iota(p.begin(), p.end(), 0u);
        auto find = [&](uint32_t x) {
            while (p[x] != x) x = p[x] = p[p[x]];
            return x;
        };
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t a : adj[i]) {
                const uint32_t ra = find(i), rb = find(a);
                if (ra != rb) p[std::max(ra, rb)] = std::min(ra, rb);
            }
        for (uint32_t i = 0; i < n; i++) comp[i] = find(i);
    }
    for (uint32_t t = 0; t < num_trees; t++) {
        std::vector<uint32_t> ids(n);
        std::iota(ids.begin(), ids.end(), 0u);
        const auto leaves = clusterTreeLeaves(base, dim, ids, leaf_size, mixSeed(seed, 200 + t), nthreads);
        std::vector<std::vector<std::pair<uint32_t, uint32_t>>> found(leaves.size());
        parallelFor(leaves.size(), nthreads, [&](size_t li, int) {
            const uint32_t b = leaves[li].first, m = leaves[li].second - b;
            const uint32_t* L = ids.data() + b;
            // rep[i]: first leaf position of i's component; edges (d, lower, upper position)
            std::vector<uint32_t> rep(m);
            std::vector<uint64_t> edges;
            for (uint32_t i = 0; i < m; i++) {
                rep[i] = i;
                uint64_t best = ~0ull;
                const bool room = bridges[L[i]] < kBridgeCap;
                for (uint32_t j = 0; j < m; j++) {
                    if (comp[L[j]] == comp[L[i]]) {
                        rep[i] = std::min(rep[i], j);
                        continue;
                    }
                    if (!room || bridges[L[j]] >= kBridgeCap) continue;
                    const uint64_t key = ((uint64_t)(uint32_t)distI8(vec(L[i]), vec(L[j]), dim) << 32) | j;
                    best = std::min(best, key);
                }
                if (best != ~0ull) {
                    const uint32_t j = (uint32_t)best;
                    edges.push_back(((best >> 32) << 20) | ((uint64_t)std::min(i, j) << 10) | std::max(i, j));
                }
            }
            std::sort(edges.begin(), edges.end());
            edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
            std::vector<uint32_t> p(rep), used(m, 0);
            auto find = [&](uint32_t x) {
                while (p[x] != x) x = p[x] = p[p[x]];
                return x;
            };
            for (const uint64_t e : edges) {
                const uint32_t x = (uint32_t)(e >> 10) & 1023u, y = (uint32_t)e & 1023u;
                const uint32_t rx = find(x), ry = find(y);
                if (rx == ry) continue;
                if (bridges[L[x]] + used[x] >= kBridgeCap || bridges[L[y]] + used[y] >= kBridgeCap) continue;
                p[rx] = ry;
                used[x]++;
                used[y]++;
                found[li].push_back({L[x], L[y]});
            }
        }, 1);
        for (const auto& f : found)
            for (const auto& e : f) {
                adj[e.first].push_back(e.second);
                adj[e.second].push_back(e.first);
                bridges[e.first]++;
                bridges[e.second]++;
            }
    }
    // whatever the leaves left unreachable from node 0
    bridgeFromStart(base, n, dim, adj, nthreads);
    return writeGraphLists(out_path, adj);
}
