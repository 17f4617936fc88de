// build_vamana.cpp -- DiskANN/Vamana graph builder (host, multi-threaded).
//
// Algorithm as published for ParlayANN's Vamana (Manohar et al., PPoPP'24, and the
// DiskANN paper): points are inserted in a random order in batches whose size doubles
// from 1 up to 2% of n ("prefix doubling"); each point of a batch runs a beam search of
// width L from start point 0 on the graph as it was before the batch, its visited set is
// robust-pruned with alpha to at most R out-neighbours, then reverse edges are added and
// any node pushed over R is pruned again.  With two_pass the whole insertion runs once
// with alpha = 1 and once with the requested alpha.  The search always starts at node 0
// (cuda/busqueda.cu:166-173), so the build uses node 0 as its start point as well.
//
// Every phase is deterministic for a seed: searches read the pre-batch graph, prunes
// write only their own row, reverse edges are grouped by target in sorted order.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/pann_tools.h"
#include "tools_common.h"

namespace pann {
namespace tools {
namespace {

struct Vamana {
    const int8_t* base;
    uint32_t n, dim, R, L;
    int nthreads;
    std::vector<uint32_t> adj;  // n x R
    std::vector<uint32_t> deg;

    const int8_t* vec(uint32_t i) const { return base + (size_t)i * dim; }

    // ParlayANN robustPrune: candidates plus p's current out-neighbours, nearest first;
    // keep p* and drop every p' with alpha * d(p*, p') <= d(p, p').
    void robustPrune(uint32_t p, std::vector<Nbr>& cand, float alpha, std::vector<uint32_t>& out,
                     bool addCurrent) const {
        if (addCurrent)
            for (uint32_t j = 0; j < deg[p]; j++) {
                const uint32_t a = adj[(size_t)p * R + j];
                cand.push_back({distI8(vec(p), vec(a), dim), a});
            }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end(),
                               [](const Nbr& x, const Nbr& y) { return x.id == y.id; }),
                   cand.end());
        out.clear();
        std::vector<char> pruned(cand.size(), 0);
        for (size_t i = 0; i < cand.size() && out.size() < R; i++) {
            if (pruned[i] || cand[i].id == p) continue;
            const uint32_t s = cand[i].id;
            out.push_back(s);
            for (size_t j = i + 1; j < cand.size(); j++) {
                if (pruned[j] || cand[j].id == p) continue;
                const float dsp = (float)distI8(vec(s), vec(cand[j].id), dim);
                if (alpha * dsp <= (float)cand[j].d) pruned[j] = 1;
            }
        }
    }

    void insertPass(const std::vector<uint32_t>& order, float alpha) {
        const uint32_t maxBatch = std::max<uint32_t>(1, (uint32_t)(0.02 * n));
        std::vector<BuildSearcher> searchers;
        for (int t = 0; t < nthreads; t++) searchers.emplace_back(n);
        std::vector<std::vector<uint32_t>> newOut;
        size_t count = 0;
        uint32_t bsize = 1;
        while (count < order.size()) {
            const size_t size = std::min<size_t>(bsize, order.size() - count);
            bsize = std::min(maxBatch, bsize * 2);
            const uint32_t* batch = order.data() + count;
            count += size;
            newOut.assign(size, {});
            // 1. search + prune on the pre-batch graph
            parallelFor(size, nthreads, [&](size_t i, int t) {
                const uint32_t p = batch[i];
                std::vector<Nbr> visited;
                const uint32_t start = 0;
                searchers[t].search(base, dim, vec(p), &start, 1, L,
                                    [&](uint32_t u, auto&& visit) {
                                        const uint32_t* row = adj.data() + (size_t)u * R;
                                        for (uint32_t j = 0; j < deg[u]; j++) visit(row[j]);
                                    },
                                    visited);
                robustPrune(p, visited, alpha, newOut[i], true);
            }, 1);
            // 2. install the new out-neighbourhoods
            for (size_t i = 0; i < size; i++) {
                const uint32_t p = batch[i];
                deg[p] = (uint32_t)newOut[i].size();
                std::copy(newOut[i].begin(), newOut[i].end(), adj.begin() + (size_t)p * R);
                std::fill(adj.begin() + (size_t)p * R + deg[p], adj.begin() + (size_t)(p + 1) * R, 0xFFFFFFFFu);
            }
            // 3. reverse edges, grouped by target
            std::vector<std::pair<uint32_t, uint32_t>> rev;
            for (size_t i = 0; i < size; i++)
                for (uint32_t t : newOut[i]) rev.push_back({t, batch[i]});
            std::sort(rev.begin(), rev.end());
            std::vector<size_t> starts;
            for (size_t i = 0; i < rev.size(); i++)
                if (i == 0 || rev[i].first != rev[i - 1].first) starts.push_back(i);
            starts.push_back(rev.size());
            parallelFor(starts.size() - 1, nthreads, [&](size_t gi, int) {
                const uint32_t t = rev[starts[gi]].first;
                std::vector<uint32_t> add;
                for (size_t j = starts[gi]; j < starts[gi + 1]; j++) {
                    const uint32_t s = rev[j].second;
                    bool present = false;
                    for (uint32_t x = 0; x < deg[t] && !present; x++)
                        present = adj[(size_t)t * R + x] == s;
                    if (!present && s != t) add.push_back(s);
                }
                if (add.empty()) return;
                if (deg[t] + add.size() <= R) {
                    for (uint32_t s : add) adj[(size_t)t * R + deg[t]++] = s;
                } else {
                    std::vector<Nbr> cand;
                    for (uint32_t s : add) cand.push_back({distI8(vec(t), vec(s), dim), s});
                    std::vector<uint32_t> out;
                    robustPrune(t, cand, alpha, out, true);
                    deg[t] = (uint32_t)out.size();
                    std::copy(out.begin(), out.end(), adj.begin() + (size_t)t * R);
                    std::fill(adj.begin() + (size_t)t * R + deg[t], adj.begin() + (size_t)(t + 1) * R, 0xFFFFFFFFu);
                }
            }, 4);
        }
    }
};

}  // namespace
}  // namespace tools
}  // namespace pann

extern "C" int pann_build_vamana(const int8_t* base, uint32_t n, uint32_t dim, uint32_t R,
                                 uint32_t L, float alpha, int two_pass, uint64_t seed,
                                 int nthreads, uint32_t* out_degrees, uint32_t* out_edges) {
    using namespace pann::tools;
    if (!base || !out_degrees || !out_edges || n == 0 || dim == 0 || R == 0 || L == 0)
        return setError(1, "invalid argument");
    Vamana v{base, n, dim, R, L, std::max(1, nthreads), {}, {}};
    v.adj.assign((size_t)n * R, 0xFFFFFFFFu);
    v.deg.assign(n, 0);
    const std::vector<uint32_t> order = permutation(n, mixSeed(seed, 1));
    if (two_pass) v.insertPass(order, 1.0f);
    v.insertPass(order, alpha);
    memcpy(out_degrees, v.deg.data(), 4ull * n);
    memcpy(out_edges, v.adj.data(), 4ull * n * R);
    return 0;
}
