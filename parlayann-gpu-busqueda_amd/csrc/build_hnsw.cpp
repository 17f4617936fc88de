// build_hnsw.cpp -- HNSW graph builder (host, multi-threaded), writes the v3 model file.
//
// Algorithm as published (Malkov & Yashunin, HNSW; ParlayANN algorithms/HNSW): node levels
// floor(-ln U * ml); M neighbours per upper layer and 2M on layer 0; insertion descends
// greedily (ef = 1) from the entry point to the node's level, then runs an efc-wide beam
// per layer and keeps neighbours with the distance heuristic (a candidate is dropped when
// alpha * d(candidate, kept) < d(candidate, node)); reverse edges are added and an
// overfull list is re-selected with the same heuristic.
//
// Insertion is batched with prefix doubling (batch sizes 1, 2, 4, ... up to 2% of n),
// each batch searching the graph as it was before the batch, so the graph depends only
// on the seed and not on the thread count.  Node i holds base point i.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/pann_tools.h"
#include "hnsw_plan.h"
#include "tools_common.h"
#include <cstdlib>
#include <chrono>

namespace pann {
namespace tools {
namespace {

struct Hnsw {
    const int8_t* base;
    uint32_t n, dim, M, M0, efc;
    float alpha;
    int nthreads;
    std::vector<uint32_t> level;
    std::vector<uint64_t> upStart;  // first upper block of node i (level >= 1)
    std::vector<uint32_t> adj0, deg0;    // n x M0
    std::vector<uint32_t> adjUp, degUp;  // blocks x M
    uint32_t entry = 0, top = 0;

    const int8_t* vec(uint32_t i) const { return base + (size_t)i * dim; }
    uint32_t cap(uint32_t l) const { return l == 0 ? M0 : M; }
    uint32_t* row(uint32_t node, uint32_t l) {
        return l == 0 ? adj0.data() + (size_t)node * M0
                      : adjUp.data() + (upStart[node] + l - 1) * M;
    }
    uint32_t& deg(uint32_t node, uint32_t l) {
        return l == 0 ? deg0[node] : degUp[upStart[node] + l - 1];
    }

    // heuristic selection over candidates sorted by distance to `p`
    void select(uint32_t p, std::vector<Nbr>& cand, uint32_t m, std::vector<uint32_t>& out) const {
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end(),
                               [](const Nbr& a, const Nbr& b) { return a.id == b.id; }),
                   cand.end());
        out.clear();
        std::vector<uint32_t> kept;
        for (const Nbr& c : cand) {
            if (out.size() >= m) break;
            if (c.id == p) continue;
            bool good = true;
            for (uint32_t s : out) {
                const float dcs = (float)distI8(vec(c.id), vec(s), dim);
                if (alpha * dcs < (float)c.d) {
                    good = false;
                    break;
                }
            }
            if (good) out.push_back(c.id);
        }
    }

    void insertBatch(const uint32_t* batch, size_t size, std::vector<BuildSearcher>& searchers) {
        // out[i][l] for l <= min(level, top)
        std::vector<std::vector<std::vector<uint32_t>>> out(size);
        parallelFor(size, nthreads, [&](size_t i, int t) {
            const uint32_t p = batch[i];
            const uint32_t lp = level[p];
            std::vector<Nbr> visited, W;
            std::vector<uint32_t> eps{entry};
            for (int l = (int)top; l > (int)lp; l--) {
                searchers[t].search(base, dim, vec(p), eps.data(), (uint32_t)eps.size(), 1,
                                    [&](uint32_t u, auto&& visit) {
                                        const uint32_t* r = row(u, l);
                                        for (uint32_t j = 0; j < deg(u, l); j++) visit(r[j]);
                                    },
                                    visited, &W);
                eps.assign(1, W[0].id);
            }
            out[i].resize(std::min(lp, top) + 1);
            for (int l = (int)std::min(lp, top); l >= 0; l--) {
                searchers[t].search(base, dim, vec(p), eps.data(), (uint32_t)eps.size(), efc,
                                    [&](uint32_t u, auto&& visit) {
                                        const uint32_t* r = row(u, l);
                                        for (uint32_t j = 0; j < deg(u, l); j++) visit(r[j]);
                                    },
                                    visited, &W);
                std::vector<Nbr> cand = W;
                select(p, cand, cap(l) == M0 ? M0 : M, out[i][l]);
                eps.clear();
                for (const Nbr& w : W) eps.push_back(w.id);
            }
        }, 1);
        for (size_t i = 0; i < size; i++) {
            const uint32_t p = batch[i];
            for (uint32_t l = 0; l < out[i].size(); l++) {
                deg(p, l) = (uint32_t)out[i][l].size();
                std::copy(out[i][l].begin(), out[i][l].end(), row(p, l));
            }
        }
        // reverse edges grouped by (layer, target)
        struct Rev {
            uint32_t l, t, s;
            bool operator<(const Rev& o) const {
                return l != o.l ? l < o.l : t != o.t ? t < o.t : s < o.s;
            }
        };
        std::vector<Rev> rev;
        for (size_t i = 0; i < size; i++)
            for (uint32_t l = 0; l < out[i].size(); l++)
                for (uint32_t t : out[i][l]) rev.push_back({l, t, batch[i]});
        std::sort(rev.begin(), rev.end());
        std::vector<size_t> starts;
        for (size_t i = 0; i < rev.size(); i++)
            if (i == 0 || rev[i].l != rev[i - 1].l || rev[i].t != rev[i - 1].t) starts.push_back(i);
        starts.push_back(rev.size());
        parallelFor(starts.size() - 1, nthreads, [&](size_t gi, int) {
            const uint32_t l = rev[starts[gi]].l, t = rev[starts[gi]].t;
            uint32_t* r = row(t, l);
            uint32_t& d = deg(t, l);
            std::vector<uint32_t> add;
            for (size_t j = starts[gi]; j < starts[gi + 1]; j++) {
                const uint32_t s = rev[j].s;
                if (s == t) continue;
                bool present = false;
                for (uint32_t x = 0; x < d && !present; x++) present = r[x] == s;
                if (!present) add.push_back(s);
            }
            if (add.empty()) return;
            if (d + add.size() <= cap(l)) {
                for (uint32_t s : add) r[d++] = s;
            } else {
                std::vector<Nbr> cand;
                for (uint32_t x = 0; x < d; x++) cand.push_back({distI8(vec(t), vec(r[x]), dim), r[x]});
                for (uint32_t s : add) cand.push_back({distI8(vec(t), vec(s), dim), s});
                std::vector<uint32_t> sel;
                select(t, cand, cap(l), sel);
                d = (uint32_t)sel.size();
                std::copy(sel.begin(), sel.end(), r);
            }
        }, 4);
    }

    int write(const char* path, float ml) {
        const int r = writeHnswV3(path, n, dim, ml, M, efc, alpha, level,
                                  [&](uint32_t i, uint32_t l) { return deg(i, l); },
                                  [&](uint32_t i, uint32_t l) { return row(i, l); }, entry);
        if (r == 1) return setError(2, std::string("cannot write ") + path);
        if (r != 0) return setError(2, std::string("short write to ") + path);
        return 0;
    }
};

}  // namespace
}  // namespace tools
}  // namespace pann

extern "C" int pann_build_hnsw(const int8_t* base, uint32_t n, uint32_t dim, uint32_t M,
                               uint32_t efc, float ml, float alpha, uint64_t seed, int nthreads,
                               const char* out_path) {
    using namespace pann::tools;
    if (!base || !out_path || n == 0 || dim == 0 || M == 0 || efc == 0 || !(ml > 0))
        return setError(1, "invalid argument");
    Hnsw h;
    h.base = base;
    h.n = n;
    h.dim = dim;
    h.M = M;
    h.M0 = 2 * M;
    h.efc = efc;
    h.alpha = alpha > 0 ? alpha : 1.0f;
    h.nthreads = std::max(1, nthreads);
    HnswPlan plan = hnswPlan(n, ml, seed);
    h.level = std::move(plan.level);
    h.upStart = std::move(plan.upStart);
    const uint64_t blocks = plan.blocks;
    h.adj0.assign((size_t)n * h.M0, 0xFFFFFFFFu);
    h.deg0.assign(n, 0);
    h.adjUp.assign(std::max<uint64_t>(1, blocks) * M, 0xFFFFFFFFu);
    h.degUp.assign(std::max<uint64_t>(1, blocks), 0);
    h.entry = plan.entry;
    h.top = plan.top;

    const std::vector<uint32_t>& order = plan.order;
    std::vector<BuildSearcher> searchers;
    for (int t = 0; t < h.nthreads; t++) searchers.emplace_back(n);
    const uint32_t maxBatch = plan.maxBatch;
    size_t count = 0;
    uint32_t bsize = 1;
    // PANN_BUILD_PROGRESS=1: a line on stderr at every tenth of the points inserted
    const char* pe = getenv("PANN_BUILD_PROGRESS");
    const bool progress = pe && pe[0] == '1';
    const auto t0 = std::chrono::steady_clock::now();
    size_t nextReport = order.size() / 10;
    while (count < order.size()) {
        const size_t size = std::min<size_t>(bsize, order.size() - count);
        h.insertBatch(order.data() + count, size, searchers);
        count += size;
        bsize = std::min(maxBatch, bsize * 2);
        if (progress && count >= nextReport) {
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            fprintf(stderr, "[hnsw build] %zu of %zu points inserted, %.1f s\n", count, order.size(), s);
            fflush(stderr);
            nextReport += order.size() / 10;
        }
    }
    return h.write(out_path, ml);
}
