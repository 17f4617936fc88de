// oracle.cpp -- CPU restatement of ParlayANN's beam search and HNSW search, written
// from the reference's own restatement in its CUDA kernels (lines cited per function).
// TEST INFRASTRUCTURE ONLY (tests/, smoke check, CPU-baseline leg of bench.py).
//
// The flat search is written the way ParlayANN writes it (sorted std::vector
// frontier, std::set_union / std::set_difference, sorted visited list), not the way a
// GPU kernel would, so that it checks the kernel rather than mirroring it.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_set>
#include <vector>

namespace {

// ParlayANN's id_dist = std::pair<indexType, float>; distances of int8 points are exact
// integers stored as float (all < 2^24, so exactly representable).
struct IdDist {
    uint32_t id;
    float d;
};

// less(a, b) = a.d < b.d || (a.d == b.d && a.id < b.id)        cuda/utiles_gpu.cuh:33-36
inline bool lessPair(const IdDist& a, const IdDist& b) {
    return a.d < b.d || (a.d == b.d && a.id < b.id);
}

inline int32_t dist(const int8_t* a, const int8_t* b, uint32_t dim) {
    int32_t s = 0;
    for (uint32_t i = 0; i < dim; i++) {
        int32_t x = (int32_t)a[i] - (int32_t)b[i];
        s += x * x;
    }
    return s;
}

template <class F>
void parallelFor(uint32_t n, int nthreads, F&& body) {
    if (nthreads <= 1 || n < 2) {
        for (uint32_t i = 0; i < n; i++) body(i);
        return;
    }
    std::atomic<uint32_t> next{0};
    const uint32_t chunk = 16;
    auto worker = [&]() {
        while (true) {
            uint32_t s = next.fetch_add(chunk);
            if (s >= n) break;
            uint32_t e = std::min(n, s + chunk);
            for (uint32_t i = s; i < e; i++) body(i);
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; t++) ts.emplace_back(worker);
    for (auto& t : ts) t.join();
}

}  // namespace

extern "C" {

// cuda/utiles_gpu.cuh:12-17 (parlaylib hash64_2)
uint64_t oracle_hash64_2(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    x = x ^ (x >> 31);
    return x;
}

// cuda/utiles_gpu.cuh:21-31 (parlaylib hash64, Numerical Recipes Ranhash)
uint64_t oracle_hash64(uint64_t u) {
    uint64_t v = u * 3935559000370003845ull + 2691343689449507681ull;
    v ^= v >> 21;
    v ^= v << 37;
    v ^= v >> 4;
    v *= 4768777513237032717ull;
    v ^= v << 20;
    v ^= v >> 41;
    v ^= v << 5;
    return v;
}

// cuda/lectores.cpp:172-179
int32_t oracle_dist_i8(const int8_t* a, const int8_t* b, uint32_t dim) { return dist(a, b, dim); }

// ParlayANN beam_search for one query, as restated by cuda/busqueda.cu:95-506.
//   - start point is always node 0 (busqueda.cu:166-173); it is not written to the filter
//   - neighbours equal to the query index are skipped and not written (busqueda.cu:214-231)
//   - lossy filter: slot = hash64_2(a) & (2^bits - 1), bits = max(10, ceil(log2(Ls*Ls)) - 2)
//     (busqueda.cu:536-538)
//   - cutoff = frontier full ? last distance : INT_MAX, taken before the merge (:211-212)
//   - multiset set_union, trim to Ls, then cut at (float)(cut * (double)F[k].d) (:364-464)
//   - next = first element of set_difference(frontier, visited) (:466-484)
int oracle_beam_flat(const int8_t* base, uint32_t n, uint32_t dim, const uint32_t* degrees,
                     const uint64_t* offsets, const uint32_t* edges, const int8_t* query,
                     const oracle_query_params* p, uint32_t* out_ids, int32_t* out_dists,
                     int32_t* out_comps, int32_t* out_flags, int64_t* out_edges,
                     int32_t* out_hops) {
    if (n == 0 || p->Ls < 1 || p->k < 0) return 1;
    const int Ls = p->Ls;
    const int64_t limit = (p->visit_limit > 0) ? p->visit_limit : (int64_t)n;
    const uint32_t qid = p->qid;

    const int bits = std::max(10, (int)std::ceil(std::log2((double)(Ls * Ls))) - 2);
    const uint32_t mask = (1u << bits) - 1;
    std::vector<uint32_t> filter((size_t)1 << bits, 0xFFFFFFFFu);
    auto has_been_seen = [&](uint32_t a) -> bool {
        uint32_t loc = (uint32_t)(oracle_hash64_2(a) & mask);
        if (filter[loc] == a) return true;
        filter[loc] = a;
        return false;
    };

    std::vector<IdDist> frontier;
    frontier.push_back({0u, (float)dist(query, base, dim)});
    std::vector<IdDist> unvisited(frontier);
    std::vector<IdDist> visited;
    std::vector<IdDist> candidates, newFrontier;
    std::vector<uint32_t> filtered;
    int64_t comps = 1, numVisited = 0, edgesRead = 0;

    while (!unvisited.empty() && numVisited < limit) {
        IdDist current = unvisited[0];
        visited.insert(std::upper_bound(visited.begin(), visited.end(), current, lessPair),
                       current);
        numVisited++;
        const bool frontierFull = (int)frontier.size() == Ls;

        filtered.clear();
        const uint32_t deg = degrees[current.id];
        const uint32_t* row = edges + offsets[current.id];
        edgesRead += deg;
        for (uint32_t i = 0; i < deg; i++) {
            uint32_t a = row[i];
            if (a == qid || has_been_seen(a)) continue;
            filtered.push_back(a);
        }
        const float cutoff = frontierFull ? frontier.back().d : (float)INT_MAX;
        candidates.clear();
        for (uint32_t a : filtered) {
            float d = (float)dist(query, base + (size_t)a * dim, dim);
            comps++;
            if (d >= cutoff) continue;
            candidates.push_back({a, d});
        }
        std::sort(candidates.begin(), candidates.end(), lessPair);

        newFrontier.resize(frontier.size() + candidates.size());
        size_t nfs = std::set_union(frontier.begin(), frontier.end(), candidates.begin(),
                                    candidates.end(), newFrontier.begin(), lessPair) -
                     newFrontier.begin();
        nfs = std::min<size_t>((size_t)Ls, nfs);
        if (p->k > 0 && nfs > (size_t)p->k) {
            IdDist thr{0u, (float)(p->cut * (double)newFrontier[p->k].d)};
            nfs = std::upper_bound(newFrontier.begin(), newFrontier.begin() + nfs, thr,
                                   lessPair) -
                  newFrontier.begin();
        }
        frontier.assign(newFrontier.begin(), newFrontier.begin() + nfs);

        unvisited.resize(frontier.size());
        size_t rem = std::set_difference(frontier.begin(), frontier.end(), visited.begin(),
                                         visited.end(), unvisited.begin(), lessPair) -
                     unvisited.begin();
        unvisited.resize(rem);
    }

    int flags = 0;
    if ((int)frontier.size() < p->k) flags |= ORACLE_FLAG_SHORT;
    for (int j = 0; j < p->k; j++) {
        bool have = j < (int)frontier.size();
        out_ids[j] = have ? frontier[j].id : ORACLE_NO_NEIGHBOR;
        out_dists[j] = have ? (int32_t)frontier[j].d : INT_MAX;
    }
    if (out_comps) *out_comps = (int32_t)comps;
    if (out_flags) *out_flags = flags;
    if (out_edges) *out_edges = edgesRead;
    if (out_hops) *out_hops = (int32_t)numVisited;
    return 0;
}

int oracle_beam_flat_batch(const int8_t* base, uint32_t n, uint32_t dim,
                           const uint32_t* degrees, const uint64_t* offsets,
                           const uint32_t* edges, const int8_t* queries, uint32_t nq,
                           const oracle_query_params* params, int shared_params,
                           uint32_t out_stride, uint32_t* out_ids, int32_t* out_dists,
                           int32_t* out_comps, int32_t* out_flags, int64_t* out_edges,
                           int32_t* out_hops, int nthreads) {
    std::atomic<int> err{0};
    parallelFor(nq, nthreads, [&](uint32_t q) {
        const oracle_query_params* p = shared_params ? params : params + q;
        oracle_query_params pq = *p;
        if (shared_params) pq.qid = q;
        int r = oracle_beam_flat(base, n, dim, degrees, offsets, edges,
                                 queries + (size_t)q * dim, &pq, out_ids + (size_t)q * out_stride,
                                 out_dists + (size_t)q * out_stride,
                                 out_comps ? out_comps + q : nullptr,
                                 out_flags ? out_flags + q : nullptr,
                                 out_edges ? out_edges + q : nullptr,
                                 out_hops ? out_hops + q : nullptr);
        if (r) err = r;
    });
    return err.load();
}

// ParlayANN HNSW::search as restated by cuda/busqueda_hnsw.cu:179-457 and
// cuda/busqueda_hnsw.cu:551-559, 609-610 (parameters):
//   layers top..1: ef = 1, 2-bit visited filter, no visit limit; layer 0: ef = Ls,
//   bits = Ls > 2 ? 2*ceil(log2 Ls) - 2 : 2, limit = visit (or n).
//   entry points are marked with hash64, neighbours with hash64_2 (:250-264, :286-336).
//   cand is ordered and unique by distance only (:79-90); W is ordered by (d, id).
//   acceptance is sequential (:393-418); the next layer starts from W's farthest (:434-440).
int oracle_search_hnsw(const int8_t* base, uint32_t dim, const oracle_hnsw* h,
                       const int8_t* query, const oracle_query_params* p, uint32_t* out_ids,
                       int32_t* out_dists, int32_t* out_comps, int32_t* out_flags,
                       int64_t* out_edges, int32_t* out_hops) {
    if (h->num_entries < 1 || p->Ls < 1 || p->k < 0) return 1;
    const uint32_t n = h->n;
    const int Ls = p->Ls;
    const int64_t lim0 = (p->visit_limit > 0 && p->visit_limit < (int64_t)n) ? p->visit_limit
                                                                             : (int64_t)n;
    auto nodeDist = [&](uint32_t node) -> int32_t {
        return dist(query, base + (size_t)h->node_point[node] * dim, dim);
    };
    std::vector<uint32_t> eps(h->entries, h->entries + h->num_entries);
    const int top = (int)h->level[eps[0]];
    for (uint32_t e : eps)
        if (e >= n || (int)h->level[e] != top) return 1;  // entry points live on the top layer
    int64_t comps = 0, edgesRead = 0, hops = 0;
    std::vector<IdDist> W;

    for (int layer = top; layer >= 0; layer--) {
        const int ef = layer > 0 ? 1 : Ls;
        int bits = 2;
        if (layer == 0 && Ls > 2) bits = (int)std::ceil(std::log2((double)Ls)) * 2 - 2;
        const uint64_t mask = ((uint64_t)1 << bits) - 1;
        const int64_t limit = layer > 0 ? INT64_MAX : lim0;
        std::vector<uint32_t> vis((size_t)1 << bits, 0xFFFFFFFFu);
        std::unordered_set<uint32_t> inw;
        std::vector<IdDist> cand;  // sorted by d, unique d
        W.clear();

        auto candInsert = [&](const IdDist& x) {
            auto it = std::lower_bound(cand.begin(), cand.end(), x,
                                       [](const IdDist& a, const IdDist& b) { return a.d < b.d; });
            if (it != cand.end() && it->d == x.d) return;
            cand.insert(it, x);
        };
        auto wInsert = [&](const IdDist& x) {
            W.insert(std::upper_bound(W.begin(), W.end(), x, lessPair), x);
        };

        for (uint32_t e : eps) {
            vis[oracle_hash64(e) & mask] = e;
            inw.insert(e);
            IdDist x{e, (float)nodeDist(e)};
            candInsert(x);
            wInsert(x);
            comps++;
        }
        int64_t cnt = 0;
        while (!cand.empty()) {
            if (cand.front().d > W.back().d) break;
            if (++cnt > limit) break;
            const uint32_t u = cand.front().id;
            cand.erase(cand.begin());
            hops++;
            const uint64_t li = h->node_first_list[u] + (uint64_t)layer;
            const uint32_t* row = h->list_ids + h->list_offsets[li];
            const uint32_t deg = h->list_sizes[li];
            edgesRead += deg;
            std::vector<uint32_t> P;
            for (uint32_t i = 0; i < deg; i++) {
                uint32_t a = row[i];
                uint64_t slot = oracle_hash64_2(a) & mask;
                if (vis[slot] == a) continue;
                vis[slot] = a;
                P.push_back(a);
            }
            comps += (int64_t)P.size();
            for (uint32_t a : P) {
                float d = (float)nodeDist(a);
                if (!((int)W.size() < ef || d < W.back().d)) continue;
                if (!inw.insert(a).second) continue;
                IdDist x{a, d};
                candInsert(x);
                wInsert(x);
                if ((int)W.size() > ef) W.pop_back();
                if ((int)cand.size() > ef) cand.pop_back();
            }
        }
        if (layer > 0) eps.assign(1, W.back().id);
    }

    int flags = 0;
    if ((int)W.size() < p->k) flags |= ORACLE_FLAG_SHORT;
    for (int j = 0; j < p->k; j++) {
        bool have = j < (int)W.size();
        out_ids[j] = have ? h->node_point[W[j].id] : ORACLE_NO_NEIGHBOR;
        out_dists[j] = have ? (int32_t)W[j].d : INT_MAX;
    }
    if (out_comps) *out_comps = (int32_t)comps;
    if (out_flags) *out_flags = flags;
    if (out_edges) *out_edges = edgesRead;
    if (out_hops) *out_hops = (int32_t)hops;
    return 0;
}

int oracle_search_hnsw_batch(const int8_t* base, uint32_t dim, const oracle_hnsw* h,
                             const int8_t* queries, uint32_t nq,
                             const oracle_query_params* params, int shared_params,
                             uint32_t out_stride, uint32_t* out_ids, int32_t* out_dists,
                             int32_t* out_comps, int32_t* out_flags, int64_t* out_edges,
                             int32_t* out_hops, int nthreads) {
    std::atomic<int> err{0};
    parallelFor(nq, nthreads, [&](uint32_t q) {
        const oracle_query_params* p = shared_params ? params : params + q;
        int r = oracle_search_hnsw(base, dim, h, queries + (size_t)q * dim, p,
                                   out_ids + (size_t)q * out_stride,
                                   out_dists + (size_t)q * out_stride,
                                   out_comps ? out_comps + q : nullptr,
                                   out_flags ? out_flags + q : nullptr,
                                   out_edges ? out_edges + q : nullptr,
                                   out_hops ? out_hops + q : nullptr);
        if (r) err = r;
    });
    return err.load();
}

int oracle_bruteforce_knn(const int8_t* base, uint32_t n, uint32_t dim, const int8_t* queries,
                          uint32_t nq, uint32_t k, uint32_t* out_ids, int32_t* out_dists,
                          int nthreads) {
    if (k > n) return 1;
    parallelFor(nq, nthreads, [&](uint32_t q) {
        std::vector<std::pair<int32_t, uint32_t>> all(n);
        const int8_t* qv = queries + (size_t)q * dim;
        for (uint32_t i = 0; i < n; i++) all[i] = {dist(qv, base + (size_t)i * dim, dim), i};
        std::partial_sort(all.begin(), all.begin() + k, all.end());
        for (uint32_t j = 0; j < k; j++) {
            out_ids[(size_t)q * k + j] = all[j].second;
            out_dists[(size_t)q * k + j] = all[j].first;
        }
    });
    return 0;
}

// cuda/evaluacion.cpp:56-80
double oracle_recall(const uint32_t* gt_ids, const float* gt_dists, uint32_t nq, uint32_t gt_k,
                     const uint32_t* ids, uint32_t ids_stride, uint32_t k) {
    long long total = 0;
    for (uint32_t q = 0; q < nq; q++) {
        const uint32_t* truth = gt_ids + (size_t)q * gt_k;
        const float* dTruth = gt_dists + (size_t)q * gt_k;
        int end = (int)gt_k;
        if (gt_k != k) {
            float ref = dTruth[k - 1];
            for (int i = (int)k; i < (int)gt_k; i++)
                if (!(std::fabs(ref - dTruth[i]) < 1e-6f)) {
                    end = i;
                    break;
                }
        }
        std::vector<uint32_t> mine(ids + (size_t)q * ids_stride, ids + (size_t)q * ids_stride + k);
        std::sort(mine.begin(), mine.end());
        mine.erase(std::unique(mine.begin(), mine.end()), mine.end());
        for (uint32_t x : mine)
            if (std::find(truth, truth + end, x) != truth + end) total++;
    }
    return (double)total / nq / k;
}

}  // extern "C"
