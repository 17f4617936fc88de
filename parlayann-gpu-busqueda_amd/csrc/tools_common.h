// tools_common.h -- helpers shared by the host graph builders.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

namespace pann {
namespace tools {

int setError(int code, const std::string& msg);

// Exact int32 squared L2 between two int8 rows (auto-vectorised by the compiler).
inline int32_t distI8(const int8_t* __restrict a, const int8_t* __restrict b, uint32_t dim) {
    int32_t s = 0;
    for (uint32_t i = 0; i < dim; i++) {
        const int32_t x = (int32_t)a[i] - (int32_t)b[i];
        s += x * x;
    }
    return s;
}

// Deterministic 64-bit generator (SplitMix64 stream).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t m) { return (uint32_t)((next() >> 11) % m); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline uint64_t mixSeed(uint64_t seed, uint64_t salt) {
    Rng r(seed ^ (salt * 0xd1b54a32d192ed03ull));
    return r.next();
}

// Fisher-Yates permutation of 0..n-1.
inline std::vector<uint32_t> permutation(uint32_t n, uint64_t seed) {
    std::vector<uint32_t> p(n);
    for (uint32_t i = 0; i < n; i++) p[i] = i;
    Rng r(seed);
    for (uint32_t i = n; i > 1; i--) std::swap(p[i - 1], p[r.below(i)]);
    return p;
}

// parallel loop over [0, n) with dynamic chunks; body(i, thread_index)
template <class F>
void parallelFor(size_t n, int nthreads, F&& body, size_t chunk = 16) {
    if (nthreads <= 1 || n <= chunk) {
        for (size_t i = 0; i < n; i++) body(i, 0);
        return;
    }
    const int T = (int)std::min<size_t>((size_t)nthreads, (n + chunk - 1) / chunk);
    std::atomic<size_t> next{0};
    auto worker = [&](int t) {
        while (true) {
            const size_t s = next.fetch_add(chunk);
            if (s >= n) break;
            const size_t e = std::min(n, s + chunk);
            for (size_t i = s; i < e; i++) body(i, t);
        }
    };
    std::vector<std::thread> ts;
    ts.reserve(T - 1);
    for (int t = 1; t < T; t++) ts.emplace_back(worker, t);
    worker(0);
    for (auto& th : ts) th.join();
}

int writeParlayGraph(const char* path, uint32_t n, uint32_t maxDeg, const uint32_t* degrees,
                     const uint32_t* dense, uint32_t stride);

int writeGraphLists(const char* path, const std::vector<std::vector<uint32_t>>& adj);

struct Nbr {
    int32_t d;
    uint32_t id;
    bool operator<(const Nbr& o) const { return d < o.d || (d == o.d && id < o.id); }
    bool operator==(const Nbr& o) const { return d == o.d && id == o.id; }
};

// Marks every node reachable from s (depth-first over the out-lists).
inline void floodFrom(uint32_t s, const std::vector<std::vector<uint32_t>>& adj,
                      std::vector<uint8_t>& reach) {
    std::vector<uint32_t> stack{s};
    reach[s] = 1;
    while (!stack.empty()) {
        const uint32_t u = stack.back();
        stack.pop_back();
        for (uint32_t a : adj[u])
            if (!reach[a]) {
                reach[a] = 1;
                stack.push_back(a);
            }
    }
}

inline bool allReachableFromStart(const std::vector<std::vector<uint32_t>>& adj) {
    std::vector<uint8_t> reach(adj.size(), 0);
    if (!adj.empty()) floodFrom(0, adj, reach);
    return std::find(reach.begin(), reach.end(), 0) == reach.end();
}

// The search always starts at node 0.  For every node not reachable from 0 (in id order),
// adds one edge from its nearest reachable node (rows of `dim` bytes), then marks what
// became reachable.  Returns the number of edges added.
inline size_t bridgeFromStart(const int8_t* base, uint32_t n, uint32_t dim,
                              std::vector<std::vector<uint32_t>>& adj, int nthreads) {
    std::vector<uint8_t> reach(n, 0);
    floodFrom(0, adj, reach);
    size_t added = 0;
    const int T = std::max(1, nthreads);
    for (uint32_t x = 0; x < n; x++) {
        if (reach[x]) continue;
        std::vector<Nbr> best(T, Nbr{INT32_MAX, 0xFFFFFFFFu});
        const int8_t* xv = base + (size_t)x * dim;
        parallelFor(n, T, [&](size_t y, int t) {
            if (!reach[y]) return;
            const Nbr c{distI8(xv, base + y * dim, dim), (uint32_t)y};
            if (c < best[t]) best[t] = c;
        }, 4096);
        const Nbr b = *std::min_element(best.begin(), best.end());
        adj[b.id].push_back(x);
        added++;
        floodFrom(x, adj, reach);
    }
    return added;
}

// ParlayANN graph file from per-node lists ([n][maxDeg][degrees][edges]).  Returns 0,
// 1 (cannot open) or 2 (short write).
inline int writeGraphListsFile(const char* path, const std::vector<std::vector<uint32_t>>& adj) {
    const uint32_t n = (uint32_t)adj.size();
    uint32_t maxDeg = 0;
    for (const auto& a : adj) maxDeg = std::max<uint32_t>(maxDeg, (uint32_t)a.size());
    FILE* f = fopen(path, "wb");
    if (!f) return 1;
    const uint32_t hdr[2] = {n, maxDeg};
    bool ok = fwrite(hdr, 4, 2, f) == 2;
    for (uint32_t i = 0; ok && i < n; i++) {
        const uint32_t d = (uint32_t)adj[i].size();
        ok = fwrite(&d, 4, 1, f) == 1;
    }
    for (uint32_t i = 0; ok && i < n; i++)
        ok = fwrite(adj[i].data(), 4, adj[i].size(), f) == adj[i].size();
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : 2;
}

// Greedy beam search used while building (exact visited set, not ParlayANN's lossy one).
// Returns the expanded nodes with their distances in `visited`.
class BuildSearcher {
   public:
    BuildSearcher(uint32_t n) : mark_(n, 0) {}
    template <class NbrFn>
    void search(const int8_t* base, uint32_t dim, const int8_t* q, const uint32_t* starts,
                uint32_t nstarts, uint32_t L, NbrFn&& neighbours, std::vector<Nbr>& visited,
                std::vector<Nbr>* frontierOut = nullptr) {
        if (++epoch_ == 0) {
            std::fill(mark_.begin(), mark_.end(), 0);
            epoch_ = 1;
        }
        frontier_.clear();
        expanded_.clear();
        visited.clear();
        for (uint32_t i = 0; i < nstarts; i++) {
            const uint32_t s = starts[i];
            if (mark_[s] == epoch_) continue;
            mark_[s] = epoch_;
            frontier_.push_back({distI8(q, base + (size_t)s * dim, dim), s});
        }
        std::sort(frontier_.begin(), frontier_.end());
        if (frontier_.size() > L) frontier_.resize(L);
        expanded_.assign(frontier_.size(), 0);
        size_t pos = 0;
        while (true) {
            while (pos < frontier_.size() && expanded_[pos]) pos++;
            if (pos >= frontier_.size()) break;
            const Nbr cur = frontier_[pos];
            expanded_[pos] = 1;
            visited.push_back(cur);
            size_t lowest = frontier_.size();
            neighbours(cur.id, [&](uint32_t a) {
                if (mark_[a] == epoch_) return;
                mark_[a] = epoch_;
                const int32_t d = distI8(q, base + (size_t)a * dim, dim);
                if (frontier_.size() >= L && !(Nbr{d, a} < frontier_.back())) return;
                const Nbr x{d, a};
                auto it = std::upper_bound(frontier_.begin(), frontier_.end(), x);
                const size_t ip = it - frontier_.begin();
                frontier_.insert(it, x);
                expanded_.insert(expanded_.begin() + ip, 0);
                if (frontier_.size() > L) {
                    frontier_.pop_back();
                    expanded_.pop_back();
                }
                lowest = std::min(lowest, ip);
            });
            pos = std::min(pos, lowest);
        }
        if (frontierOut) *frontierOut = frontier_;
    }

   private:
    std::vector<uint32_t> mark_;
    uint32_t epoch_ = 0;
    std::vector<Nbr> frontier_;
    std::vector<char> expanded_;
};

}  // namespace tools
}  // namespace pann
