// build_trees.h -- random two-pivot cluster trees used by the HCNNG and PyNNDescent builders.
#pragma once

#include <stdint.h>

#include <vector>

#include "tools_common.h"

namespace pann {
namespace tools {

// Splits ids[0..n) into leaves of at most leafSize points: each inner node draws two
// pivots and sends every point to the nearer one (ties to the first); a degenerate split
// halves the segment.  Returns leaf boundaries into the permuted ids (deterministic for a
// seed; segments of one level are split in parallel).
inline std::vector<std::pair<uint32_t, uint32_t>> clusterTreeLeaves(const int8_t* base, uint32_t dim,
                                                                    std::vector<uint32_t>& ids,
                                                                    uint32_t leafSize,
                                                                    uint64_t seed, int nthreads) {
    std::vector<std::pair<uint32_t, uint32_t>> leaves, todo{{0u, (uint32_t)ids.size()}};
    std::vector<uint32_t> scratch(ids.size());
    uint64_t round = 0;
    while (!todo.empty()) {
        std::vector<std::pair<uint32_t, uint32_t>> next;
        std::vector<std::pair<uint32_t, uint32_t>> split;
        for (auto s : todo) {
            if (s.second - s.first <= leafSize) leaves.push_back(s);
            else split.push_back(s);
        }
        std::vector<uint32_t> mid(split.size());
        parallelFor(split.size(), nthreads, [&](size_t si, int) {
            const uint32_t b = split[si].first, e = split[si].second, m = e - b;
            Rng r(mixSeed(seed, (round << 32) ^ b));
            const uint32_t pa = ids[b + r.below(m)];
            uint32_t pb = ids[b + r.below(m)];
            for (int tries = 0; pb == pa && tries < 8; tries++) pb = ids[b + r.below(m)];
            uint32_t lo = b, hi = e;
            for (uint32_t i = b; i < e; i++) {
                const uint32_t x = ids[i];
                const int32_t da = distI8(base + (size_t)x * dim, base + (size_t)pa * dim, dim);
                const int32_t db = distI8(base + (size_t)x * dim, base + (size_t)pb * dim, dim);
                if (da <= db) scratch[lo++] = x;
                else scratch[--hi] = x;
            }
            // keep the far side in input order
            std::reverse(scratch.begin() + hi, scratch.begin() + e);
            uint32_t cut = lo;
            if (cut == b || cut == e) cut = b + m / 2;
            std::copy(scratch.begin() + b, scratch.begin() + e, ids.begin() + b);
            mid[si] = cut;
        }, 1);
        for (size_t si = 0; si < split.size(); si++) {
            next.push_back({split[si].first, mid[si]});
            next.push_back({mid[si], split[si].second});
        }
        todo.swap(next);
        round++;
    }
    std::sort(leaves.begin(), leaves.end());
    return leaves;
}

}  // namespace tools
}  // namespace pann
