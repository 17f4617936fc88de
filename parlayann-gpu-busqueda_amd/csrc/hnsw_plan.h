// hnsw_plan.h -- what the HNSW builders share: the seeded node levels, layout and insertion
// order, and the writer of the HNSW model v3 file (the layout cuda/hnsw_lector.cpp:28-111
// reads).  Host code only.
#pragma once

#include <stdint.h>
#include <stdio.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "tools_common.h"

namespace pann {
namespace tools {

struct HnswPlan {
    std::vector<uint32_t> level;    // floor(-ln U * ml), at most 30
    std::vector<uint64_t> upStart;  // first upper-layer block of each node (level >= 1)
    uint64_t blocks = 0;            // upper-layer rows over all nodes
    uint32_t entry = 0, top = 0;    // first node of the highest level, and that level
    std::vector<uint32_t> order;    // insertion order: a seeded permutation without the entry
    uint32_t maxBatch = 1;          // prefix doubling stops at 2% of n
};

inline HnswPlan hnswPlan(uint32_t n, float ml, uint64_t seed) {
    HnswPlan h;
    h.level.resize(n);
    Rng lr(mixSeed(seed, 2));
    for (uint32_t i = 0; i < n; i++) {
        double u = 1.0 - lr.uniform();  // (0, 1]
        h.level[i] = (uint32_t)std::min(30.0, std::floor(-std::log(u) * ml));
    }
    h.upStart.assign(n, 0);
    for (uint32_t i = 0; i < n; i++) {
        h.upStart[i] = h.blocks;
        h.blocks += h.level[i];
    }
    for (uint32_t i = 1; i < n; i++)
        if (h.level[i] > h.level[h.entry]) h.entry = i;
    h.top = h.level[h.entry];
    h.order = permutation(n, mixSeed(seed, 3));
    h.order.erase(std::find(h.order.begin(), h.order.end(), h.entry));
    h.maxBatch = std::max<uint32_t>(1, (uint32_t)(0.02 * n));
    return h;
}

// Writes the v3 model: header, (level, point) per node, every node's lists from layer 0 up
// (count, ids), one entry point.  deg(i, l) and row(i, l) give node i's list on layer l.
// Returns 0, 1 (cannot open) or 2 (short write).
template <class DegFn, class RowFn>
int writeHnswV3(const char* path, uint32_t n, uint32_t dim, float ml, uint32_t M, uint32_t efc,
                float alpha, const std::vector<uint32_t>& level, DegFn deg, RowFn row, uint32_t entry) {
    FILE* f = fopen(path, "wb");
    if (!f) return 1;
    bool ok = fwrite("HNSW", 1, 4, f) == 4;
    const uint32_t version = 3;
    const uint64_t zero = 0;
    ok = ok && fwrite(&version, 4, 1, f) == 1 && fwrite(&zero, 8, 1, f) == 1 && fwrite(&zero, 8, 1, f) == 1;
    ok = ok && fwrite(&dim, 4, 1, f) == 1 && fwrite(&ml, 4, 1, f) == 1 && fwrite(&M, 4, 1, f) == 1 &&
         fwrite(&efc, 4, 1, f) == 1 && fwrite(&alpha, 4, 1, f) == 1 && fwrite(&n, 4, 1, f) == 1;
    for (uint32_t i = 0; ok && i < n; i++) {
        const uint32_t rec[2] = {level[i], i};
        ok = fwrite(rec, 4, 2, f) == 2;
    }
    for (uint32_t i = 0; ok && i < n; i++)
        for (uint32_t l = 0; ok && l <= level[i]; l++) {
            const uint64_t c = deg(i, l);
            ok = fwrite(&c, 8, 1, f) == 1 && fwrite(row(i, l), 4, c, f) == c;
        }
    const uint64_t ne = 1;
    ok = ok && fwrite(&ne, 8, 1, f) == 1 && fwrite(&entry, 4, 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : 2;
}

}  // namespace tools
}  // namespace pann
