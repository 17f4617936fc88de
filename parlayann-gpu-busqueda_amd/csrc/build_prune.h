// build_prune.h -- ParlayANN's robust prune on one wavefront, shared by the GPU graph builds.
#pragma once

#include "device_common.h"
#include "pann_internal.h"

namespace pann {
namespace {

using namespace dev;

constexpr uint32_t kSmallCap = 512;        // candidates of a prune in the common tier
constexpr uint32_t kSortCap = 2048;        // candidates of a prune in the deferred tier
constexpr int kPruneWaves = 4;             // wavefronts per common-tier prune block

__device__ __forceinline__ uint64_t keyOf(int d, uint32_t id) {
    return ((uint64_t)(uint32_t)d << 32) | id;
}

// ---- robust prune (one wavefront, LDS working set) ----

template <uint32_t CAP>
struct PruneLds {
    uint64_t keys[CAP];
    union {
        uint32_t list[CAP];                // reverse phase: the sources to add (read before
        uint16_t idx[CAP];                 // the prune lists live candidates here)
    };
    unsigned char flags[CAP];              // 1: pruned, or the center itself
    uint32_t out[128];
    uint32_t rowT[128];                    // the center's current row
};
static_assert(sizeof(PruneLds<kSmallCap>) * kPruneWaves <= 32 * 1024, "common-tier block LDS");
static_assert(sizeof(PruneLds<kSortCap>) <= 64 * 1024, "deferred-tier block LDS");

// Distance-keyed candidates keys[0..nc) of `center` (any order, duplicates allowed,
// nc <= CAP) -> at most m out-neighbours in out[0..return), nearest first: a candidate is
// kept unless an earlier kept s has alpha * d(s, c) <= d(center, c) (Vamana's robustPrune)
// or, STRICT, alpha * d(s, c) < d(center, c) (HNSW's neighbour heuristic).  After each kept
// s the candidates still live are listed, and their rows gathered 64 at a time (eight
// rows per 8-lane group in flight).
template <uint32_t CAP, bool STRICT>
__device__ uint32_t robustPruneWave(const int8_t* vec, int normsInRow, uint32_t m, uint32_t center,
                                    float alpha, PruneLds<CAP>& S, uint32_t nc, int lane) {
    const int sub = lane & 7, grp = lane >> 3;
    const int normMask = (normsInRow && sub == 7) ? -1 : 0;
    // bitonic sort of the keys in LDS, padded with ~0 to a power of two
    uint32_t P = 64;
    while (P < nc) P <<= 1;
    for (uint32_t t = nc + lane; t < P; t += 64) S.keys[t] = ~0ull;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = lane; t < P / 2; t += 64) {
                const uint32_t i0 = 2 * t - (t & (j - 1));
                const uint32_t i1 = i0 + j;
                const bool asc = (i0 & k) == 0;
                const uint64_t x = S.keys[i0], y = S.keys[i1];
                if ((x > y) == asc) {
                    S.keys[i0] = y;
                    S.keys[i1] = x;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    // drop repeats (the same id has the same distance, so repeats are adjacent and equal)
    uint32_t nu = 0;
    for (uint32_t b = 0; b < nc; b += 64) {
        const uint32_t t = b + lane;
        const uint64_t x = t < nc ? S.keys[t] : ~0ull;
        const uint64_t prev = (t < nc && t > 0) ? S.keys[t - 1] : ~0ull;
        const bool keep = t < nc && (t == 0 || x != prev);
        const uint64_t m = ballot(keep);
        __builtin_amdgcn_wave_barrier();
        if (keep) S.keys[nu + mbcnt(m)] = x;
        nu += (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
    }
    for (uint32_t t = lane; t < nu; t += 64) S.flags[t] = (uint32_t)S.keys[t] == center ? 1 : 0;
    __builtin_amdgcn_wave_barrier();

    uint32_t outc = 0, i = 0;
    while (outc < m) {
        int found = -1;
        for (uint32_t b = i; b < nu && found < 0; b += 64) {
            const uint32_t t = b + lane;
            const uint64_t m = ballot(t < nu && S.flags[t] == 0);
            if (m) found = (int)b + __ffsll((unsigned long long)m) - 1;
        }
        if (found < 0) break;
        const uint32_t s = (uint32_t)S.keys[found];
        if (lane == 0) S.out[outc] = s;
        outc++;
        if (outc >= m) break;
        int4 sv = *reinterpret_cast<const int4*>(vec + (size_t)s * kVecBytes + sub * 16);
        sv.w &= ~normMask;
        const int ss = groupSum8(sliceNorm(sv));
        uint32_t cnt = 0;
        for (uint32_t b = (uint32_t)found + 1; b < nu; b += 64) {
            const uint32_t t = b + lane;
            const bool live = t < nu && S.flags[t] == 0;
            const uint64_t m = ballot(live);
            if (live) S.idx[cnt + mbcnt(m)] = (uint16_t)t;
            cnt += (uint32_t)__popcll(m);
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t g0 = 0; g0 < cnt; g0 += 64) {
            const uint32_t nb = min(64u, cnt - g0);
            int4 v[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                v[r] = make_int4(0, 0, 0, 0);
                const uint32_t t = (uint32_t)r * 8 + grp;
                if (t < nb) {
                    const uint32_t id = (uint32_t)S.keys[S.idx[g0 + t]];
                    v[r] = *reinterpret_cast<const int4*>(vec + (size_t)id * kVecBytes + sub * 16);
                }
            }
            int part8[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                v[r].w &= ~normMask;
                part8[r] = slicePartial(v[r], sv);
            }
            const int tot = transposeReduce8(part8, lane);
            const int mine = __shfl(tot, (lane & 7) * 8 + (lane >> 3), 64);
            if ((uint32_t)lane < nb) {
                const uint32_t j = S.idx[g0 + lane];
                const int dsp = ss + mine;
                const int dj = (int)(S.keys[j] >> 32);
                const float lhs = alpha * (float)dsp;
                if (STRICT ? lhs < (float)dj : lhs <= (float)dj) S.flags[j] = 1;
            }
        }
        __builtin_amdgcn_wave_barrier();
        i = (uint32_t)found + 1;
    }
    __builtin_amdgcn_wave_barrier();
    return outc;
}

// keys[t] = (d(center, ids[t]), ids[t]) for t < count; ids in LDS.
__device__ void keysFromIds(const int8_t* vec, int normsInRow, uint32_t center, const uint32_t* ids,
                            uint32_t count, uint64_t* keys, int lane) {
    const int sub = lane & 7, grp = lane >> 3;
    const int normMask = (normsInRow && sub == 7) ? -1 : 0;
    int4 qv = *reinterpret_cast<const int4*>(vec + (size_t)center * kVecBytes + sub * 16);
    qv.w &= ~normMask;
    const int qq = groupSum8(sliceNorm(qv));
    for (uint32_t t0 = 0; t0 < count; t0 += 8) {
        const uint32_t t = t0 + grp;
        const uint32_t id = t < count ? ids[t] : 0;
        int4 v = make_int4(0, 0, 0, 0);
        if (t < count) v = *reinterpret_cast<const int4*>(vec + (size_t)id * kVecBytes + sub * 16);
        v.w &= ~normMask;
        const int d = qq + groupSum8(slicePartial(v, qv));
        if (t < count && sub == 0) keys[t] = keyOf(d, id);
    }
    __builtin_amdgcn_wave_barrier();
}

}  // namespace
}  // namespace pann
