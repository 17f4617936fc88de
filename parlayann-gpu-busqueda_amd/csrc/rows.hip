// rows.hip -- streaming kernels over 128-byte rows: squared norms, node-order gathers and
// a plain copy whose byte count is known (used to scale hardware traffic counters).
#include "pann_internal.h"

namespace pann {
namespace {

__global__ __launch_bounds__(256) void rowNormsKernel(int8_t* rows, uint32_t n, int32_t* out,
                                                     bool embed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int4* r = reinterpret_cast<int4*>(rows + (size_t)i * kVecBytes);
    int s = 0;
#pragma unroll
    for (int j = 0; j < kVecBytes / 16; j++) {
        int4 v = r[j];
        if (embed && j == kVecBytes / 16 - 1) v.w = 0;  // the norm's own bytes
        s = __builtin_amdgcn_sdot4(v.x, v.x, s, false);
        s = __builtin_amdgcn_sdot4(v.y, v.y, s, false);
        s = __builtin_amdgcn_sdot4(v.z, v.z, s, false);
        s = __builtin_amdgcn_sdot4(v.w, v.w, s, false);
    }
    out[i] = s;
    if (embed) reinterpret_cast<int32_t*>(rows + (size_t)i * kVecBytes)[kVecBytes / 4 - 1] = s;
}

// dst row i = src row idx[i]: eight lanes move one 128-byte row as 16-byte pieces, so a
// wavefront writes eight whole rows per step.
__global__ __launch_bounds__(256) void gatherRowsKernel(int4* __restrict__ dst,
                                                       const int4* __restrict__ src,
                                                       const uint32_t* __restrict__ idx,
                                                       uint32_t n, uint32_t nsrc) {
    constexpr int kPieces = kVecBytes / 16;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t p = t; p < (size_t)n * kPieces; p += stride) {
        const size_t row = p / kPieces;
        const uint32_t s = idx[row];
        if (s < nsrc) dst[p] = src[(size_t)s * kPieces + p % kPieces];
    }
}

// Random-gather ceiling: distance i reads row r(i) = hi32(hash64_2(seed + i)) * nrows >> 32
// and query row i mod nq (nq a power of two); eight lanes each take one 16-byte piece of both and one
// v_dot4_i32_i8 per dword, as the search kernels do.  Each lane keeps kInFlight rows in
// flight.  Only the gathered rows touch memory beyond the (cached) queries: indices come
// from the hash and the per-lane sums of q.p are folded into one 64-bit checksum.
constexpr int kGatherInFlight = 4;

__device__ __forceinline__ uint64_t mixSplit(uint64_t x) {
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void gatherBenchKernel(const int4* __restrict__ rows,
                                                        uint32_t nrows,
                                                        const int4* __restrict__ queries,
                                                        uint32_t nq, uint64_t count,
                                                        uint64_t seed,
                                                        unsigned long long* __restrict__ sum) {
    constexpr int kPieces = kVecBytes / 16;
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / kPieces;
    const int piece = threadIdx.x % kPieces;
    uint64_t acc = 0;
    for (uint64_t i0 = t / kPieces; i0 < count; i0 += groups * kGatherInFlight) {
        int4 v[kGatherInFlight], q[kGatherInFlight];
#pragma unroll
        for (int u = 0; u < kGatherInFlight; u++) {
            const uint64_t i = i0 + (uint64_t)u * groups;
            const bool on = i < count;
            const uint32_t r = (uint32_t)(((mixSplit(seed + i) >> 32) * (uint64_t)nrows) >> 32);
            v[u] = on ? rows[(size_t)r * kPieces + piece] : make_int4(0, 0, 0, 0);
            q[u] = queries[(size_t)(on ? ((uint32_t)i & (nq - 1)) : 0) * kPieces + piece];
        }
#pragma unroll
        for (int u = 0; u < kGatherInFlight; u++) {
            int d = __builtin_amdgcn_sdot4(v[u].x, q[u].x, 0, false);
            d = __builtin_amdgcn_sdot4(v[u].y, q[u].y, d, false);
            d = __builtin_amdgcn_sdot4(v[u].z, q[u].z, d, false);
            d = __builtin_amdgcn_sdot4(v[u].w, q[u].w, d, false);
            acc += (uint64_t)(int64_t)d;
        }
    }
    // wave sum, one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)acc, off, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(acc >> 32), off, 64);
        acc += ((uint64_t)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0) atomicAdd(sum, (unsigned long long)acc);
}

__global__ __launch_bounds__(256) void copyKernel(const int4* __restrict__ src,
                                                  int4* __restrict__ dst, size_t n16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

hipError_t launchCopy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    const size_t n16 = bytes / 16;
    if (n16 == 0) return hipSuccess;
    const size_t want = (n16 + 255) / 256, blocks = want < 256 * 64 ? want : 256 * 64;
    hipLaunchKernelGGL(copyKernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<const int4*>(src), reinterpret_cast<int4*>(dst), n16);
    return hipGetLastError();
}

hipError_t launchRowNorms(int8_t* rows, uint32_t n, int32_t* out, bool embed, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(rowNormsKernel, dim3((n + 255) / 256), dim3(256), 0, s, rows, n, out, embed);
    return hipGetLastError();
}

hipError_t launchGatherBench(const int8_t* rows, uint32_t nrows, const int8_t* queries,
                             uint32_t nq, uint64_t count, uint64_t seed,
                             unsigned long long* sum, int blocksPerCU, int cus, hipStream_t s) {
    if (count == 0 || nrows == 0 || nq == 0) return hipErrorInvalidValue;
    const uint64_t want = (count * (kVecBytes / 16) + 255) / 256;
    const uint64_t cap = (uint64_t)blocksPerCU * (uint64_t)cus;
    const unsigned blocks = (unsigned)(want < cap ? want : cap);
    hipLaunchKernelGGL(gatherBenchKernel, dim3(blocks), dim3(256), 0, s,
                       reinterpret_cast<const int4*>(rows), nrows,
                       reinterpret_cast<const int4*>(queries), nq, count, seed, sum);
    return hipGetLastError();
}

hipError_t launchGatherRows(int8_t* dst, const int8_t* src, const uint32_t* idx, uint32_t n,
                            uint32_t nsrc, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t pieces = (size_t)n * (kVecBytes / 16);
    const size_t want = (pieces + 255) / 256, blocks = want < 256 * 64 ? want : 256 * 64;
    hipLaunchKernelGGL(gatherRowsKernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       reinterpret_cast<int4*>(dst), reinterpret_cast<const int4*>(src), idx, n,
                       nsrc);
    return hipGetLastError();
}

}  // namespace pann
