// Per-tile depth order for global mode (same list as the reference's 64-bit (tile | depth)
// sort, gpu_radix_sort.cpp:132-195 with splat_tile_renderer.cpp:740-760).
//
// gfx950 design: the global radix sort only orders the list by its tile bits (2 passes for a
// 13-bit tile id instead of 6 for the full 45-bit key). Within a tile the entries are then in
// duplication (gaussian-index) order, and this kernel orders each tile's range by
// (depth bits, gaussian index), which is the stable (tile, depth) order of the full sort:
//   - one workgroup per tile; ranges of up to kChunk entries are sorted in LDS with a bitonic
//     network over 64-bit (depth << 32 | index) words (unique within a tile: a gaussian
//     appears at most once per tile), written back in place;
//   - longer ranges are sorted chunk by chunk the same way, then merged in rounds of doubling
//     run width through the scratch buffers (merge path: every thread finds its output slice
//     by a binary search, then merges it sequentially).
#include "gsr_internal.hpp"

namespace gsr {

namespace {

constexpr int kTsThreads = 256;
constexpr uint32_t kChunk = 2048;  // entries sorted in LDS at once (16 KB)

__device__ __forceinline__ uint64_t orderWord(uint64_t key, uint32_t value) {
    return (key << 32) | value;  // (depth bits, gaussian index); the tile bits shift out
}

__device__ __forceinline__ void bitonicSortLds(uint64_t* s, uint32_t n2) {
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2 / 2; i += kTsThreads) {
                const uint32_t a = ((i & ~(j - 1u)) << 1) | (i & (j - 1u));
                const uint32_t b = a + j;
                const bool up = (a & k) == 0u;
                const uint64_t x = s[a], y = s[b];
                if ((x > y) == up) {
                    s[a] = y;
                    s[b] = x;
                }
            }
            __syncthreads();
        }
    }
}

// Number of elements of run A (length la) among the first d outputs of merge(A, B).
__device__ __forceinline__ uint32_t coRank(const uint64_t* ka, const uint32_t* va, uint32_t la, const uint64_t* kb,
                                           const uint32_t* vb, uint32_t lb, uint32_t d) {
    uint32_t lo = d > lb ? d - lb : 0u, hi = d < la ? d : la;
    while (lo < hi) {
        const uint32_t i = (lo + hi) >> 1;
        if (orderWord(ka[i], va[i]) < orderWord(kb[d - i - 1], vb[d - i - 1])) lo = i + 1;
        else hi = i;
    }
    return lo;
}

__global__ __launch_bounds__(kTsThreads) void kTileDepthSort(uint64_t* __restrict__ keys, uint32_t* __restrict__ values,
                                                             uint64_t* __restrict__ keysTmp,
                                                             uint32_t* __restrict__ valuesTmp,
                                                             const uint2* __restrict__ ranges, uint32_t tileCount) {
    __shared__ uint64_t sW[kChunk];
    const uint32_t tile = blockIdx.x;
    if (tile >= tileCount) return;
    const uint2 range = ranges[tile];
    const uint32_t n = range.y - range.x;
    if (n <= 1u) return;  // uniform
    const uint64_t tileHi = uint64_t(tile) << 32;

    // 1. Chunks of up to kChunk entries, sorted in LDS, in place.
    for (uint32_t c0 = 0; c0 < n; c0 += kChunk) {
        const uint32_t m = n - c0 < kChunk ? n - c0 : kChunk;
        uint32_t n2 = 2;
        while (n2 < m) n2 <<= 1;
        const uint32_t base = range.x + c0;
        for (uint32_t i = threadIdx.x; i < n2; i += kTsThreads)
            sW[i] = i < m ? orderWord(keys[base + i], values[base + i]) : ~uint64_t(0);
        __syncthreads();
        bitonicSortLds(sW, n2);
        for (uint32_t i = threadIdx.x; i < m; i += kTsThreads) {
            const uint64_t w = sW[i];
            keys[base + i] = tileHi | (w >> 32);
            values[base + i] = uint32_t(w);
        }
        __syncthreads();
    }
    if (n <= kChunk) return;

    // 2. Merge rounds: runs of width w -> 2w, alternating between the list and the scratch.
    uint64_t* ks = keys + range.x;
    uint32_t* vs = values + range.x;
    uint64_t* kd = keysTmp + range.x;
    uint32_t* vd = valuesTmp + range.x;
    for (uint32_t w = kChunk; w < n; w <<= 1) {
        for (uint32_t lo = 0; lo < n; lo += 2 * w) {
            const uint32_t mid = lo + w < n ? lo + w : n;
            const uint32_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            const uint32_t la = mid - lo, lb = hi - mid, total = hi - lo;
            const uint32_t per = (total + kTsThreads - 1) / kTsThreads;
            const uint32_t d0 = threadIdx.x * per < total ? threadIdx.x * per : total;
            const uint32_t d1 = d0 + per < total ? d0 + per : total;
            uint32_t ia = coRank(ks + lo, vs + lo, la, ks + mid, vs + mid, lb, d0);
            uint32_t ib = d0 - ia;
            for (uint32_t d = d0; d < d1; ++d) {
                bool takeA;
                if (ia >= la) takeA = false;
                else if (ib >= lb) takeA = true;
                else takeA = orderWord(ks[lo + ia], vs[lo + ia]) < orderWord(ks[mid + ib], vs[mid + ib]);
                if (takeA) {
                    kd[lo + d] = ks[lo + ia];
                    vd[lo + d] = vs[lo + ia];
                    ++ia;
                } else {
                    kd[lo + d] = ks[mid + ib];
                    vd[lo + d] = vs[mid + ib];
                    ++ib;
                }
            }
        }
        __syncthreads();  // the round's output is the next round's input (same workgroup)
        uint64_t* tk = ks;
        ks = kd;
        kd = tk;
        uint32_t* tv = vs;
        vs = vd;
        vd = tv;
    }
    if (ks != keys + range.x) {  // the result is in the scratch: copy it back
        for (uint32_t i = threadIdx.x; i < n; i += kTsThreads) {
            keys[range.x + i] = ks[i];
            values[range.x + i] = vs[i];
        }
    }
}

}  // namespace

void launchTileDepthSort(uint64_t* keys, uint32_t* values, uint64_t* keysTmp, uint32_t* valuesTmp, const uint2* ranges,
                         uint32_t tileCount, hipStream_t stream) {
    if (tileCount == 0) return;
    hipLaunchKernelGGL(kTileDepthSort, dim3(tileCount), dim3(kTsThreads), 0, stream, keys, values, keysTmp, valuesTmp,
                       ranges, tileCount);
}

}  // namespace gsr
