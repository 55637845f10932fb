// Binning-mode front half (the reference's default TileSortMode::Binning,
// splat_tile_renderer.h:8-26 and 246-266): the visible splats are compacted in index order,
// sorted by depth (32-bit keys, stable, so ties keep index order), and their tile counts are
// gathered and scanned in that depth order. The duplicate pass then writes each splat's tiles
// in depth order with 32-bit tile keys, and one stable sort by tile yields the same
// (tile, depth, index) order as global mode.
//
// gfx950 layout: the compaction reuses the preprocess block statistics (kFrameSetup scans the
// per-block visible counts), so no separate flag scan runs; the gather writes the counts in
// depth order so the duplicate reads them coalesced.
#include "gsr_internal.hpp"

namespace gsr {

namespace {

constexpr int kBinThreads = 256;

// One thread per gaussian: visible ones (tile count != 0) go to their rank among the
// visible, in index order: blockOffsets[b] + (visible before them in the block).
__global__ __launch_bounds__(kBinThreads) void kCompactVisible(const float4* __restrict__ splats,
                                                               const uint32_t* __restrict__ tileCounts,
                                                               const uint32_t* __restrict__ blockOffsets,
                                                               uint32_t count, uint32_t* __restrict__ depthKeys,
                                                               uint32_t* __restrict__ depthValues) {
    __shared__ uint32_t sWave[kBinThreads / 64];
    const uint32_t i = blockIdx.x * kBinThreads + threadIdx.x;
    const bool vis = i < count && tileCounts[i] != 0u;
    const uint64_t mask = __ballot(vis);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) sWave[wave] = uint32_t(__popcll(mask));
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += sWave[w];
    before += uint32_t(__popcll(mask & ((uint64_t(1) << lane) - 1u)));
    if (vis) {
        const uint32_t slot = blockOffsets[blockIdx.x] + before;
        depthKeys[slot] = __float_as_uint(splats[3 * size_t(i)].z);  // depth > 0.2: bits are monotonic
        depthValues[slot] = i;
    }
}

// Per 256 depth-sorted visible splats: counts in depth order and their block sum.
__global__ __launch_bounds__(kBinThreads) void kGatherCounts(const uint32_t* __restrict__ tileCounts,
                                                             const uint32_t* __restrict__ order,
                                                             const gsr_control* __restrict__ control,
                                                             uint32_t* __restrict__ sortedCounts,
                                                             uint32_t* __restrict__ blockSums) {
    const uint32_t visible = control->visible;
    if (blockIdx.x * kBinThreads >= visible) return;  // uniform
    const uint32_t j = blockIdx.x * kBinThreads + threadIdx.x;
    uint32_t c = 0;
    if (j < visible) {
        c = tileCounts[order[j]];
        sortedCounts[j] = c;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    __shared__ uint32_t sWave[kBinThreads / 64];
    if ((threadIdx.x & 63u) == 0) sWave[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) blockSums[blockIdx.x] = sWave[0] + sWave[1] + sWave[2] + sWave[3];
}

}  // namespace

void launchCompactVisible(const float4* splats, const uint32_t* tileCounts, const uint32_t* blockOffsets,
                          uint32_t count, uint32_t* depthKeys, uint32_t* depthValues, hipStream_t stream) {
    const uint32_t blocks = (count + kBinThreads - 1) / kBinThreads;
    if (blocks == 0) return;
    static_assert(kBinThreads == kPreprocessThreads, "compaction blocks are the preprocess blocks");
    hipLaunchKernelGGL(kCompactVisible, dim3(blocks), dim3(kBinThreads), 0, stream, splats, tileCounts, blockOffsets,
                       count, depthKeys, depthValues);
    markKernel("kCompactVisible", stream);
}

void launchGatherCounts(const uint32_t* tileCounts, const uint32_t* order, const gsr_control* control,
                        uint32_t maxVisible, uint32_t* sortedCounts, uint32_t* blockSums, hipStream_t stream) {
    const uint32_t blocks = (maxVisible + kBinThreads - 1) / kBinThreads;
    if (blocks == 0) return;
    hipLaunchKernelGGL(kGatherCounts, dim3(blocks), dim3(kBinThreads), 0, stream, tileCounts, order, control,
                       sortedCounts, blockSums);
    markKernel("kGatherCounts", stream);
}

}  // namespace gsr
