// Per-tile depth order for global mode (same list as the reference's 64-bit (tile | depth)
// sort, gpu_radix_sort.cpp:132-195 with splat_tile_renderer.cpp:740-760).
//
// gfx950 design: the global radix sort only orders the list by its 32-bit tile ids (2 passes
// for a 13-bit tile id instead of 6 for the full 45-bit key; 8 B per entry moved instead of 12).
// Within a tile the entries are then in duplication (gaussian-index) order, and this kernel
// orders each tile's range by (depth bits, gaussian index), which is the stable
// (tile, depth, index) order of the full sort. The depth bits are looked up per gaussian
// (depthBits[g], written once per listed gaussian by the duplicate pass) when a range is loaded:
// whole rows of 64 list values first (the value buffers are padded past the list's end, so these
// loads need no per-lane bounds branch), then the lookups, then the padding selects.
//
// One wave per tile, no LDS and no workgroup barrier (the 4 waves of a workgroup work on 4
// different tiles and never wait for each other):
//   - a range of n <= 2048 entries is loaded into registers, E = 1, 2, 4, ..., 32 64-bit
//     (depth << 32 | index) words per lane (the smallest E with 64 E >= n; padding words sort
//     last), sorted by a bitonic network and written back. The words are unique within a tile
//     (a gaussian appears at most once per tile), so the order is total and deterministic.
//     Network positions are i = lane * E + r: partners less than E apart are in the lane's own
//     registers, partners further apart are in another lane: for E <= 16, fewer than 16 lanes
//     apart (17 of the 21 lane stages) a DPP move within the row per 32-bit half, VALU only;
//     16 or more lanes apart, and every lane stage of E = 32, a lane shuffle (ds_bpermute). The network is
//     the flip formulation (the first stage of every merge level pairs i with i ^ (k - 1),
//     every later stage pairs i with i ^ j, and the lower position always keeps the minimum),
//     so no per-element direction bits are needed. The stage loops stay rolled (only the
//     in-register stages unroll): the code stays small enough for the instruction cache when
//     waves of every size class run side by side;
//   - a longer range (a tile covered by more than 1024 splats: the dense central tiles) is
//     sorted afterwards by the whole workgroup: up to 8192 entries in registers, 4 waves of 8,
//     16 or 32 words per lane with the cross-wave stages through LDS; longer ones by LDS chunks
//     of 2048 (bitonic), then merge rounds of doubling run width through the scratch buffers
//     (merge path: every thread finds its output slice by a binary search, then merges it
//     sequentially). Workgroup b takes positions b, b + G, b + 2G, b + 3G of the tile order, so
//     the dense tiles, first in the order, are spread one per workgroup; every thread reads the
//     workgroup's four range lengths first, so the choice is uniform and a workgroup without a
//     long range never waits at a barrier. The register forms are separate (noinline)
//     functions: a loop around an inlined copy makes the compiler keep the words' addresses
//     live across iterations, doubling the kernel's registers.
// The loads are striped (word r of lane l is input entry r * 64 + l: coalesced; the input order
// of a sort is arbitrary), the stores follow the network positions. Tiles are taken in the
// order the caller gives (the frame passes the tiles centre-out: dense tiles, the long sorts,
// tend to be central and so start first rather than last; the lists do not depend on it).
#include "tilesort.hpp"

namespace gsr {

namespace {

// 5 workgroups per CU: 96 VGPRs, the most the 32-word sort fits in without spills.
__global__ __launch_bounds__(kTsThreads, 5) void kTileDepthSort(uint32_t* __restrict__ values,
                                                                const uint32_t* __restrict__ depthBits,
                                                                uint32_t* __restrict__ valuesTmp,
                                                                const uint2* __restrict__ ranges,
                                                                const uint32_t* __restrict__ order,
                                                                uint32_t tileCount) {
    __shared__ uint64_t sW[kLdsChunk];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t groupMask = 0;  // the workgroup's ranges longer than 2048 (same in every thread)
    uint32_t tiles[kTsWaves];
#pragma unroll
    for (uint32_t v = 0; v < uint32_t(kTsWaves); ++v) {
        const uint32_t at = blockIdx.x * kTsWaves + v;
        tiles[v] = at < tileCount ? order[at] : tileCount;
        if (tiles[v] < tileCount) {
            const uint2 rg = ranges[tiles[v]];
            if (rg.y - rg.x > kWaveAlone) groupMask |= 1u << v;
        }
    }
    // 1. One range of up to 2048 entries per wave (the tile, its range and size class are
    // uniform per wave: scalar registers).
    const uint32_t tile = __builtin_amdgcn_readfirstlane(tiles[wave]);
    if (tile < tileCount) {
        const uint2 range = ranges[tile];
        const uint32_t n = range.y - range.x;
        if (n > 1u && n <= kWaveAlone) waveSortAny(values, depthBits, range.x, n, lane);
    }
    if (groupMask == 0u) return;  // uniform per workgroup
    // 2. The longer ones, one at a time by the whole workgroup: up to 8192 entries in registers
    // (16 or 32 words per lane), longer ones by LDS chunks and merge rounds.
    __syncthreads();
    for (uint32_t v = 0; v < uint32_t(kTsWaves); ++v) {
        if (!(groupMask & (1u << v))) continue;
        const uint2 rg = ranges[tiles[v]];
        const uint32_t n = rg.y - rg.x;
        if (n <= 4096u) groupSortRegisters<16>(values, depthBits, rg, wave, lane, sW);
        else if (n <= kGroupChunk) groupSortRegisters<32>(values, depthBits, rg, wave, lane, sW);
        else groupSortLong(values, depthBits, valuesTmp, rg, sW);
        __syncthreads();  // sW is reused by the next range
    }
}

}  // namespace

void launchTileDepthSort(uint32_t* values, const uint32_t* depthBits, uint32_t* valuesTmp, const uint2* ranges,
                         const uint32_t* order, uint32_t tileCount, hipStream_t stream) {
    if (tileCount == 0) return;
    const uint32_t blocks = (tileCount + kTsWaves - 1) / kTsWaves;
    hipLaunchKernelGGL(kTileDepthSort, dim3(blocks), dim3(kTsThreads), 0, stream, values, depthBits, valuesTmp, ranges,
                       order, tileCount);
    markKernel("kTileDepthSort", stream);
}

}  // namespace gsr
