// K5 tile_duplicate, global 64-bit keys, per entry (reference shaders/tile_duplicate.comp:119-173,
// GSR_TILE_GLOBAL64 + GSR_DUPLICATE_PER_ENTRY) and K7 tile_ranges (shaders/tile_ranges.comp:22-43).
//
// Duplicate: one workgroup per 256 gaussians (the preprocess's blocks). It finishes the
// reduce-then-scan itself: the in-block exclusive scan of the tile counts plus the block's
// scanned offset gives every gaussian its first entry, so no per-gaussian offsets array is
// written or binary-searched in HBM. Each gaussian's coverage (rectangle, or the cutoff
// ellipse and its rows) is set up once into LDS; then the block's entries are handed out
// to all 256 threads (load-balanced over big splats) and each finds its gaussian by an
// 8-step binary search in LDS and its tile by walking that coverage's rows. Keys are
// (tile << 32 | depth bits), the depth bits read from the preprocess's depthBits[index], or the
// tile id alone (the list then moves 8 B per entry through the tile sort, and the per-tile depth
// sort looks the depth up in depthBits); values the gaussian index; entries >= capacity are
// dropped (index order), the tile id #tiles marks an entry without a tile (never produced).
#include "gsr_device.hpp"

namespace gsr {

namespace {

constexpr int kDupThreads = kPreprocessThreads;

constexpr int kCachedRows = kSpanRows;  // row spans kept in LDS per gaussian; longer splats walk the rest

// BINNING = false: block b = gaussians [256 b, 256 b + 256) in index order, 64-bit keys
// (tile << 32 | depth bits), or 32-bit keys = tile id with each listed gaussian's depth bits in
// depthBits[index]. BINNING = true: block b = the depth-sorted visible splats [256 b, ...) of
// `order` (count control->visible), tileCounts in that order, 32-bit keys = tile id.
__device__ __forceinline__ uint32_t waveSum64(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Global mode, workgroup 0 of the duplicate pass: the control block from the superblock sums (N,
// V, M, entries, cull counts, overflow; the counters the passes accumulate into were zeroed by the
// preprocess). Every thread of the workgroup calls it (barrier).
__device__ void writeControl(const DupFrame& frame, uint32_t count, gsr_control* control) {
    __shared__ uint32_t sSum[7];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    for (uint32_t f = wave; f < 7u; f += kDupThreads / 64) {  // field f of BlockStats, summed by a wave
        uint32_t acc = 0;
        for (uint32_t sb = lane; sb < frame.numSuper; sb += 64)
            acc += reinterpret_cast<const uint32_t*>(&frame.superStats[sb])[f];
        acc = waveSum64(acc);
        if (lane == 0) sSum[f] = acc;
    }
    __syncthreads();
    if (t == 0) {
        const uint32_t m = sSum[0];
        const uint32_t entries = m < frame.capacity ? m : frame.capacity;
        control->gaussian_count = count;
        control->visible = sSum[1];
        control->duplicates = m;
        control->entries = entries;
        for (int r = 0; r < 5; ++r) control->cull_counts[r] = sSum[2 + r];
        control->overflow = m - entries;
        control->ranges_groups[0] = (entries + 255u) / 256u;
        control->ranges_groups[1] = 1u;
        control->ranges_groups[2] = 1u;
    }
}

// BINNING = false (global mode): the block offsets come from frame (see DupFrame) and workgroup 0
// writes the control block; BINNING = true: from blockOffsets and control (kFrameSetup).
template <bool BINNING, typename Key>
__global__ __launch_bounds__(kDupThreads) void kDuplicate(const float4* __restrict__ splats,
                                                          const uint32_t* __restrict__ tileCounts,
                                                          const uint32_t* __restrict__ order,
                                                          const uint32_t* __restrict__ blockOffsets,
                                                          const gsr_control* __restrict__ control, DupFrame frame,
                                                          const uint32_t* __restrict__ spanInfo,
                                                          const uint4* __restrict__ spanPlanes,
                                                          uint32_t count, uint32_t bounds, BinGrid grid,
                                                          Key* __restrict__ keys, uint32_t* __restrict__ values,
                                                          const uint32_t* __restrict__ depthBits,
                                                          gsr_control* controlRW) {
    __shared__ uint32_t sScan[kDupThreads];
    __shared__ uint32_t sWave[kDupThreads / 64];
    __shared__ uint32_t sFrame[3];  // global mode: superblocks before this block, earlier blocks of it, M
    // Per gaussian: first row, row count, depth bits (64-bit keys), and either the rectangle's x0/width
    // (one span for every row) or the first kCachedRows row spans (x0 | width << 16). For the
    // rows past them (long splats) the binning form also keeps the ellipse: its blocks hold
    // depth-ordered splats, so the near, long ones come together and rebuilding their coverage
    // per entry costs more than the LDS. The global form (blocks in index order, long splats
    // spread out) rebuilds it from the splat, which keeps its LDS small enough for 8 waves/SIMD.
    __shared__ int sRow0[kDupThreads];
    __shared__ int sRows[kDupThreads];
    __shared__ uint32_t sDepth[kDupThreads];
    __shared__ uint32_t sSpan[kCachedRows][kDupThreads];
    __shared__ int sBinX[2][kDupThreads];
    __shared__ float sEl[BINNING ? 9 : 1][BINNING ? kDupThreads : 1];  // cx cy A B C D t exPad eyPad

    const uint32_t t = threadIdx.x;
    const int S = int(grid.binSize);
    const uint32_t j = blockIdx.x * kDupThreads + t;
    const uint32_t live = BINNING ? control->visible : count;
    if (BINNING && blockIdx.x * kDupThreads >= live) return;  // uniform: past the visible splats
    const uint32_t i = BINNING ? (j < live ? order[j] : 0u) : j;
    const uint32_t c = j < live ? tileCounts[BINNING ? j : i] : 0u;  // binning: counts in depth order
    __shared__ uint32_t sIndex[kDupThreads];
    sIndex[t] = i;  // read after the barrier below
    // Global mode: every gaussian's coverage record (and, for 64-bit keys, the preprocess's depth
    // bits) is loaded with its count, before the block offsets are known (a gaussian without tiles
    // holds a stale record, which is never used).
    uint32_t rInfo = 0u, rDepth = 0u;
    uint4 rPlane0 = make_uint4(0u, 0u, 0u, 0u);
    if (!BINNING && j < live) {
        rInfo = spanInfo[i];
        rPlane0 = spanPlanes[i];
        if (sizeof(Key) == 8) rDepth = depthBits[i];
    }

    // Block exclusive scan of the counts (the downsweep of the reduce-then-scan).
    const uint32_t lane = t & 63u, wave = t >> 6;
    if (!BINNING) {
        // The block's first entry: the tile sums of the superblocks before it (and of all: M) by
        // wave 0, of the earlier blocks of its own superblock by wave 1.
        const uint32_t sb = blockIdx.x / kSuperBlocks;
        if (wave == 0) {
            // Up to 256 superblocks in one round of loads (4 a lane), any further ones in a loop.
            uint32_t before = 0, all = 0, m[4];
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                const uint32_t q = lane + 64u * u;
                m[u] = q < frame.numSuper ? frame.superStats[q].tileSum : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                all += m[u];
                before += lane + 64u * u < sb ? m[u] : 0u;
            }
            for (uint32_t q = lane + 256u; q < frame.numSuper; q += 64) {
                const uint32_t x = frame.superStats[q].tileSum;
                all += x;
                before += q < sb ? x : 0u;
            }
            before = waveSum64(before);
            all = waveSum64(all);
            if (lane == 0) {
                sFrame[0] = before;
                sFrame[2] = all;
            }
        } else if (wave == 1) {
            const uint32_t b = sb * kSuperBlocks + lane;
            uint32_t m = b < blockIdx.x ? frame.blockStats[b].tileSum : 0u;
            m = waveSum64(m);
            if (lane == 0) sFrame[1] = m;
        }
    }
    uint32_t v = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t n = __shfl_up(v, off, 64);
        if (lane >= uint32_t(off)) v += n;
    }
    if (lane == 63) sWave[wave] = v;
    __syncthreads();
    uint32_t prefix = 0;
    for (uint32_t w = 0; w < wave; ++w) prefix += sWave[w];
    const uint32_t total = sWave[0] + sWave[1] + sWave[2] + sWave[3];
    const uint32_t excl = prefix + v - c;
    sScan[t] = excl;

    const uint32_t base = BINNING ? blockOffsets[blockIdx.x] : sFrame[0] + sFrame[1];
    const uint32_t entries = BINNING ? control->entries : (sFrame[2] < frame.capacity ? sFrame[2] : frame.capacity);
    // Uniform: the whole block is past the capacity or empty.
    const bool work = !(base >= entries || total == 0);
    if (BINNING && !work) return;

    // The coverage as the preprocess recorded it (first row, row code, the first kCachedRows
    // spans); only a binning-mode splat of more rows sets its ellipse up again (for the rows past
    // the cached ones, which its entries walk from LDS).
    bool recorded = false;
    if (work && c != 0u) {
        const uint32_t info = BINNING ? spanInfo[i] : rInfo;
        const int code = int(int16_t(uint16_t(info >> 16)));
        recorded = !BINNING || code <= kCachedRows;
        if (recorded) {
            sRow0[t] = int(info & 0xFFFFu);
            sRows[t] = code;
            if (!BINNING && sizeof(Key) == 8) sDepth[t] = rDepth;
            const uint4 p0 = BINNING ? spanPlanes[i] : rPlane0;
            if (code < 0) {  // a rectangle: x0 and width of every row
                sBinX[0][t] = int(p0.x & 0xFFFFu);
                sBinX[1][t] = int(p0.x >> 16);
            } else {
                sSpan[0][t] = p0.x;
                sSpan[1][t] = p0.y;
                sSpan[2][t] = p0.z;
                sSpan[3][t] = p0.w;
                if (code > 4) {
                    const uint4 p1 = spanPlanes[size_t(count) + i];
                    sSpan[4][t] = p1.x;
                    sSpan[5][t] = p1.y;
                    sSpan[6][t] = p1.z;
                    sSpan[7][t] = p1.w;
                }
            }
        }
    }
    if (work && c != 0u && !recorded) {
        const float4* s = splats + 3 * size_t(i);
        const float4 o0 = s[0], o1 = s[1], o2 = s[2];
        const Rect rect = tileRect(o0.x, o0.y, o2.w, grid.tilesX, grid.tilesY);
        const Coverage cv = makeCoverage(bounds, rect, S, int(grid.binsX), int(grid.binsY), o0.x, o0.y, o1.x, o1.y,
                                         o1.z, o1.w);
        int r0, r1;
        coverageRows(cv, S, &r0, &r1);
        sRow0[t] = r0;
        sRows[t] = r1 - r0;
        if (!cv.useEllipse) {
            sBinX[0][t] = cv.bins.x0;
            sBinX[1][t] = cv.bins.x1 - cv.bins.x0;
            sRows[t] = -(r1 - r0) - 1;  // negative: a rectangle
        } else {
            if constexpr (BINNING) {
                sEl[0][t] = cv.e.cx;
                sEl[1][t] = cv.e.cy;
                sEl[2][t] = cv.e.A;
                sEl[3][t] = cv.e.B;
                sEl[4][t] = cv.e.C;
                sEl[5][t] = cv.e.D;
                sEl[6][t] = cv.e.t;
                sEl[7][t] = cv.e.exPad;
                sEl[8][t] = cv.e.eyPad;
                sBinX[0][t] = cv.bins.x0;
                sBinX[1][t] = cv.bins.x1;
            }
            // Walk the rows once: cache the first spans, and check the count/write
            // agreement (tile_duplicate.comp:375-376) from the same walk.
            uint32_t sum = 0;
            for (int row = r0; row < r1; ++row) {
                int x0, x1;
                coverageRowSpan(cv, row, S, &x0, &x1);
                if (row - r0 < kCachedRows) sSpan[row - r0][t] = uint32_t(x0) | (uint32_t(x1 - x0) << 16);
                sum += uint32_t(x1 - x0);
            }
            if (sum != c && base + excl < entries) atomicAdd(&controlRW->coverage_mismatch, 1u);
        }
    }
    __syncthreads();

    const uint32_t end = (base + total < entries) ? base + total : entries;
    const uint32_t binsX = grid.binsX;
    const uint32_t gridTiles = binsX * grid.binsY;
    for (uint32_t k = base + t; k < end; k += kDupThreads) {
        const uint32_t rl = k - base;
        uint32_t lo = 0, hi = kDupThreads;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sScan[mid] <= rl) lo = mid;
            else hi = mid;
        }
        const uint32_t r = rl - sScan[lo];
        uint32_t tile = gridTiles;
        const int rows = sRows[lo];
        const int r0 = sRow0[lo];
        if (rows < 0) {  // rectangle
            const uint32_t nr = uint32_t(-rows - 1);
            const uint32_t w = uint32_t(sBinX[1][lo]);
            if (w > 0 && r < w * nr) tile = (uint32_t(r0) + r / w) * binsX + uint32_t(sBinX[0][lo]) + r % w;
        } else {
            uint32_t before = 0;
            const int cached = rows < kCachedRows ? rows : kCachedRows;
            int row = 0;
            for (; row < cached; ++row) {
                const uint32_t sp = sSpan[row][lo];
                const uint32_t w = sp >> 16;
                if (r < before + w) {
                    tile = uint32_t(r0 + row) * binsX + (sp & 0xFFFFu) + (r - before);
                    break;
                }
                before += w;
            }
            if (tile == gridTiles && row == kCachedRows) {  // a long splat: walk the remaining rows
                // of its coverage (binning: the cached ellipse; global: rebuilt from the splat as
                // above, the same inputs and so the same bits)
                Coverage cv;
                if constexpr (BINNING) {
                    cv.bins = Rect{sBinX[0][lo], 0, sBinX[1][lo], 0};
                    cv.useEllipse = true;
                    cv.e.cx = sEl[0][lo];
                    cv.e.cy = sEl[1][lo];
                    cv.e.A = sEl[2][lo];
                    cv.e.B = sEl[3][lo];
                    cv.e.C = sEl[4][lo];
                    cv.e.D = sEl[5][lo];
                    cv.e.t = sEl[6][lo];
                    cv.e.exPad = sEl[7][lo];
                    cv.e.eyPad = sEl[8][lo];
                    cv.e.valid = true;
                } else {
                    const float4* s = splats + 3 * size_t(sIndex[lo]);
                    const float4 o0 = s[0], o1 = s[1], o2 = s[2];
                    const Rect rect = tileRect(o0.x, o0.y, o2.w, grid.tilesX, grid.tilesY);
                    cv = makeCoverage(bounds, rect, S, int(grid.binsX), int(grid.binsY), o0.x, o0.y, o1.x, o1.y, o1.z,
                                      o1.w);
                }
                for (int rr = r0 + kCachedRows; rr < r0 + rows; ++rr) {
                    int x0, x1;
                    coverageRowSpan(cv, rr, S, &x0, &x1);
                    const uint32_t w = uint32_t(x1 - x0);
                    if (r < before + w) {
                        tile = uint32_t(rr) * binsX + uint32_t(x0) + (r - before);
                        break;
                    }
                    before += w;
                }
            }
        }
        if (BINNING || sizeof(Key) == 4) keys[k] = Key(tile);
        else keys[k] = Key((uint64_t(tile) << 32) | uint64_t(sDepth[lo]));
        values[k] = sIndex[lo];
    }
    if (!BINNING && blockIdx.x == 0) writeControl(frame, count, controlRW);
}

constexpr int kRangesThreads = 256;

template <typename Key>
__device__ __forceinline__ uint32_t tileOfKey(Key k) {
    return sizeof(Key) == 8 ? uint32_t(uint64_t(k) >> 32) : uint32_t(k);
}

// Also zeroes zeroWords[0, zeroCount) (workgroup 0): global mode's superblock sums, which the
// frame's duplicate pass has read, for the next frame's preprocess.
template <typename Key>
__global__ __launch_bounds__(kRangesThreads) void kRanges(const Key* __restrict__ keys,
                                                          const gsr_control* __restrict__ control, uint32_t tileCount,
                                                          uint2* __restrict__ ranges, uint32_t* __restrict__ zeroWords,
                                                          uint32_t zeroCount) {
    if (blockIdx.x == 0)
        for (uint32_t w = threadIdx.x; w < zeroCount; w += kRangesThreads) zeroWords[w] = 0u;
    const uint32_t n = control->entries;
    for (uint32_t k = blockIdx.x * kRangesThreads + threadIdx.x; k < n; k += gridDim.x * kRangesThreads) {
        const uint32_t tile = tileOfKey(keys[k]);
        if (tile >= tileCount) {
            if (k > 0) {
                const uint32_t prev = tileOfKey(keys[k - 1]);
                if (prev < tileCount) ranges[prev].y = k;
            }
            continue;
        }
        if (k == 0) {
            ranges[tile].x = 0u;
        } else {
            const uint32_t prev = tileOfKey(keys[k - 1]);
            if (prev != tile) {
                if (prev < tileCount) ranges[prev].y = k;
                ranges[tile].x = k;
            }
        }
        if (k == n - 1) ranges[tile].y = n;
    }
}

}  // namespace

// Global mode: at least one workgroup, so that the control block is written for an empty scene.
void launchDuplicate(const float4* splats, const uint32_t* tileCounts, const DupFrame& frame, const uint32_t* spanInfo,
                     const uint4* spanPlanes, const uint32_t* depthBits, uint32_t count, uint32_t bounds,
                     const BinGrid& grid, uint64_t* keys, uint32_t* values, gsr_control* control, hipStream_t stream) {
    const uint32_t blocks = count > 0 ? (count + kDupThreads - 1) / kDupThreads : 1u;
    hipLaunchKernelGGL((kDuplicate<false, uint64_t>), dim3(blocks), dim3(kDupThreads), 0, stream, splats, tileCounts,
                       nullptr, nullptr, nullptr, frame, spanInfo, spanPlanes, count, bounds, grid, keys, values,
                       depthBits, control);
    markKernel("kDuplicate", stream);
}

void launchDuplicateTiles(const float4* splats, const uint32_t* tileCounts, const DupFrame& frame,
                          const uint32_t* spanInfo, const uint4* spanPlanes, uint32_t count, uint32_t bounds,
                          const BinGrid& grid, uint32_t* keys, uint32_t* values, gsr_control* control,
                          hipStream_t stream) {
    const uint32_t blocks = count > 0 ? (count + kDupThreads - 1) / kDupThreads : 1u;
    hipLaunchKernelGGL((kDuplicate<false, uint32_t>), dim3(blocks), dim3(kDupThreads), 0, stream, splats, tileCounts,
                       nullptr, nullptr, nullptr, frame, spanInfo, spanPlanes, count, bounds, grid, keys, values,
                       nullptr, control);
    markKernel("kDuplicate", stream);
}

void launchDuplicateBinning(const float4* splats, const uint32_t* sortedCounts, const uint32_t* order,
                            const uint32_t* blockOffsets, const gsr_control* control, const uint32_t* spanInfo,
                            const uint4* spanPlanes, uint32_t maxVisible, uint32_t bounds, const BinGrid& grid,
                            uint32_t* keys, uint32_t* values, gsr_control* controlRW, hipStream_t stream) {
    const uint32_t blocks = (maxVisible + kDupThreads - 1) / kDupThreads;
    if (blocks == 0) return;
    hipLaunchKernelGGL((kDuplicate<true, uint32_t>), dim3(blocks), dim3(kDupThreads), 0, stream, splats, sortedCounts,
                       order, blockOffsets, control, DupFrame{}, spanInfo, spanPlanes, maxVisible, bounds, grid, keys,
                       values, nullptr, controlRW);
    markKernel("kDuplicate", stream);
}

template <typename Key>
void launchRangesT(const Key* keys, const gsr_control* control, uint32_t capacity, uint32_t tileCount, uint2* ranges,
                   uint32_t* zeroWords, uint32_t zeroCount, hipStream_t stream) {
    uint32_t blocks = (capacity + kRangesThreads - 1) / kRangesThreads;
    if (blocks > 2048) blocks = 2048;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(kRanges<Key>, dim3(blocks), dim3(kRangesThreads), 0, stream, keys, control, tileCount, ranges,
                       zeroWords, zeroCount);
    markKernel("kRanges", stream);
}

void launchRanges(const uint64_t* keys, const gsr_control* control, uint32_t capacity, uint32_t tileCount,
                  uint2* ranges, uint32_t* zeroWords, uint32_t zeroCount, hipStream_t stream) {
    launchRangesT(keys, control, capacity, tileCount, ranges, zeroWords, zeroCount, stream);
}

void launchRanges32(const uint32_t* keys, const gsr_control* control, uint32_t capacity, uint32_t tileCount,
                    uint2* ranges, uint32_t* zeroWords, uint32_t zeroCount, hipStream_t stream) {
    launchRangesT(keys, control, capacity, tileCount, ranges, zeroWords, zeroCount, stream);
}

}  // namespace gsr
