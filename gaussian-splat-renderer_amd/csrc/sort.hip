// K6 radix sort (reference GpuRadixSort, render/gpu_radix_sort.{h,cpp}; shaders
// radix_histogram.comp, radix_scan.comp, radix_scatter_reorder.comp): stable LSD sort of
// (64-bit or 32-bit key, u32 value) pairs by digits of at most 8 bits.
//
// gfx950 design. The key count lives on the device (the frame's entries = min(M, capacity)):
// every kernel reads it there; grids are sized for the capacity and idle blocks exit.
//
// Reduce-then-scan (the default, GSR_SORT_REDUCE_THEN_SCAN): per digit of at most 8 bits (a
// key's bits split evenly over the passes) three kernels:
//   - kSortHistogram: each block counts the digits of its contiguous chunk of keys in LDS and
//     writes one count per (digit, block) for the 2^bits digits of the pass;
//   - kSortScan: one workgroup per digit scans that digit's block counts;
//   - kSortScatter: each wave64 ranks its keys with one 64-lane ballot per digit bit (the
//     lanes holding the same digit) and a running per-wave digit counter in LDS, so ranks
//     follow the input order (stable); the block's tile of 2048 keys is reordered by digit in
//     LDS and written out as contiguous per-digit runs (coalesced stores).
//
// One-sweep (GSR_SORT_ONESWEEP, 64-bit keys only): one up-front pass builds the global
// histograms of every digit position; then ONE kernel per digit over tiles of 2048 keys ranks
// as above and takes the tile's per-digit offset among the earlier tiles from a decoupled
// look-back over per-(tile, digit) status words (aggregate / inclusive prefix, tagged with the
// pass's epoch, so the status array never needs clearing). Tiles take their index from an
// atomic ticket, so a tile only waits on tiles that are already running; the wait is bounded
// and a timeout is counted, never hung on. Slower here than reduce-then-scan: the look-back
// crosses the XCDs' separate L2s.
#include "gsr_internal.hpp"

namespace gsr {

namespace {

constexpr uint64_t kFlagAggregate = 1;
constexpr uint64_t kFlagPrefix = 2;
constexpr uint32_t kSpinLimit = 1u << 20;

__device__ __forceinline__ uint32_t sortCount(const uint32_t* countDevice, uint32_t maxCount) {
    if (!countDevice) return maxCount;
    const uint32_t c = *countDevice;
    return c < maxCount ? c : maxCount;
}

// The reduce-then-scan grid of G blocks: block b sorts keys [b chunk, b chunk + chunk), one tile
// of kSortTile keys while the count allows (chunk > kSortTile only past G tiles); `used` blocks
// hold keys. Every kernel of a pass derives the same values from the count on the device, so the
// grid can be sized for the capacity while the work follows the count.
struct SortChunks {
    uint32_t chunk, used;
};
__device__ __forceinline__ SortChunks sortChunks(uint32_t count, uint32_t G) {
    uint32_t chunk = (count + G - 1) / G;
    chunk = chunk < uint32_t(kSortTile) ? uint32_t(kSortTile) : chunk;
    return SortChunks{chunk, (count + chunk - 1) / chunk};
}

template <typename Key>
__device__ __forceinline__ uint32_t digitOf(Key key, uint32_t shift) { return uint32_t(key >> shift) & 0xFFu; }
// A digit of `bits` <= 8 bits (the reduce-then-scan passes split a key's bits evenly).
template <typename Key>
__device__ __forceinline__ uint32_t digitOf(Key key, uint32_t shift, uint32_t bits) {
    return uint32_t(key >> shift) & ((1u << bits) - 1u);
}

__device__ __forceinline__ uint64_t packStatus(uint32_t epoch, uint64_t flag, uint32_t value) {
    return (uint64_t(epoch) << 34) | (flag << 32) | uint64_t(value);
}

// Exclusive scan of v over the 256 threads of a workgroup (total optional).
__device__ __forceinline__ uint32_t blockExclusiveScan256(uint32_t v, uint32_t* sWave, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t n = __shfl_up(inc, off, 64);
        if (lane >= uint32_t(off)) inc += n;
    }
    if (lane == 63) sWave[wave] = inc;
    __syncthreads();
    uint32_t prefix = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w)
        if (uint32_t(w) < wave) prefix += sWave[w];
    if (total) *total = sWave[0] + sWave[1] + sWave[2] + sWave[3];
    __syncthreads();
    return prefix + inc - v;
}

// Zeroes the global histograms and the tickets of this sort.
__global__ __launch_bounds__(kSortThreads) void kSortInit(uint32_t* __restrict__ globalHist,
                                                          uint32_t* __restrict__ tickets, uint32_t passes) {
    for (uint32_t i = threadIdx.x; i < passes * kRadix; i += kSortThreads) globalHist[i] = 0u;
    if (threadIdx.x < kMaxSortPasses) tickets[threadIdx.x] = 0u;
}

// Global digit histograms of every pass, in one read of the keys.
__global__ __launch_bounds__(kSortThreads) void kSortUpfront(const uint64_t* __restrict__ keys,
                                                             const uint32_t* __restrict__ countDevice,
                                                             uint32_t maxCount, uint32_t firstBit, uint32_t passes,
                                                             uint32_t* __restrict__ globalHist) {
    __shared__ uint32_t sHist[kMaxSortPasses][kRadix];
    for (uint32_t i = threadIdx.x; i < kMaxSortPasses * kRadix; i += kSortThreads) (&sHist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t count = sortCount(countDevice, maxCount);
    for (uint32_t idx = blockIdx.x * kSortThreads + threadIdx.x; idx < count; idx += gridDim.x * kSortThreads) {
        const uint64_t k = keys[idx];
        for (uint32_t p = 0; p < passes; ++p) atomicAdd(&sHist[p][digitOf(k, firstBit + p * kRadixBits)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < passes * kRadix; i += kSortThreads) {
        const uint32_t c = (&sHist[0][0])[i];
        if (c) atomicAdd(&globalHist[i], c);
    }
}

__global__ __launch_bounds__(kSortThreads) void kSortOnesweep(const uint64_t* __restrict__ keysIn,
                                                              const uint32_t* __restrict__ valsIn,
                                                              uint64_t* __restrict__ keysOut,
                                                              uint32_t* __restrict__ valsOut,
                                                              const uint32_t* __restrict__ countDevice,
                                                              uint32_t maxCount, uint32_t shift,
                                                              const uint32_t* __restrict__ passHist,
                                                              uint64_t* __restrict__ status, uint32_t* ticket,
                                                              uint32_t epoch, uint32_t* timeouts) {
    __shared__ uint64_t sKeys[kSortTile];
    __shared__ uint32_t sVals[kSortTile];
    __shared__ uint32_t sWaveHist[kSortWaves][kRadix];
    __shared__ uint32_t sGlobalBase[kRadix];
    __shared__ uint32_t sLocalStart[kRadix];
    __shared__ uint32_t sWave[kSortWaves];
    __shared__ uint32_t sTile;

    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    if (t == 0) sTile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tileIdx = sTile;
    const uint32_t count = sortCount(countDevice, maxCount);
    const uint32_t tileBegin = tileIdx * uint32_t(kSortTile);
    if (tileBegin >= count) return;  // uniform: past the keys
    const uint32_t tileN = count - tileBegin < uint32_t(kSortTile) ? count - tileBegin : uint32_t(kSortTile);

    for (uint32_t d = lane; d < kRadix; d += 64) sWaveHist[wave][d] = 0u;
    const uint64_t laneMaskLt = (uint64_t(1) << lane) - 1u;
    uint64_t key[kSortKeysPerLane];
    uint32_t val[kSortKeysPerLane], dig[kSortKeysPerLane], rank[kSortKeysPerLane];
#pragma unroll
    for (int r = 0; r < kSortKeysPerLane; ++r) {
        const uint32_t pos = wave * (kSortTile / kSortWaves) + r * 64 + lane;
        const bool valid = pos < tileN;
        key[r] = valid ? keysIn[tileBegin + pos] : 0ull;
        val[r] = valid ? valsIn[tileBegin + pos] : 0u;
    }
#pragma unroll
    for (int r = 0; r < kSortKeysPerLane; ++r) {
        const uint32_t pos = wave * (kSortTile / kSortWaves) + r * 64 + lane;
        const bool valid = pos < tileN;
        const uint32_t d = digitOf(key[r], shift);
        dig[r] = d;
        // Lanes holding the same digit: 8 ballots, one per digit bit.
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        if (valid) {
            const uint64_t lower = peers & laneMaskLt;
            const uint32_t old = sWaveHist[wave][d];
            rank[r] = old + uint32_t(__popcll(lower));
            if (lower == 0u) sWaveHist[wave][d] = old + uint32_t(__popcll(peers));
        }
    }
    __syncthreads();

    // Thread d: the tile's count of digit d, its offset per wave, and the decoupled look-back.
    const uint32_t d = t;  // kSortThreads == kRadix
    uint32_t tileCount = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t c = sWaveHist[w][d];
        sWaveHist[w][d] = tileCount;
        tileCount += c;
    }
    uint64_t* myStatus = status + size_t(tileIdx) * kRadix + d;
    if (tileIdx == 0) {
        __hip_atomic_store(myStatus, packStatus(epoch, kFlagPrefix, tileCount), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(myStatus, packStatus(epoch, kFlagAggregate, tileCount), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    sLocalStart[d] = blockExclusiveScan256(tileCount, sWave, nullptr);
    const uint32_t digitStart = blockExclusiveScan256(passHist[d], sWave, nullptr);
    uint32_t before = 0;
    if (tileIdx > 0) {
        int32_t j = int32_t(tileIdx) - 1;
        uint32_t spins = 0;
        while (j >= 0) {
            const uint64_t s = __hip_atomic_load(status + size_t(j) * kRadix + d, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t se = uint32_t(s >> 34);
            const uint64_t flag = (s >> 32) & 3u;
            if (se != epoch || flag == 0) {
                if (++spins > kSpinLimit) {  // never expected: count it and give up rather than hang
                    atomicAdd(timeouts, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            before += uint32_t(s);
            if (flag == kFlagPrefix) break;
            --j;
        }
        __hip_atomic_store(myStatus, packStatus(epoch, kFlagPrefix, before + tileCount), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    sGlobalBase[d] = digitStart + before;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kSortKeysPerLane; ++r) {
        const uint32_t pos = wave * (kSortTile / kSortWaves) + r * 64 + lane;
        if (pos < tileN) {
            const uint32_t lr = sLocalStart[dig[r]] + sWaveHist[wave][dig[r]] + rank[r];
            sKeys[lr] = key[r];
            sVals[lr] = val[r];
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < tileN; i += kSortThreads) {
        const uint64_t k = sKeys[i];
        const uint32_t dd = digitOf(k, shift);
        const uint32_t dst = sGlobalBase[dd] + (i - sLocalStart[dd]);
        keysOut[dst] = k;
        valsOut[dst] = sVals[i];
    }
}


// --- Reduce-then-scan passes: histogram, per-digit scan, scatter (3 kernels per digit) ---

template <typename Key>
__global__ __launch_bounds__(kSortThreads) void kSortHistogram(const Key* __restrict__ keys,
                                                               const uint32_t* __restrict__ countDevice,
                                                               uint32_t maxCount, uint32_t shift, uint32_t bits,
                                                               uint32_t* __restrict__ hist) {
    __shared__ uint32_t sHist[kSortWaves][kRadix];
    const uint32_t t = threadIdx.x, wave = t >> 6;
    for (uint32_t d = t; d < kSortWaves * kRadix; d += kSortThreads) (&sHist[0][0])[d] = 0u;
    __syncthreads();
    const uint32_t count = sortCount(countDevice, maxCount);
    const SortChunks sc = sortChunks(count, gridDim.x);
    if (blockIdx.x >= sc.used) return;  // uniform: no keys (the scan reads only the used rows)
    const uint32_t begin = blockIdx.x * sc.chunk;
    const uint32_t end = begin + sc.chunk < count ? begin + sc.chunk : count;
    // four keys in flight per thread, then their counts
    for (uint32_t idx = begin + t; idx < end; idx += 4 * kSortThreads) {
        Key k[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t e = idx + uint32_t(u) * kSortThreads;
            k[u] = e < end ? keys[e] : Key(0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (idx + uint32_t(u) * kSortThreads < end) atomicAdd(&sHist[wave][digitOf(k[u], shift, bits)], 1u);
    }
    __syncthreads();
    const uint32_t R = 1u << bits;
    for (uint32_t d = t; d < R; d += kSortThreads) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) s += sHist[w][d];
        hist[d * gridDim.x + blockIdx.x] = s;
    }
}

// One workgroup per digit d < 2^bits: exclusive scan of the used blocks' counts of digit d
// (hist[d * G + b], b < used), in place; the digit's total into digitTotals[d].
__global__ __launch_bounds__(kSortThreads) void kSortScan(uint32_t* __restrict__ hist,
                                                          const uint32_t* __restrict__ countDevice,
                                                          uint32_t maxCount, uint32_t G, uint32_t bits,
                                                          uint32_t* __restrict__ digitTotals) {
    __shared__ uint32_t sWave[kSortWaves];
    const uint32_t d = blockIdx.x;
    const uint32_t used = sortChunks(sortCount(countDevice, maxCount), G).used;
    uint32_t* row = hist + size_t(d) * G;
    constexpr int kPer = int(kMaxSortBlocks) / kSortThreads;  // up to 16 block counts per thread
    const uint32_t first = threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {  // all loads first (independent), then the sum
        const uint32_t b = first + uint32_t(j);
        v[j] = b < used ? row[b] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) s += v[j];
    uint32_t total;
    uint32_t run = blockExclusiveScan256(s, sWave, &total);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t b = first + uint32_t(j);
        if (b < used) row[b] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) digitTotals[d] = total;
}

template <typename Key>
__global__ __launch_bounds__(kSortThreads) void kSortScatter(const Key* __restrict__ keysIn,
                                                             const uint32_t* __restrict__ valsIn,
                                                             Key* __restrict__ keysOut,
                                                             uint32_t* __restrict__ valsOut,
                                                             const uint32_t* __restrict__ countDevice,
                                                             uint32_t maxCount, uint32_t shift, uint32_t bits,
                                                             const uint32_t* __restrict__ hist,
                                                             const uint32_t* __restrict__ digitTotals) {
    __shared__ Key sKeys[kSortTile];
    __shared__ uint32_t sVals[kSortTile];
    __shared__ uint32_t sWaveHist[kSortWaves][kRadix];
    __shared__ uint32_t sGlobalBase[kRadix];
    __shared__ uint32_t sLocalStart[kRadix];
    __shared__ uint32_t sTileCount[kRadix];
    __shared__ uint32_t sWave[kSortWaves];

    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t count = sortCount(countDevice, maxCount);
    const SortChunks sc = sortChunks(count, gridDim.x);
    if (blockIdx.x >= sc.used) return;  // uniform: no keys
    const uint32_t begin = blockIdx.x * sc.chunk;
    const uint32_t end = begin + sc.chunk < count ? begin + sc.chunk : count;

    {  // global base of (digit, block) = exclusive prefix of the digit totals + the in-digit scan
        const uint32_t d = t;  // kSortThreads == kRadix; digits >= 2^bits do not occur
        const uint32_t R = 1u << bits;
        const uint32_t start = blockExclusiveScan256(d < R ? digitTotals[d] : 0u, sWave, nullptr);
        sGlobalBase[d] = start + (d < R ? hist[d * gridDim.x + blockIdx.x] : 0u);
    }
    const uint64_t laneMaskLt = (uint64_t(1) << lane) - 1u;

    for (uint32_t tileBegin = begin; tileBegin < end; tileBegin += kSortTile) {
        const uint32_t tileN = end - tileBegin < uint32_t(kSortTile) ? end - tileBegin : uint32_t(kSortTile);
        for (uint32_t d = lane; d < kRadix; d += 64) sWaveHist[wave][d] = 0u;

        Key key[kSortKeysPerLane];
        uint32_t val[kSortKeysPerLane], dig[kSortKeysPerLane], rank[kSortKeysPerLane];
#pragma unroll
        for (int r = 0; r < kSortKeysPerLane; ++r) {
            const uint32_t pos = wave * (kSortTile / kSortWaves) + r * 64 + lane;
            const bool valid = pos < tileN;
            key[r] = valid ? keysIn[tileBegin + pos] : Key(0);
            val[r] = valid ? valsIn[tileBegin + pos] : 0u;
        }
#pragma unroll
        for (int r = 0; r < kSortKeysPerLane; ++r) {
            const uint32_t pos = wave * (kSortTile / kSortWaves) + r * 64 + lane;
            const bool valid = pos < tileN;
            const uint32_t d = digitOf(key[r], shift, bits);
            dig[r] = d;
            // Lanes holding the same digit: one ballot per digit bit.
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < kRadixBits; ++b) {
                if (uint32_t(b) >= bits) break;  // uniform
                const bool bit = (d >> b) & 1u;
                const uint64_t bal = __ballot(bit);
                peers &= bit ? bal : ~bal;
            }
            if (valid) {
                const uint64_t lower = peers & laneMaskLt;
                const uint32_t old = sWaveHist[wave][d];
                rank[r] = old + uint32_t(__popcll(lower));
                if (lower == 0u) sWaveHist[wave][d] = old + uint32_t(__popcll(peers));
            }
        }
        __syncthreads();
        {
            const uint32_t d = t;
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) {
                const uint32_t c = sWaveHist[w][d];
                sWaveHist[w][d] = run;
                run += c;
            }
            sTileCount[d] = run;
            sLocalStart[d] = blockExclusiveScan256(run, sWave, nullptr);
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSortKeysPerLane; ++r) {
            const uint32_t pos = wave * (kSortTile / kSortWaves) + r * 64 + lane;
            if (pos < tileN) {
                const uint32_t lr = sLocalStart[dig[r]] + sWaveHist[wave][dig[r]] + rank[r];
                sKeys[lr] = key[r];
                sVals[lr] = val[r];
            }
        }
        __syncthreads();
        for (uint32_t i = t; i < tileN; i += kSortThreads) {
            const Key k = sKeys[i];
            const uint32_t d = digitOf(k, shift, bits);
            const uint32_t dst = sGlobalBase[d] + (i - sLocalStart[d]);
            keysOut[dst] = k;
            valsOut[dst] = sVals[i];
        }
        __syncthreads();
        sGlobalBase[t] += sTileCount[t];
        __syncthreads();
    }
}

// Copies the first `count` pairs (count on the device): moves an odd-pass result back into
// the caller's buffers without touching the keys past the count.
template <typename Key>
__global__ __launch_bounds__(kSortThreads) void kCopyPairs(const Key* __restrict__ ks,
                                                           const uint32_t* __restrict__ vs, Key* __restrict__ kd,
                                                           uint32_t* __restrict__ vd,
                                                           const uint32_t* __restrict__ countDevice,
                                                           uint32_t maxCount) {
    const uint32_t count = sortCount(countDevice, maxCount);
    for (uint32_t i = blockIdx.x * kSortThreads + threadIdx.x; i < count; i += gridDim.x * kSortThreads) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

}  // namespace

// Digit width of pass p when `total` key bits are sorted in `passes` passes: the bits split as
// evenly as the passes allow, the wider digits first (13 bits: 7 + 6; 32: 8 + 8 + 8 + 8). Narrower
// digits give the scatter longer runs per digit (coalesced stores); any split gives the same
// stable order.
static uint32_t passBits(uint32_t total, uint32_t passes, uint32_t p) { return (total + passes - 1u - p) / passes; }

size_t sortStatusBytes(uint32_t maxCount) {
    const size_t tiles = (size_t(maxCount) + kSortTile - 1) / kSortTile;
    return (tiles > 0 ? tiles : 1) * kRadix * sizeof(uint64_t);
}

void launchCopyPairs(const uint64_t* ks, const uint32_t* vs, uint64_t* kd, uint32_t* vd, const uint32_t* countDevice,
                     uint32_t maxCount, hipStream_t stream) {
    uint32_t blocks = (maxCount + kSortThreads - 1) / kSortThreads;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) return;
    hipLaunchKernelGGL(kCopyPairs<uint64_t>, dim3(blocks), dim3(kSortThreads), 0, stream, ks, vs, kd, vd, countDevice,
                       maxCount);
    markKernel("kCopyPairs", stream);
}

// Reduce-then-scan LSD sort of 32-bit keys (depth keys, tile ids): 3 kernels per digit.
int launchRadixSort32(const SortWork& w, uint32_t* keysA, uint32_t* valuesA, uint32_t* keysB, uint32_t* valuesB,
                      uint32_t maxCount, const uint32_t* countDevice, uint32_t keyBits, hipStream_t stream) {
    if (maxCount == 0) return 0;
    if (keyBits > 32) keyBits = 32;
    const uint32_t passes = (keyBits + kRadixBits - 1) / kRadixBits;
    uint32_t blocks = (maxCount + kSortTile - 1) / kSortTile;
    if (blocks > kMaxSortBlocks) blocks = kMaxSortBlocks;
    uint32_t* kin = keysA;
    uint32_t* vin = valuesA;
    uint32_t* kout = keysB;
    uint32_t* vout = valuesB;
    uint32_t shift = 0;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t bits = passBits(keyBits, passes, p);
        hipLaunchKernelGGL(kSortHistogram<uint32_t>, dim3(blocks), dim3(kSortThreads), 0, stream, kin, countDevice,
                           maxCount, shift, bits, w.blockHist);
        markKernel("kSortHistogram", stream);
        hipLaunchKernelGGL(kSortScan, dim3(1u << bits), dim3(kSortThreads), 0, stream, w.blockHist, countDevice,
                           maxCount, blocks, bits, w.globalHist);
        markKernel("kSortScan", stream);
        hipLaunchKernelGGL(kSortScatter<uint32_t>, dim3(blocks), dim3(kSortThreads), 0, stream, kin, vin, kout, vout,
                           countDevice, maxCount, shift, bits, w.blockHist, w.globalHist);
        markKernel("kSortScatter", stream);
        shift += bits;
        uint32_t* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    return int(passes & 1u);
}

int launchRadixSort(const SortWork& w, uint64_t* keysA, uint32_t* valuesA, uint64_t* keysB, uint32_t* valuesB,
                    uint32_t maxCount, const uint32_t* countDevice, uint32_t firstBit, uint32_t keyBits,
                    uint32_t* epoch, hipStream_t stream) {
    if (maxCount == 0 || keyBits <= firstBit) return 0;
    uint32_t passes = (keyBits - firstBit + kRadixBits - 1) / kRadixBits;
    if (passes > kMaxSortPasses) passes = kMaxSortPasses;
    uint64_t* kin = keysA;
    uint32_t* vin = valuesA;
    uint64_t* kout = keysB;
    uint32_t* vout = valuesB;
    if (w.algorithm == kSortAlgReduceThenScan) {
        uint32_t blocks = (maxCount + kSortTile - 1) / kSortTile;
        if (blocks > kMaxSortBlocks) blocks = kMaxSortBlocks;
        const uint32_t bitsTotal = passes * kRadixBits < keyBits - firstBit ? passes * kRadixBits : keyBits - firstBit;
        uint32_t shift = firstBit;
        for (uint32_t p = 0; p < passes; ++p) {
            const uint32_t bits = passBits(bitsTotal, passes, p);
            hipLaunchKernelGGL(kSortHistogram<uint64_t>, dim3(blocks), dim3(kSortThreads), 0, stream, kin, countDevice,
                               maxCount, shift, bits, w.blockHist);
            markKernel("kSortHistogram", stream);
            hipLaunchKernelGGL(kSortScan, dim3(1u << bits), dim3(kSortThreads), 0, stream, w.blockHist, countDevice,
                               maxCount, blocks, bits, w.globalHist);
            markKernel("kSortScan", stream);
            hipLaunchKernelGGL(kSortScatter<uint64_t>, dim3(blocks), dim3(kSortThreads), 0, stream, kin, vin, kout,
                               vout, countDevice, maxCount, shift, bits, w.blockHist, w.globalHist);
            markKernel("kSortScatter", stream);
            shift += bits;
            uint64_t* tk = kin;
            kin = kout;
            kout = tk;
            uint32_t* tv = vin;
            vin = vout;
            vout = tv;
        }
        return int(passes & 1u);
    }
    const uint32_t tiles = (maxCount + kSortTile - 1) / kSortTile;
    hipLaunchKernelGGL(kSortInit, dim3(1), dim3(kSortThreads), 0, stream, w.globalHist, w.tickets, passes);
    markKernel("kSortInit", stream);
    uint32_t upBlocks = (maxCount + kSortThreads * 16 - 1) / (kSortThreads * 16);
    if (upBlocks > 512) upBlocks = 512;
    hipLaunchKernelGGL(kSortUpfront, dim3(upBlocks), dim3(kSortThreads), 0, stream, keysA, countDevice, maxCount,
                       firstBit, passes, w.globalHist);
    markKernel("kSortUpfront", stream);
    for (uint32_t p = 0; p < passes; ++p) {
        *epoch = (*epoch + 1u) & 0x3FFFFFFFu;
        if (*epoch == 0) *epoch = 1;  // 0 is the value of never-written status words
        hipLaunchKernelGGL(kSortOnesweep, dim3(tiles), dim3(kSortThreads), 0, stream, kin, vin, kout, vout, countDevice,
                           maxCount, firstBit + p * kRadixBits, w.globalHist + p * kRadix, w.status, w.tickets + p, *epoch,
                           w.timeouts);
        markKernel("kSortOnesweep", stream);
        uint64_t* tk = kin;
        kin = kout;
        kout = tk;
        uint32_t* tv = vin;
        vin = vout;
        vout = tv;
    }
    return int(passes & 1u);
}

}  // namespace gsr
