// K2 prefix sums (reference GpuScan, render/gpu_scan.h:25-76, shaders/scan_*.comp) and K4
// tile_setup (shaders/tile_setup.comp:16-24).
//
// Reduce-then-scan. In the frame, the reduce is fused into the preprocess (one BlockStats
// per 256 gaussians) and the downsweep into the duplicate pass, so the frame only runs the
// middle step here: one workgroup scans the block sums and writes the control block
// (N, V, M, entries = min(M, capacity), overflow, cull counts), which also replaces the
// per-frame control reset. The stand-alone GpuScan (gsr_exclusive_scan) runs all three.
#include "gsr_internal.hpp"

namespace gsr {

namespace {

constexpr int kSetupThreads = 1024;

// Inclusive scan of one value per thread over a 1024-thread block; returns the block total
// through *total. Uses wave shuffles then one LDS step across the 16 waves.
__device__ __forceinline__ uint32_t blockInclusiveScan1024(uint32_t v, uint32_t* sWave, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t n = __shfl_up(v, off, 64);
        if (lane >= uint32_t(off)) v += n;
    }
    if (lane == 63) sWave[wave] = v;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < kSetupThreads / 64 ? sWave[lane] : 0u;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            const uint32_t n = __shfl_up(w, off, 64);
            if (lane >= uint32_t(off)) w += n;
        }
        if (lane < kSetupThreads / 64) sWave[lane] = w;
    }
    __syncthreads();
    const uint32_t prefix = wave > 0 ? sWave[wave - 1] : 0u;
    *total = sWave[kSetupThreads / 64 - 1];
    __syncthreads();
    return v + prefix;
}

constexpr uint32_t kSetupPer = 8;  // block records per thread and round

// Rounds of 8 x 1024 block records: every record of the round is loaded at once (coalesced:
// record j * 1024 + t by thread t), its scanned value staged in LDS; then thread t scans the 8
// consecutive values 8 t .. 8 t + 7 and the workgroup scan adds the threads' sums. One load
// latency and one workgroup scan per round (a single round up to 2M gaussians).
__global__ __launch_bounds__(kSetupThreads) void kFrameSetup(const BlockStats* __restrict__ stats, uint32_t numBlocks,
                                                             uint32_t gaussianCount, uint32_t capacity,
                                                             uint32_t scanVisible, uint32_t* __restrict__ blockOffsets,
                                                             gsr_control* __restrict__ control) {
    __shared__ uint32_t sWave[kSetupThreads / 64];
    __shared__ uint32_t sCull[7];  // 5 cull reasons, non-zero counts, M
    __shared__ uint32_t sV[kSetupPer * kSetupThreads];
    if (threadIdx.x < 7) sCull[threadIdx.x] = 0u;
    __syncthreads();
    uint32_t carry = 0, mTotal = 0;
    uint32_t cull[6] = {0, 0, 0, 0, 0, 0};  // [0..4] reasons, [5] non-zero counts
    const uint32_t t = threadIdx.x;
    for (uint32_t base = 0; base < numBlocks; base += kSetupPer * kSetupThreads) {
        BlockStats rec[kSetupPer];
#pragma unroll
        for (uint32_t j = 0; j < kSetupPer; ++j) {
            const uint32_t b = base + j * kSetupThreads + t;
            rec[j] = b < numBlocks ? stats[b] : BlockStats{};
        }
#pragma unroll
        for (uint32_t j = 0; j < kSetupPer; ++j) {
            sV[j * kSetupThreads + t] = scanVisible ? rec[j].nonZero : rec[j].tileSum;
            mTotal += rec[j].tileSum;
            cull[5] += rec[j].nonZero;
#pragma unroll
            for (int r = 0; r < 5; ++r) cull[r] += rec[j].cull[r];
        }
        __syncthreads();
        uint32_t v[kSetupPer], sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < kSetupPer; ++k) {
            v[k] = sV[t * kSetupPer + k];
            sum += v[k];
        }
        uint32_t total;
        uint32_t run = carry + blockInclusiveScan1024(sum, sWave, &total) - sum;
#pragma unroll
        for (uint32_t k = 0; k < kSetupPer; ++k) {
            const uint32_t b = base + t * kSetupPer + k;
            if (b < numBlocks) blockOffsets[b] = run;
            run += v[k];
        }
        carry += total;
        __syncthreads();  // sV is rewritten by the next round
    }
    {  // M, reduced across the workgroup
        uint32_t m = mTotal;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m += __shfl_xor(m, off, 64);
        if ((threadIdx.x & 63u) == 0 && m) atomicAdd(&sCull[6], m);
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        uint32_t c = cull[r];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
        if ((threadIdx.x & 63u) == 0 && c) atomicAdd(&sCull[r], c);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t m = sCull[6];
        const uint32_t entries = m < capacity ? m : capacity;
        gsr_control c{};
        c.gaussian_count = gaussianCount;
        c.visible = sCull[5];
        c.duplicates = m;
        c.entries = entries;
        for (int r = 0; r < 5; ++r) c.cull_counts[r] = sCull[r];
        c.overflow = m - entries;
        c.ranges_groups[0] = (entries + 255u) / 256u;
        c.ranges_groups[1] = 1u;
        c.ranges_groups[2] = 1u;
        c.coverage_mismatch = 0u;
        *control = c;
    }
}

// --- Stand-alone exclusive scan ----------------------------------------------------------------

__device__ __forceinline__ uint32_t blockInclusiveScan256(uint32_t v, uint32_t* sWave) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t n = __shfl_up(v, off, 64);
        if (lane >= uint32_t(off)) v += n;
    }
    if (lane == 63) sWave[wave] = v;
    __syncthreads();
    uint32_t prefix = 0;
    for (uint32_t w = 0; w < wave; ++w) prefix += sWave[w];
    __syncthreads();
    return v + prefix;
}

__global__ __launch_bounds__(kScanThreads) void kScanReduce(const uint32_t* __restrict__ in, uint32_t count,
                                                            uint32_t* __restrict__ blockSums) {
    const uint32_t base = blockIdx.x * kScanBlock;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t idx = base + j * kScanThreads + threadIdx.x;
        if (idx < count) s += in[idx];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ uint32_t sw[kScanThreads / 64];
    if ((threadIdx.x & 63u) == 0) sw[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += sw[w];
        blockSums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kSetupThreads) void kScanBlockSums(uint32_t* __restrict__ blockSums, uint32_t numBlocks,
                                                                uint32_t* __restrict__ total) {
    __shared__ uint32_t sWave[kSetupThreads / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < numBlocks; base += kSetupThreads) {
        const uint32_t b = base + threadIdx.x;
        const uint32_t v = b < numBlocks ? blockSums[b] : 0u;
        uint32_t t;
        const uint32_t inc = blockInclusiveScan1024(v, sWave, &t);
        if (b < numBlocks) blockSums[b] = carry + inc - v;
        carry += t;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// Binning mode: exclusive scan of the per-256 sums of the depth-ordered tile counts.
__global__ __launch_bounds__(kSetupThreads) void kScanVisibleBlocks(uint32_t* __restrict__ blockSums,
                                                                    const gsr_control* __restrict__ control) {
    __shared__ uint32_t sWave[kSetupThreads / 64];
    const uint32_t numBlocks = (control->visible + 255u) / 256u;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < numBlocks; base += kSetupThreads) {
        const uint32_t b = base + threadIdx.x;
        const uint32_t v = b < numBlocks ? blockSums[b] : 0u;
        uint32_t t;
        const uint32_t inc = blockInclusiveScan1024(v, sWave, &t);
        if (b < numBlocks) blockSums[b] = carry + inc - v;
        carry += t;
    }
}

// Each thread scans kScanItems consecutive elements; the block adds its scanned offset.
__global__ __launch_bounds__(kScanThreads) void kScanDownsweep(const uint32_t* __restrict__ in,
                                                               uint32_t* __restrict__ out, uint32_t count,
                                                               const uint32_t* __restrict__ blockOffsets) {
    __shared__ uint32_t sWave[kScanThreads / 64];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = base + j < count ? in[base + j] : 0u;
        s += v[j];
    }
    const uint32_t inc = blockInclusiveScan256(s, sWave);
    uint32_t run = blockOffsets[blockIdx.x] + inc - s;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < count) out[base + j] = run;
        run += v[j];
    }
}

}  // namespace

void launchFrameSetup(const BlockStats* blockStats, uint32_t numBlocks, uint32_t gaussianCount, uint32_t capacity,
                      bool scanVisible, uint32_t* blockOffsets, gsr_control* control, hipStream_t stream) {
    hipLaunchKernelGGL(kFrameSetup, dim3(1), dim3(kSetupThreads), 0, stream, blockStats, numBlocks, gaussianCount,
                       capacity, scanVisible ? 1u : 0u, blockOffsets, control);
    markKernel("kFrameSetup", stream);
}

void launchScanVisibleBlocks(uint32_t* blockSums, const gsr_control* control, hipStream_t stream) {
    hipLaunchKernelGGL(kScanVisibleBlocks, dim3(1), dim3(kSetupThreads), 0, stream, blockSums, control);
    markKernel("kScanVisibleBlocks", stream);
}

void launchExclusiveScan(const uint32_t* in, uint32_t* out, uint32_t count, uint32_t* blockSums, uint32_t* total,
                         hipStream_t stream) {
    const uint32_t blocks = (count + kScanBlock - 1) / kScanBlock;
    if (blocks > 0) hipLaunchKernelGGL(kScanReduce, dim3(blocks), dim3(kScanThreads), 0, stream, in, count, blockSums);
    hipLaunchKernelGGL(kScanBlockSums, dim3(1), dim3(kSetupThreads), 0, stream, blockSums, blocks, total);
    markKernel("kScanBlockSums", stream);
    if (blocks > 0)
        hipLaunchKernelGGL(kScanDownsweep, dim3(blocks), dim3(kScanThreads), 0, stream, in, out, count, blockSums);
        markKernel("kScanDownsweep", stream);
}

}  // namespace gsr
