// Scene upload layout (the reference's GpuScene::upload, render/gpu_scene.h:42-78, builds its
// device buffers once per scene; this is that step for HIP).
//
// gfx950 layout: the per-gaussian records wider than 16 B (f32 SH: up to 48 floats; f16
// geometry: 2 uint4; f16 colour: up to 6 uint4) are stored as planes of 16-B quads,
// plane q = quad q of every gaussian. A preprocess load of quad q by a wave then reads 64
// consecutive 16-B words (1 KB, 8 cache lines) instead of 64 words 48-192 B apart (64 lines),
// which is what bounds the address path of the record-per-gaussian layout.
#include "gsr_internal.hpp"

namespace gsr {

namespace {

constexpr int kLayoutThreads = 256;

// planes[q * count + i] = floats [4q, 4q + 4) of gaussian i's record of `floatsPer` floats
// (zero past the record's end).
__global__ __launch_bounds__(kLayoutThreads) void kFloatsToPlanes(const float* __restrict__ src, uint32_t count,
                                                                  uint32_t floatsPer, uint32_t quads,
                                                                  float4* __restrict__ planes) {
    const size_t total = size_t(count) * quads;
    for (size_t k = size_t(blockIdx.x) * kLayoutThreads + threadIdx.x; k < total;
         k += size_t(gridDim.x) * kLayoutThreads) {
        const uint32_t q = uint32_t(k / count), i = uint32_t(k % count);
        const float* r = src + size_t(i) * floatsPer;
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t j = 4 * q + uint32_t(e);
            v[e] = j < floatsPer ? r[j] : 0.0f;
        }
        planes[k] = make_float4(v[0], v[1], v[2], v[3]);
    }
}

// planes[q * count + i] = src[i * quads + q] (records that are whole uint4s).
__global__ __launch_bounds__(kLayoutThreads) void kQuadsToPlanes(const uint4* __restrict__ src, uint32_t count,
                                                                 uint32_t quads, uint4* __restrict__ planes) {
    const size_t total = size_t(count) * quads;
    for (size_t k = size_t(blockIdx.x) * kLayoutThreads + threadIdx.x; k < total;
         k += size_t(gridDim.x) * kLayoutThreads) {
        const uint32_t q = uint32_t(k / count), i = uint32_t(k % count);
        planes[k] = src[size_t(i) * quads + q];
    }
}

uint32_t layoutBlocks(size_t total) {
    const size_t b = (total + kLayoutThreads - 1) / kLayoutThreads;
    return uint32_t(b < 65536 ? (b ? b : 1) : 65536);
}

}  // namespace

void launchFloatsToPlanes(const float* src, uint32_t count, uint32_t floatsPer, uint32_t quads, float4* planes,
                          hipStream_t stream) {
    if (count == 0 || quads == 0) return;
    hipLaunchKernelGGL(kFloatsToPlanes, dim3(layoutBlocks(size_t(count) * quads)), dim3(kLayoutThreads), 0, stream,
                       src, count, floatsPer, quads, planes);
}

void launchQuadsToPlanes(const uint4* src, uint32_t count, uint32_t quads, uint4* planes, hipStream_t stream) {
    if (count == 0 || quads == 0) return;
    hipLaunchKernelGGL(kQuadsToPlanes, dim3(layoutBlocks(size_t(count) * quads)), dim3(kLayoutThreads), 0, stream,
                       src, count, quads, planes);
}

}  // namespace gsr
