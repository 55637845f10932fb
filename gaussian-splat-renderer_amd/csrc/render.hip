// K8 tile_render + K9 resolve (reference shaders/tile_render.comp:147-199, 281-378; CPU twin
// splat/reference_renderer.cpp:190-222): one 256-thread workgroup per 16x16 tile, one
// thread per pixel, front to back over the tile's range of the sorted list.
//
//   - batches of 256 splats are staged in LDS (centre, conic + opacity, colour), every lane
//     then reads the same LDS word (broadcast, no bank conflicts);
//   - whole-tile early exit: the workgroup stops once all 256 pixels are done
//     (__syncthreads_count, INRIA's __syncthreads_count(done));
//   - per pixel: alpha = min(0.99, o exp(power)), skip power > 0 and alpha < 1/255,
//     test-before-accept T (1 - alpha) < 1e-4;
//   - the result C + T * black, alpha 1 - T, is written once: RGBA8 (clamp, round(v * 255))
//     or RGBA32F;
//   - XCD-aware tile order: workgroup b runs on XCD b % 8, so b -> tile maps each XCD to a
//     contiguous band of tiles and neighbouring tiles (which share splats) share an L2.
#include "gsr_internal.hpp"

namespace gsr {

namespace {

constexpr int kRenderThreads = kTileSize * kTileSize;
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t quantize8(float v) {
    v = v > 0.0f ? v : 0.0f;  // also NaN -> 0
    v = v < 1.0f ? v : 1.0f;
    return uint32_t(roundf(v * 255.0f));
}

template <int FORMAT>
__global__ __launch_bounds__(kRenderThreads) void kRender(const float4* __restrict__ splats,
                                                          const uint32_t* __restrict__ values,
                                                          const uint2* __restrict__ ranges, uint32_t width,
                                                          uint32_t height, uint32_t tilesX, uint32_t tileCount,
                                                          uint32_t tilesPerXcd, uint32_t gaussianCount,
                                                          void* __restrict__ out) {
    __shared__ float2 sCenter[kRenderThreads];
    __shared__ float4 sConic[kRenderThreads];
    __shared__ float4 sColor[kRenderThreads];

    const uint32_t tile = (blockIdx.x % kXcds) * tilesPerXcd + blockIdx.x / kXcds;
    if (tile >= tileCount) return;  // uniform
    const uint32_t t = threadIdx.x;
    const uint32_t px = (tile % tilesX) * kTileSize + (t % kTileSize);
    const uint32_t py = (tile / tilesX) * kTileSize + (t / kTileSize);
    const bool inside = px < width && py < height;
    const float pxf = float(px), pyf = float(py);

    const uint2 range = ranges[tile];
    float T = 1.0f;
    float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    bool done = !inside;

    for (uint32_t batch = range.x; batch < range.y; batch += kRenderThreads) {
        // Whole-tile vote; also the barrier that protects the previous batch's LDS.
        if (__syncthreads_count(done) == kRenderThreads) break;
        const uint32_t e = batch + t;
        if (e < range.y) {
            uint32_t g = values[e];
            g = g < gaussianCount ? g : 0u;  // list values are gaussian indices; never read past the scene
            const float4* s = splats + 3 * size_t(g);
            const float4 o0 = s[0];
            sCenter[t] = make_float2(o0.x, o0.y);
            sConic[t] = s[1];
            sColor[t] = s[2];
        }
        __syncthreads();
        const uint32_t n = range.y - batch < uint32_t(kRenderThreads) ? range.y - batch : uint32_t(kRenderThreads);
        for (uint32_t j = 0; !done && j < n; ++j) {
            const float2 c = sCenter[j];
            const float4 co = sConic[j];
            const float dx = c.x - pxf;
            const float dy = c.y - pyf;
            const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
            if (power > 0.0f) continue;
            const float ex = co.w * expf(power);
            const float alpha = kMaxAlpha < ex ? kMaxAlpha : ex;
            if (alpha < kMinAlpha) continue;
            const float testT = T * (1.0f - alpha);
            if (testT < kMinTransmittance) {
                done = true;
                break;
            }
            const float4 col = sColor[j];
            C0 += col.x * alpha * T;
            C1 += col.y * alpha * T;
            C2 += col.z * alpha * T;
            T = testT;
        }
    }
    if (!inside) return;
    const size_t p = size_t(py) * width + px;
    if (FORMAT == GSR_TARGET_RGBA8) {
        const uint32_t rgba = quantize8(C0) | (quantize8(C1) << 8) | (quantize8(C2) << 16) | (quantize8(1.0f - T) << 24);
        static_cast<uint32_t*>(out)[p] = rgba;
    } else {
        static_cast<float4*>(out)[p] = make_float4(C0, C1, C2, 1.0f - T);
    }
}

}  // namespace

void launchRender(const float4* splats, const uint32_t* values, const uint2* ranges, const DevCamera& cam,
                  uint32_t gaussianCount, gsr_target_format format, void* out, hipStream_t stream) {
    const uint32_t tileCount = cam.tilesX * cam.tilesY;
    if (tileCount == 0) return;
    const uint32_t perXcd = (tileCount + kXcds - 1) / kXcds;
    const uint32_t blocks = perXcd * kXcds;
    if (format == GSR_TARGET_RGBA8)
        hipLaunchKernelGGL(kRender<GSR_TARGET_RGBA8>, dim3(blocks), dim3(kRenderThreads), 0, stream, splats, values,
                           ranges, cam.width, cam.height, cam.tilesX, tileCount, perXcd, gaussianCount, out);
    else
        hipLaunchKernelGGL(kRender<GSR_TARGET_RGBA32F>, dim3(blocks), dim3(kRenderThreads), 0, stream, splats,
                           values, ranges, cam.width, cam.height, cam.tilesX, tileCount, perXcd, gaussianCount, out);
}

}  // namespace gsr
