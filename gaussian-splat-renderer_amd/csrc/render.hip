// K8 tile_render + K9 resolve (reference shaders/tile_render.comp:147-199, 281-378; CPU twin
// splat/reference_renderer.cpp:190-222): one workgroup per 16x16 tile, front to back over
// the tile's range of the sorted list, INRIA's blend:
//   alpha = min(0.99, o exp(power)), skip power > 0 and alpha < 1/255,
//   test-before-accept T (1 - alpha) < 1e-4, C += colour alpha T, output C + T * black and
//   alpha 1 - T, as RGBA8 (clamp, round(v * 255)) or RGBA32F.
//
// gfx950 design (VALU-bound: every (pixel, splat) pair costs a few dozen lane ops):
//   - batches of 256 splats are staged in LDS; all lanes read the same LDS words
//     (broadcast, no bank conflicts);
//   - PPT pixels per thread (1, 2 or 4 rows of the same column): the LDS reads, the x
//     offset and the loop overhead are shared by the thread's pixels; 256 / PPT threads per
//     tile;
//   - alpha in the log2 domain (FAST): at fetch time each splat's conic is scaled by
//     -log2(e)/2 and log2(o) is stored, so a pair costs 5 FMA-class ops to get
//     q = log2(o exp(power)); pairs with q < log2(1/255) (alpha < 1/255) are skipped before
//     any exponential, the rest take one v_exp_f32. EXACT keeps the reference's operation
//     order and expf for every pair (bitwise the CPU expression, up to expf's last ulp);
//   - whole-tile early exit: the workgroup stops once all pixels are done (__syncthreads_count);
//   - culled loop (the default): each wave owns an 8x8 quadrant of the tile. At fetch time
//     each splat gets the bounding box of its alpha >= 1/255 ellipse (padded so that it
//     also holds for the FP32 expression of q); each wave compacts the batch to the splats
//     whose box meets its quadrant and walks only those. The pairs it drops all have
//     q < log2(1/255), so the image is the same, bit for bit, as the uncompacted loop's;
//   - XCD-aware tile order: workgroup b runs on XCD b % 8; XCD x gets the tile rows x, x + 8,
//     ... so horizontally neighbouring tiles (which share splats) share an L2 while every
//     XCD gets an even share of the dense (image-centre) and sparse rows.
#include <type_traits>

#include "gsr_device.hpp"
#include "tilesort.hpp"

namespace gsr {

namespace {

constexpr uint32_t kXcds = 8;

// Workgroup b runs on XCD b % 8, which renders tile rows xcd, xcd + 8, ... (neighbouring tiles
// share that XCD's L2; every XCD gets an even share of dense rows). An XCD takes its rows
// centre-out: slot s of n is row k = m, m + 1, m - 1, m + 2, ... (m = (n - 1) / 2), so the dense
// middle of the image starts first and the sparse edges fill the tail.
__device__ __forceinline__ uint32_t centreOut(uint32_t s, uint32_t n) {
    const uint32_t m = (n - 1u) / 2u;
    return (s & 1u) ? m + (s + 1u) / 2u : m - s / 2u;
}
constexpr int kBatch = 256;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLog2MinAlpha = -7.994353436858858f;  // log2(1/255)

// The render's view of splat g: centre, (conic a, b, c, opacity), colour. RS = the record the
// preprocess wrote for the render (the 48-byte splat, or the p32 / p24 words).
template <int RS>
__device__ __forceinline__ void fetchSplat(const float4* __restrict__ splats, const uint32_t* __restrict__ packed,
                                           uint32_t g, float& cx, float& cy, float4& co, float4& col) {
    if (RS == GSR_RENDER_SPLAT_FULL) {
        const float4* s = splats + 3 * size_t(g);
        const float4 o0 = s[0];
        cx = o0.x;
        cy = o0.y;
        co = s[1];
        col = s[2];
    } else if (RS == GSR_RENDER_SPLAT_P32) {
        const uint4* w = reinterpret_cast<const uint4*>(packed) + 2 * size_t(g);
        const uint4 a = w[0], b = w[1];
        cx = __uint_as_float(a.x);
        cy = __uint_as_float(a.y);
        co = make_float4(__uint_as_float(a.z), __uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y));
        col = make_float4(halfBitsToFloat(b.z & 0xFFFFu), halfBitsToFloat(b.z >> 16), halfBitsToFloat(b.w & 0xFFFFu),
                          0.0f);
    } else {
        const uint2* w = reinterpret_cast<const uint2*>(packed) + 3 * size_t(g);
        const uint2 a = w[0], b = w[1], c = w[2];
        cx = __uint_as_float(a.x);
        cy = __uint_as_float(a.y);
        co = make_float4(halfBitsToFloat(b.x & 0xFFFFu) * kInvConicScaleP24, halfBitsToFloat(b.x >> 16) * kInvConicScaleP24,
                         halfBitsToFloat(b.y & 0xFFFFu) * kInvConicScaleP24, halfBitsToFloat(b.y >> 16));
        col = make_float4(halfBitsToFloat(c.x & 0xFFFFu), halfBitsToFloat(c.x >> 16), halfBitsToFloat(c.y & 0xFFFFu),
                          0.0f);
    }
}

// The predicated loops' per-splat record: q = log2(o exp(power)) at the pixel (x0 + X, y0 + Y)
// of the tile with origin (x0, y0), as a quadratic in the tile-local X, Y:
//   q = F + D X + E Y + A X^2 + B X Y + C Y^2,
// A, B, C = the conic scaled by -log2(e)/2 (B: -log2 e), D = -2 A dx - B dy, E = -B dx - 2 C dy,
// F = q at the tile origin, dx, dy = the centre relative to the tile origin. s0 = (A, B, C, D),
// s1 = (E, F): a pair then costs 5 FMAs against per-lane constants X, Y, X^2, XY, Y^2.
__device__ __forceinline__ void tileQuadratic(float cx, float cy, float4 co, float x0, float y0, float4& s0,
                                              float2& s1) {
    const float A = -0.5f * kLog2e * co.x, B = -kLog2e * co.y, C = -0.5f * kLog2e * co.z;
    const float L = __log2f(co.w);
    const float dx = cx - x0, dy = cy - y0;
    const float D = -2.0f * A * dx - B * dy;
    const float E = -B * dx - 2.0f * C * dy;
    const float F = fmaf(fmaf(A, dx, B * dy), dx, fmaf(C * dy, dy, L));
    s0 = make_float4(A, B, C, D);
    s1 = make_float2(E, F);
}

struct PixelQ {
    float X, Y, X2, XY, Y2;
};

// A batch slot of the culled loop: the quadratic (s0, s1), qmax and the colour, 48 B. qmax =
// log2(o) for a splat whose conic is not positive definite (a pair with q > qmax has power > 0
// and is not drawn), +inf otherwise (power <= 0 everywhere).
struct __align__(16) SplatRecord {
    float4 s0;
    float2 s1;
    float qmax, pad;
    float4 col;
};

__device__ __forceinline__ PixelQ pixelQ(uint32_t lx, uint32_t ly) {
    const float X = float(lx), Y = float(ly);
    return PixelQ{X, Y, X * X, X * Y, Y * Y};
}

__device__ __forceinline__ float evalQ(const float4& s0, const float2& s1, const PixelQ& p) {
    return fmaf(s0.x, p.X2, fmaf(s0.y, p.XY, fmaf(s0.z, p.Y2, fmaf(s0.w, p.X, fmaf(s1.x, p.Y, s1.y)))));
}

// One splat into one pixel: the lane's state is its transmittance T, colour and cutoff thr
// (log2(1/255) while alive, +inf once done or outside the image). A pair below the cutoff changes
// nothing (a = 0, so w = 0 and T' = T); otherwise the splat is blended with the reference's weight
// w = alpha T and T' = T - w, unless T' < 1e-4: then it is not blended and ends the pixel
// (test-before-accept). Termination is rare, so its selects run only when some lane of the wave
// ends at this splat; a finished lane keeps T >= 1e-4 and never ends again.
__device__ __forceinline__ void blendPredicated(float q, float4 col, float& T, float& c0, float& c1, float& c2,
                                                float& thr) {
    const float a = q >= thr ? fminf(kMaxAlpha, __builtin_amdgcn_exp2f(q)) : 0.0f;
    float w = a * T;
    float testT = T - w;
    const bool term = testT < kMinTransmittance;
    if (__builtin_amdgcn_ballot_w64(term) != 0) {  // rare: some lane of the wave ends here
        __asm__ volatile("" ::: "memory");  // keeps this a branch: the compiler would make it selects
        w = term ? 0.0f : w;
        testT = term ? T : testT;
        thr = term ? __builtin_inff() : thr;
    }
    c0 = fmaf(col.x, w, c0);
    c1 = fmaf(col.y, w, c1);
    c2 = fmaf(col.z, w, c2);
    T = testT;
}


__device__ __forceinline__ uint32_t quantize8(float v) {
    v = v > 0.0f ? v : 0.0f;  // also NaN -> 0
    v = v < 1.0f ? v : 1.0f;
    return uint32_t(roundf(v * 255.0f));
}

template <int FORMAT, bool FAST, int PPT, bool PRED, int RS>
__global__ __launch_bounds__(256 / PPT) void kRender(const float4* __restrict__ splats,
                                                     const uint32_t* __restrict__ packed,
                                                     const uint32_t* __restrict__ values,
                                                     const uint2* __restrict__ ranges, uint32_t width,
                                                     uint32_t height, uint32_t tilesX, uint32_t tileCount,
                                                     uint32_t gaussianCount,
                                                     void* __restrict__ out) {
    constexpr int kThreads = 256 / PPT;
    constexpr int kRowStep = 16 / PPT;
    // PRED: the tile-local quadratic (tileQuadratic). FAST: s0 = (cx, cy, A, B), s1 = (C, log2 o),
    // A/B/C the conic scaled by -log2(e)/2 (B: -log2 e). EXACT: s0 = (cx, cy, a, b), s1 = (c, o).
    __shared__ float4 sS0[kBatch];
    __shared__ float2 sS1[kBatch];
    __shared__ float4 sColor[kBatch];
    __shared__ float sQmax[PRED ? kBatch : 1];  // PRED: qmax as in the culled loop's records
    static_assert(!PRED || (FAST && PPT == 1), "the predicated loop is the fast one-pixel form");

    // Row-interleaved XCD mapping: workgroup b runs on XCD b % 8; XCD x renders tile rows
    // x, x + 8, x + 16, ... in row-major order.
    const uint32_t xcd = blockIdx.x % kXcds, idx = blockIdx.x / kXcds;
    const uint32_t row = xcd + kXcds * centreOut(idx / tilesX, gridDim.x / kXcds / tilesX);
    const uint32_t tile = row * tilesX + idx % tilesX;
    if (tile >= tileCount) return;  // uniform
    const uint32_t t = threadIdx.x;
    const uint32_t tx0 = (tile % tilesX) * 16, ty0 = (tile / tilesX) * 16;
    const uint32_t px = tx0 + (t % 16);
    const uint32_t py0 = ty0 + (t / 16);
    const float pxf = float(px);

    float T[PPT], C0[PPT], C1[PPT], C2[PPT], pyf[PPT];
    bool done[PPT];
    bool alive = false;
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        const uint32_t py = py0 + i * kRowStep;
        pyf[i] = float(py);
        T[i] = 1.0f;
        C0[i] = C1[i] = C2[i] = 0.0f;
        done[i] = !(px < width && py < height);
        alive = alive || !done[i];
    }

    const uint2 range = ranges[tile];
    for (uint32_t batch = range.x; batch < range.y; batch += kBatch) {
        // Whole-tile vote; also the barrier that protects the previous batch's LDS.
        if (__syncthreads_count(alive ? 0 : 1) == kThreads) break;
#pragma unroll
        for (int f = 0; f < kBatch / kThreads; ++f) {
            const uint32_t slot = t + f * kThreads;
            const uint32_t e = batch + slot;
            if (PRED && e >= range.y && slot < ((range.y - batch + 3u) & ~3u)) {
                // Padding up to a multiple of 4 for the unrolled loop: q = -inf never passes.
                sS0[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                sS1[slot] = make_float2(0.0f, -__builtin_inff());
                sQmax[slot] = __builtin_inff();
            }
            if (e < range.y) {
                uint32_t g = values[e];
                g = g < gaussianCount ? g : 0u;  // list values are gaussian indices; never read past the scene
                float cx, cy;
                float4 co, col;
                fetchSplat<RS>(splats, packed, g, cx, cy, co, col);
                sColor[slot] = col;
                if (PRED) {
                    tileQuadratic(cx, cy, co, float(tx0), float(ty0), sS0[slot], sS1[slot]);
                    sQmax[slot] = conicPositiveDefinite(co.x, co.y, co.z) ? __builtin_inff() : __log2f(co.w);
                } else if (FAST) {
                    sS0[slot] = make_float4(cx, cy, -0.5f * kLog2e * co.x, -kLog2e * co.y);
                    sS1[slot] = make_float2(-0.5f * kLog2e * co.z, __log2f(co.w));
                } else {
                    sS0[slot] = make_float4(cx, cy, co.x, co.y);
                    sS1[slot] = make_float2(co.z, co.w);
                }
            }
        }
        __syncthreads();
        const uint32_t n = range.y - batch < uint32_t(kBatch) ? range.y - batch : uint32_t(kBatch);
        if (PRED) {
            // Predicated form (FAST, one pixel per thread): the wave walks the batch in step,
            // four splats per iteration. A lane's state is folded into its cutoff `thr`
            // (log2(1/255) while alive, +inf once done or outside the image), so every branch is
            // one vector compare: the exponential and the blend run only when some lane of the
            // wave passes its cutoff, and the blend itself is branch-free (selects).
            const PixelQ pq = pixelQ(t % 16, t / 16);
            float thr = done[0] ? __builtin_inff() : kLog2MinAlpha;
            const uint32_t n4 = (n + 3u) & ~3u;
            for (uint32_t j = 0; j < n4; j += 4) {
                // The four q first (their LDS reads in flight together), then the blends in order.
                float q[4];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    q[u] = evalQ(sS0[j + u], sS1[j + u], pq);
                    q[u] = q[u] > sQmax[j + u] ? -__builtin_inff() : q[u];  // power > 0: not drawn
                }
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    if (__builtin_amdgcn_ballot_w64(q[u] >= thr) == 0) continue;
                    blendPredicated(q[u], sColor[j + u], T[0], C0[0], C1[0], C2[0], thr);
                }
                if (__builtin_amdgcn_ballot_w64(thr < __builtin_inff()) == 0) break;
            }
            done[0] = !(thr < __builtin_inff());
            alive = !done[0];
            continue;
        }
        for (uint32_t j = 0; alive && j < n; ++j) {
            const float4 s0 = sS0[j];
            const float2 s1 = sS1[j];
            const float dx = s0.x - pxf;
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                if (done[i]) continue;
                const float dy = s0.y - pyf[i];
                float alpha;
                if (FAST) {
                    // q = log2(o) + log2(e) * power, power = -0.5 (a dx^2 + c dy^2) - b dx dy.
                    // The exponential, the cutoff and the blend form one guarded block, so the
                    // (quarter-rate) v_exp_f32 is not speculated for the pairs the cutoff skips.
                    const float p = fmaf(fmaf(s0.z, dx, s0.w * dy), dx, s1.x * dy * dy);
                    const float q = p + s1.y;
                    if (q >= kLog2MinAlpha && p <= 0.0f) {
                        const float ex = __builtin_amdgcn_exp2f(q);
                        const float a = kMaxAlpha < ex ? kMaxAlpha : ex;
                        if (a >= kMinAlpha) {
                            const float testT = T[i] * (1.0f - a);
                            if (testT < kMinTransmittance) {
                                done[i] = true;
                            } else {
                                const float4 col = sColor[j];
                                const float w = a * T[i];
                                C0[i] = fmaf(col.x, w, C0[i]);
                                C1[i] = fmaf(col.y, w, C1[i]);
                                C2[i] = fmaf(col.z, w, C2[i]);
                                T[i] = testT;
                            }
                        }
                    }
                    continue;
                } else {
                    const float power = -0.5f * (s0.z * dx * dx + s1.x * dy * dy) - s0.w * dx * dy;
                    if (power > 0.0f) continue;
                    const float ex = s1.y * expf(power);
                    alpha = kMaxAlpha < ex ? kMaxAlpha : ex;
                    if (alpha < kMinAlpha) continue;
                }
                const float testT = T[i] * (1.0f - alpha);
                if (testT < kMinTransmittance) {
                    done[i] = true;
                    continue;
                }
                const float4 col = sColor[j];
                C0[i] += col.x * alpha * T[i];
                C1[i] += col.y * alpha * T[i];
                C2[i] += col.z * alpha * T[i];
                T[i] = testT;
            }
            bool any = false;
#pragma unroll
            for (int i = 0; i < PPT; ++i) any = any || !done[i];
            alive = any;
        }
    }
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
        const uint32_t py = py0 + i * kRowStep;
        if (!(px < width && py < height)) continue;
        const size_t p = size_t(py) * width + px;
        if (FORMAT == GSR_TARGET_RGBA8) {
            const uint32_t rgba = quantize8(C0[i]) | (quantize8(C1[i]) << 8) | (quantize8(C2[i]) << 16) |
                                  (quantize8(1.0f - T[i]) << 24);
            static_cast<uint32_t*>(out)[p] = rgba;
        } else {
            static_cast<float4*>(out)[p] = make_float4(C0[i], C1[i], C2[i], 1.0f - T[i]);
        }
    }
}

__device__ __forceinline__ uint32_t waveSum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ uint32_t waveMax(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Culled predicated loop (FAST, one pixel per thread, 8x8 quadrant per wave).
// STATS: the same render, also counting per tile the reference's tile statistics (TileStat,
// splat_tile_renderer.h:130-141; tile_render.comp:78-90, 380-399) into tileStats[tile * 8 + k]
// (zeroed by the caller): k = 0 list length, 1 entries fetched before the tile stopped,
// 2 (pixel, entry) pairs evaluated in the reference's loop (a pixel walks every fetched entry
// until it terminates, the terminating one included), 3 pairs blended, 4 pixels that
// terminated (T < 1e-4), 5 lane slots this kernel spent in its blend loop (per wave and batch,
// the list slots walked x 64), 6 the largest per-pixel evaluated count, 7 pixels in the image.
// The culled loop's batch of B splats in LDS: one 48-B record per slot (the tile-local quadratic of
// q, tileQuadratic, qmax and the colour; slot B is a null splat, q = -inf for every pixel, that
// pads the wave lists), the quadrant bits of each slot and each wave's list of record byte
// offsets (so a splat's reads need no address arithmetic).
template <int B>
struct CulledBatch {
    SplatRecord rec[B + 1];
    uint32_t mask[B];                    // bit w: the splat can reach wave w's quadrant
    __align__(16) uint32_t list[4][B + 4];
};
// The culled render's LDS: the batch, or, before the tile's first batch, the fused depth sort's
// words (kLdsChunk of them for the long form).
template <int B>
union CulledShared {
    CulledBatch<B> batch;
    uint64_t sortWords[kLdsChunk];
};

// The fused per-tile depth sort: the tile's range of the list by (depth bits, gaussian index), in
// place, by the whole workgroup (a range of at most 256 by wave 0 alone), then a barrier that
// hands the LDS to the render. Every thread of the workgroup calls it.
__device__ __forceinline__ void sortTileFused(const TileJob& job, uint2 range, uint32_t wave, uint32_t lane,
                                              uint64_t* sW) {
    const uint32_t n = range.y - range.x;
    if (n <= 1u) return;  // uniform
    uint32_t* values = job.sortValues;
    const uint32_t* depth = job.depthBits;
    if (n <= 256u) {
        if (wave == 0) {
            if (n <= 64u) waveSortRange<1>(values, depth, range.x, n, lane);
            else if (n <= 128u) waveSortRange<2>(values, depth, range.x, n, lane);
            else waveSortRange<4>(values, depth, range.x, n, lane);
        }
    } else if (n <= 512u) {
        groupSortRegisters<2>(values, depth, range, wave, lane, sW);
    } else if (n <= 1024u) {
        groupSortRegisters<4>(values, depth, range, wave, lane, sW);
    } else if (n <= 2048u) {
        groupSortRegisters<8>(values, depth, range, wave, lane, sW);
    } else if (n <= 4096u) {
        groupSortRegisters<16>(values, depth, range, wave, lane, sW);
    } else {
        groupSortLong(values, depth, job.sortScratch, range, sW);
    }
    __syncthreads();
}

// One 16x16 tile by the culled loop (the whole workgroup of 256 threads: it waits at barriers).
// FUSED: the code of the fused depth sort is compiled in (it runs when job.sortValues is set).
template <int FORMAT, int RS, bool STATS, int B, bool FUSED>
__device__ __forceinline__ void renderTileCulled(const TileJob& job, uint32_t tile, CulledShared<B>& shm) {
    static_assert(B % 64 == 0 && B >= 64 && B <= 512, "batches of 64..512 splats");
    constexpr uint32_t kFetch = (B + 255) / 256;  // slots each thread stages
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint2 range = job.ranges[tile];
    if (FUSED && job.sortValues) sortTileFused(job, range, wave, lane, shm.sortWords);  // uniform
    CulledBatch<B>& sh = shm.batch;
    const uint32_t tx0 = (tile % job.tilesX) * 16, ty0 = (tile / job.tilesX) * 16;
    const uint32_t px = tx0 + (wave & 1u) * 8 + (lane & 7u);
    const uint32_t py = ty0 + (wave >> 1) * 8 + (lane >> 3);
    const PixelQ pq = pixelQ(px - tx0, py - ty0);
    if (t == 0) {
        sh.rec[B].s0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        sh.rec[B].s1 = make_float2(0.0f, -__builtin_inff());
        sh.rec[B].qmax = __builtin_inff();
        sh.rec[B].col = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const char* recBase = reinterpret_cast<const char*>(sh.rec);

    float T = 1.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    float thr = (px < job.width && py < job.height) ? kLog2MinAlpha : __builtin_inff();
    const uint64_t laneMaskLt = (uint64_t(1) << lane) - 1u;

    uint32_t statFetched = 0, statTermEntry = 0xFFFFFFFFu, statBlended = 0, statSlots = 0;
    // Each thread's list values of the next batch, loaded while the current batch is walked (so a
    // batch's splat loads do not wait for its value loads).
    uint32_t nextValue[kFetch];
#pragma unroll
    for (uint32_t f = 0; f < kFetch; ++f) {
        const uint32_t e = range.x + t + f * 256u;
        nextValue[f] = e < range.y && (B % 256 == 0 || t + f * 256u < uint32_t(B)) ? job.values[e] : 0u;
    }
    for (uint32_t batch = range.x; batch < range.y; batch += B) {
        // Whole-tile vote; also the barrier that protects the previous batch's LDS.
        if (__syncthreads_count(thr < __builtin_inff() ? 0 : 1) == 256) break;
        if (STATS) statFetched += range.y - batch < uint32_t(B) ? range.y - batch : uint32_t(B);
        bool indefinite = false;
#pragma unroll
        for (uint32_t f = 0; f < kFetch; ++f) {
            const uint32_t slot = t + f * 256u;
            if (B % 256 != 0 && slot >= uint32_t(B)) break;
            const uint32_t e = batch + slot;
            uint32_t mask = 0u;
            if (e < range.y) {
                uint32_t g = nextValue[f];
                g = g < job.gaussianCount ? g : 0u;  // list values are gaussian indices; never read past the scene
                float cx, cy;
                float4 co, col;
                fetchSplat<RS>(job.splats, job.packed, g, cx, cy, co, col);
                sh.rec[slot].col = col;
                const float logO = __log2f(co.w);
                tileQuadratic(cx, cy, co, float(tx0), float(ty0), sh.rec[slot].s0, sh.rec[slot].s1);
                const bool ind = !conicPositiveDefinite(co.x, co.y, co.z);
                indefinite = indefinite || ind;
                sh.rec[slot].qmax = ind ? logO : __builtin_inff();
                // The quadrants the splat can reach (quadrantMask: box, then the exact
                // ellipse-vs-quadrant test, both with a margin sized to the rounding of q).
                const float r2 = 2.0f * (logO - kLog2MinAlpha) * (1.0f / kLog2e);
                mask = quadrantMask(cx, cy, co.x, co.y, co.z, r2, float(tx0), float(ty0));
            }
            sh.mask[slot] = mask;
        }
        // Also tells every wave whether the batch holds a splat with an indefinite conic (only
        // p24 records of thin splats): only such batches test power <= 0 per pair.
        const bool anyIndefinite = __syncthreads_or(indefinite ? 1 : 0) != 0;
        // This wave's list: the batch slots whose box meets its quadrant, in list order
        // (slot = j * 64 + lane), padded with the null slot to a multiple of 4.
        uint32_t count = 0;
#pragma unroll
        for (uint32_t j = 0; j < uint32_t(B) / 64u; ++j) {
            const uint32_t slot = j * 64 + lane;
            const bool hit = (sh.mask[slot] >> wave) & 1u;
            const uint64_t b = __builtin_amdgcn_ballot_w64(hit);
            if (hit) sh.list[wave][count + uint32_t(__popcll(b & laneMaskLt))] = slot * uint32_t(sizeof(SplatRecord));
            count += uint32_t(__popcll(b));
        }
        const uint32_t count4 = (count + 3u) & ~3u;
        if (lane < count4 - count) sh.list[wave][count + lane] = uint32_t(B) * uint32_t(sizeof(SplatRecord));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (uint32_t f = 0; f < kFetch; ++f) {
            const uint32_t e = batch + uint32_t(B) + t + f * 256u;
            nextValue[f] = e < range.y && (B % 256 == 0 || t + f * 256u < uint32_t(B)) ? job.values[e] : 0u;
        }
        if (__builtin_amdgcn_ballot_w64(thr < __builtin_inff()) != 0) {
            for (uint32_t j = 0; j < count4; j += 4) {
                const uint4 four = *reinterpret_cast<const uint4*>(&sh.list[wave][j]);
                // The four q first (their LDS reads in flight together), then the blends in order.
                const SplatRecord* rec[4] = {reinterpret_cast<const SplatRecord*>(recBase + four.x),
                                             reinterpret_cast<const SplatRecord*>(recBase + four.y),
                                             reinterpret_cast<const SplatRecord*>(recBase + four.z),
                                             reinterpret_cast<const SplatRecord*>(recBase + four.w)};
                float q[4];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) q[u] = evalQ(rec[u]->s0, rec[u]->s1, pq);
                if (anyIndefinite) {  // uniform: a batch with an indefinite conic drops power > 0 pairs
#pragma unroll
                    for (uint32_t u = 0; u < 4; ++u) q[u] = q[u] > rec[u]->qmax ? -__builtin_inff() : q[u];
                }
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    if (__builtin_amdgcn_ballot_w64(q[u] >= thr) == 0) continue;
                    const float thrBefore = thr;
                    blendPredicated(q[u], rec[u]->col, T, c0, c1, c2, thr);
                    if (STATS) {
                        const bool ended = thrBefore < __builtin_inff() && !(thr < __builtin_inff());
                        const uint32_t slot = uint32_t(reinterpret_cast<const char*>(rec[u]) - recBase) /
                                              uint32_t(sizeof(SplatRecord));
                        if (ended) statTermEntry = batch + slot;
                        statBlended += (q[u] >= thrBefore && !ended) ? 1u : 0u;
                    }
                }
                if (STATS) statSlots += 4u * 64u;
                if (__builtin_amdgcn_ballot_w64(thr < __builtin_inff()) == 0) break;
            }
        }
    }
    const bool inside = px < job.width && py < job.height;
    if (inside) {
        const size_t p = size_t(py) * job.width + px;
        if (FORMAT == GSR_TARGET_RGBA8) {
            const uint32_t rgba = quantize8(c0) | (quantize8(c1) << 8) | (quantize8(c2) << 16) |
                                  (quantize8(1.0f - T) << 24);
            static_cast<uint32_t*>(job.out)[p] = rgba;
        } else {
            static_cast<float4*>(job.out)[p] = make_float4(c0, c1, c2, 1.0f - T);
        }
    }
    if (STATS) {
        const bool terminated = inside && statTermEntry != 0xFFFFFFFFu;
        const uint32_t evaluated = !inside ? 0u : terminated ? statTermEntry - range.x + 1u : statFetched;
        const uint32_t sumEval = waveSum(evaluated), sumBlend = waveSum(inside ? statBlended : 0u);
        const uint32_t sumTerm = waveSum(terminated ? 1u : 0u), sumInside = waveSum(inside ? 1u : 0u);
        const uint32_t maxEval = waveMax(evaluated);
        uint32_t* st = job.tileStats + size_t(tile) * 8u;
        if (lane == 0) {
            atomicAdd(&st[2], sumEval);
            atomicAdd(&st[3], sumBlend);
            atomicAdd(&st[4], sumTerm);
            atomicAdd(&st[5], statSlots);
            atomicMax(&st[6], maxEval);
            atomicAdd(&st[7], sumInside);
        }
        if (t == 0) {
            st[0] = range.y - range.x;
            st[1] = statFetched;
        }
    }
}

// The culled loop over the tiles of several frames (a batch of views) in one persistent launch:
// each workgroup takes (view, tile) items from a device counter until all are done; item i
// belongs to the view v with tileStart[v] <= i < tileStart[v + 1]. Every view's lists must be
// complete (the launch follows every view's sort).
template <int FORMAT, int RS>
__global__ __launch_bounds__(256, 8) void kRenderViews(const TileJob* __restrict__ jobs,
                                                    const uint32_t* __restrict__ tileStart, uint32_t views,
                                                    uint32_t* __restrict__ counter) {
    __shared__ CulledShared<kBatch> sh;
    __shared__ uint32_t sItem;
    const uint32_t total = tileStart[views];
    for (;;) {
        if (threadIdx.x == 0) sItem = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t item = sItem;
        if (item >= total) return;  // uniform: every item is taken
        uint32_t lo = 0, hi = views;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tileStart[mid] <= item) lo = mid;
            else hi = mid;
        }
        renderTileCulled<FORMAT, RS, false, kBatch, false>(jobs[lo], item - tileStart[lo], sh);
        __syncthreads();  // every thread has read sItem and the batch before they are reused
    }
}

__global__ void kZeroCounter(uint32_t* counter) { *counter = 0u; }

// The culled loop over a frame's tiles: workgroup b takes tile rows of XCD b % 8 centre-out.
template <int FORMAT, int RS, bool STATS, int B>
__global__ __launch_bounds__(256, B == 512 ? 4 : 8) void kRenderCulled(TileJob job, uint32_t tileCount) {
    __shared__ CulledShared<B> sh;
    const uint32_t xcd = blockIdx.x % kXcds, idx = blockIdx.x / kXcds;
    const uint32_t row = xcd + kXcds * centreOut(idx / job.tilesX, gridDim.x / kXcds / job.tilesX);
    const uint32_t tile = row * job.tilesX + idx % job.tilesX;
    if (tile >= tileCount) return;  // uniform
    renderTileCulled<FORMAT, RS, STATS, B, true>(job, tile, sh);
}

// 8x8 bins (tile_size 8, tile_render.comp with an 8x8 workgroup): each wave renders one bin from
// the bin's own list, four bins per workgroup; the waves share nothing and never wait for each
// other (no workgroup barrier). Per batch of 64 entries the wave stages each entry's tile-local
// quadratic and colour as a 48-B LDS record, keeps the entries whose alpha >= 1/255 region can
// reach the bin (quadrantMask on the bin: the box, then the exact ellipse-vs-rectangle test) and
// walks them as the culled loop does, with the same early exit once all 64 pixels are done.
template <int FORMAT, int RS, int BW>
__global__ __launch_bounds__(256, BW == 128 ? 6 : 8) void kRenderBins8(const float4* __restrict__ splats,
                                                    const uint32_t* __restrict__ packed,
                                                    const uint32_t* __restrict__ values,
                                                    const uint2* __restrict__ ranges, uint32_t width, uint32_t height,
                                                    uint32_t binsX, uint32_t binCount, uint32_t gaussianCount,
                                                    void* __restrict__ out) {
    static_assert(BW == 32 || BW == 64 || BW == 128, "per-wave batches of 32, 64 or 128 splats");
    __shared__ SplatRecord sRec[4][BW + 1];               // per wave: BW batch slots + the null slot
    __shared__ __align__(16) uint32_t sList[4][BW + 4];  // per wave: record offsets in list order
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t bin = blockIdx.x * 4u + wave;
    if (bin >= binCount) return;  // uniform per wave
    const uint32_t bx0 = (bin % binsX) * 8, by0 = (bin / binsX) * 8;
    const uint32_t px = bx0 + (lane & 7u), py = by0 + (lane >> 3);
    const PixelQ pq = pixelQ(lane & 7u, lane >> 3);
    SplatRecord* rec = sRec[wave];
    if (lane == 0) {
        rec[BW].s0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rec[BW].s1 = make_float2(0.0f, -__builtin_inff());
        rec[BW].qmax = __builtin_inff();
        rec[BW].col = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const char* recBase = reinterpret_cast<const char*>(rec);
    float T = 1.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    float thr = (px < width && py < height) ? kLog2MinAlpha : __builtin_inff();
    const uint64_t laneMaskLt = (uint64_t(1) << lane) - 1u;

    const uint2 range = ranges[bin];
    for (uint32_t batch = range.x; batch < range.y; batch += BW) {
        if (__builtin_amdgcn_ballot_w64(thr < __builtin_inff()) == 0) break;  // the bin is done
        // The previous batch's records and list are no longer read (same wave, in order).
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        bool indefinite = false;
        uint32_t count = 0;
#pragma unroll
        for (uint32_t f = 0; f < uint32_t(BW + 63) / 64u; ++f) {
            const uint32_t slot = f * 64u + lane;
            const uint32_t e = batch + slot;
            bool hit = false;
            if (slot < uint32_t(BW) && e < range.y) {
                uint32_t g = values[e];
                g = g < gaussianCount ? g : 0u;  // list values are gaussian indices; never read past the scene
                float cx, cy;
                float4 co, col;
                fetchSplat<RS>(splats, packed, g, cx, cy, co, col);
                rec[slot].col = col;
                const float logO = __log2f(co.w);
                tileQuadratic(cx, cy, co, float(bx0), float(by0), rec[slot].s0, rec[slot].s1);
                const bool ind = !conicPositiveDefinite(co.x, co.y, co.z);
                indefinite = indefinite || ind;
                rec[slot].qmax = ind ? logO : __builtin_inff();
                const float r2 = 2.0f * (logO - kLog2MinAlpha) * (1.0f / kLog2e);
                hit = (quadrantMask(cx, cy, co.x, co.y, co.z, r2, float(bx0), float(by0)) & 1u) != 0u;
            }
            const uint64_t b = __builtin_amdgcn_ballot_w64(hit);
            if (hit) sList[wave][count + uint32_t(__popcll(b & laneMaskLt))] = slot * uint32_t(sizeof(SplatRecord));
            count += uint32_t(__popcll(b));
        }
        const uint32_t count4 = (count + 3u) & ~3u;
        if (lane < count4 - count) sList[wave][count + lane] = uint32_t(BW) * uint32_t(sizeof(SplatRecord));
        const bool anyIndefinite = __builtin_amdgcn_ballot_w64(indefinite) != 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t j = 0; j < count4; j += 4) {
            const uint4 four = *reinterpret_cast<const uint4*>(&sList[wave][j]);
            const SplatRecord* r[4] = {reinterpret_cast<const SplatRecord*>(recBase + four.x),
                                       reinterpret_cast<const SplatRecord*>(recBase + four.y),
                                       reinterpret_cast<const SplatRecord*>(recBase + four.z),
                                       reinterpret_cast<const SplatRecord*>(recBase + four.w)};
            float q[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) q[u] = evalQ(r[u]->s0, r[u]->s1, pq);
            if (anyIndefinite) {  // uniform: drop the power > 0 pairs of indefinite conics
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) q[u] = q[u] > r[u]->qmax ? -__builtin_inff() : q[u];
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                if (__builtin_amdgcn_ballot_w64(q[u] >= thr) == 0) continue;
                blendPredicated(q[u], r[u]->col, T, c0, c1, c2, thr);
            }
            if (__builtin_amdgcn_ballot_w64(thr < __builtin_inff()) == 0) break;
        }
    }
    if (px < width && py < height) {
        const size_t p = size_t(py) * width + px;
        if (FORMAT == GSR_TARGET_RGBA8) {
            const uint32_t rgba = quantize8(c0) | (quantize8(c1) << 8) | (quantize8(c2) << 16) |
                                  (quantize8(1.0f - T) << 24);
            static_cast<uint32_t*>(out)[p] = rgba;
        } else {
            static_cast<float4*>(out)[p] = make_float4(c0, c1, c2, 1.0f - T);
        }
    }
}

// StopThePop (tile_render.comp:201-279, GSR_STOPTHEPOP; CPU twin reference_renderer.cpp:143-196):
// the reference's alpha (expf, its operation order) and blend rule, applied in the order of a
// per-pixel window of K slots sorted by the depth of each splat's maximum along the pixel's ray,
// t* = (v . b) / (v^T P v), v = ((px - (W - 1) / 2) / fx, (py - (H - 1) / 2) / fy, 1). When the
// window is full the nearest of (window + new) is blended and leaves; at the end the window is
// blended nearest first. The window lives in registers (constant indices, K a template).
template <int K>
struct StpWindow {
    float depth[K], alpha[K], c0[K], c1[K], c2[K];
};

template <int K>
__device__ __forceinline__ void stpBlend(float a, float r, float g, float b, float& T, float& C0, float& C1, float& C2,
                                         bool& done) {
    const float testT = T * (1.0f - a);
    if (testT < kMinTransmittance) {
        done = true;
        return;
    }
    C0 += r * a * T;
    C1 += g * a * T;
    C2 += b * a * T;
    T = testT;
}

template <int K>
__device__ __forceinline__ void stpPush(StpWindow<K>& w, float depth, float a, float r, float g, float b, float& T,
                                        float& C0, float& C1, float& C2, bool& done) {
    const float kEmpty = __builtin_inff();
    if (!(depth < 3.0e38f)) depth = 3.0e38f;  // finite (NaN/Inf from a degenerate splat)
    if (w.depth[K - 1] == kEmpty) {  // not full: insertion from the back
        bool placed = false;
#pragma unroll
        for (int i = K - 1; i > 0; --i) {
            if (!placed) {
                if (w.depth[i - 1] > depth) {
                    w.depth[i] = w.depth[i - 1];
                    w.alpha[i] = w.alpha[i - 1];
                    w.c0[i] = w.c0[i - 1];
                    w.c1[i] = w.c1[i - 1];
                    w.c2[i] = w.c2[i - 1];
                } else {
                    w.depth[i] = depth;
                    w.alpha[i] = a;
                    w.c0[i] = r;
                    w.c1[i] = g;
                    w.c2[i] = b;
                    placed = true;
                }
            }
        }
        if (!placed) {
            w.depth[0] = depth;
            w.alpha[0] = a;
            w.c0[0] = r;
            w.c1[0] = g;
            w.c2[0] = b;
        }
        return;
    }
    if (depth < w.depth[0]) {  // full: the new splat is the nearest
        stpBlend<K>(a, r, g, b, T, C0, C1, C2, done);
        return;
    }
    stpBlend<K>(w.alpha[0], w.c0[0], w.c1[0], w.c2[0], T, C0, C1, C2, done);
    bool placed = false;
#pragma unroll
    for (int i = 0; i + 1 < K; ++i) {
        if (!placed) {
            if (w.depth[i + 1] <= depth) {
                w.depth[i] = w.depth[i + 1];
                w.alpha[i] = w.alpha[i + 1];
                w.c0[i] = w.c0[i + 1];
                w.c1[i] = w.c1[i + 1];
                w.c2[i] = w.c2[i + 1];
            } else {
                w.depth[i] = depth;
                w.alpha[i] = a;
                w.c0[i] = r;
                w.c1[i] = g;
                w.c2[i] = b;
                placed = true;
            }
        }
    }
    if (!placed) {
        w.depth[K - 1] = depth;
        w.alpha[K - 1] = a;
        w.c0[K - 1] = r;
        w.c1[K - 1] = g;
        w.c2[K - 1] = b;
    }
}

template <int FORMAT, int RS, int K>
__global__ __launch_bounds__(256) void kRenderStp(const float4* __restrict__ splats, const uint32_t* __restrict__ packed,
                                                  const float4* __restrict__ rayDepth,
                                                  const uint32_t* __restrict__ values, const uint2* __restrict__ ranges,
                                                  uint32_t width, uint32_t height, uint32_t tilesX, uint32_t tileCount,
                                                  uint32_t gaussianCount, float focalX, float focalY,
                                                  void* __restrict__ out) {
    __shared__ float4 sS0[kBatch];  // cx, cy, a, b
    __shared__ float2 sS1[kBatch];  // c, o
    __shared__ float4 sColor[kBatch];
    __shared__ float4 sRayA[kBatch];  // p00, p01, p02, p11
    __shared__ float4 sRayB[kBatch];  // p12, p22, b0, b1
    __shared__ float sRayC[kBatch];   // b2

    const uint32_t xcd = blockIdx.x % kXcds, idx = blockIdx.x / kXcds;
    const uint32_t row = xcd + kXcds * centreOut(idx / tilesX, gridDim.x / kXcds / tilesX);
    const uint32_t tile = row * tilesX + idx % tilesX;
    if (tile >= tileCount) return;  // uniform
    const uint32_t t = threadIdx.x;
    const uint32_t px = (tile % tilesX) * 16 + (t % 16);
    const uint32_t py = (tile / tilesX) * 16 + (t / 16);
    const float pxf = float(px), pyf = float(py);
    const float rx = (pxf - 0.5f * (float(width) - 1.0f)) / focalX;
    const float ry = (pyf - 0.5f * (float(height) - 1.0f)) / focalY;
    bool done = !(px < width && py < height);
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    StpWindow<K> w;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        w.depth[i] = __builtin_inff();
        w.alpha[i] = w.c0[i] = w.c1[i] = w.c2[i] = 0.0f;
    }

    const uint2 range = ranges[tile];
    for (uint32_t batch = range.x; batch < range.y; batch += kBatch) {
        if (__syncthreads_count(done ? 1 : 0) == 256) break;
        const uint32_t e = batch + t;
        if (e < range.y) {
            uint32_t g = values[e];
            g = g < gaussianCount ? g : 0u;  // list values are gaussian indices; never read past the scene
            float cx, cy;
            float4 co, col;
            fetchSplat<RS>(splats, packed, g, cx, cy, co, col);
            sS0[t] = make_float4(cx, cy, co.x, co.y);
            sS1[t] = make_float2(co.z, co.w);
            sColor[t] = col;
            const float4* rd = rayDepth + 3 * size_t(g);
            sRayA[t] = rd[0];
            sRayB[t] = rd[1];
            sRayC[t] = rd[2].x;
        }
        __syncthreads();
        const uint32_t n = range.y - batch < uint32_t(kBatch) ? range.y - batch : uint32_t(kBatch);
        for (uint32_t j = 0; !done && j < n; ++j) {
            const float4 s0 = sS0[j];
            const float2 s1 = sS1[j];
            const float dx = s0.x - pxf, dy = s0.y - pyf;
            const float power = -0.5f * (s0.z * dx * dx + s1.x * dy * dy) - s0.w * dx * dy;
            if (power > 0.0f) continue;
            const float ex = s1.y * expf(power);
            const float a = kMaxAlpha < ex ? kMaxAlpha : ex;
            if (a < kMinAlpha) continue;
            const float4 pa = sRayA[j], pb = sRayB[j];
            const float vPv = pa.x * rx * rx + 2.0f * pa.y * rx * ry + 2.0f * pa.z * rx + pa.w * ry * ry +
                              2.0f * pb.x * ry + pb.y;
            const float vb = pb.z * rx + pb.w * ry + sRayC[j];
            const float4 col = sColor[j];
            stpPush<K>(w, vb / vPv, a, col.x, col.y, col.z, T, C0, C1, C2, done);
        }
    }
#pragma unroll
    for (int i = 0; i < K; ++i)
        if (!done && w.depth[i] != __builtin_inff()) stpBlend<K>(w.alpha[i], w.c0[i], w.c1[i], w.c2[i], T, C0, C1, C2, done);
    if (px < width && py < height) {
        const size_t p = size_t(py) * width + px;
        if (FORMAT == GSR_TARGET_RGBA8) {
            const uint32_t rgba = quantize8(C0) | (quantize8(C1) << 8) | (quantize8(C2) << 16) |
                                  (quantize8(1.0f - T) << 24);
            static_cast<uint32_t*>(out)[p] = rgba;
        } else {
            static_cast<float4*>(out)[p] = make_float4(C0, C1, C2, 1.0f - T);
        }
    }
}

// StopThePop windows beyond the register form (K = 17..64; splat_tile_renderer.cpp:142 allows 64):
// the same blend order, with the window a binary min-heap per pixel in LDS. The register form
// keeps the window sorted by (ray depth, arrival) (an arrival is placed after the equal depths
// already there, tile_render.comp:221-279); a min-heap keyed by (ray depth, list position) pops in
// exactly that order, so the splats are blended in the same order: when the window is full the
// new splat is blended at once if it is strictly nearer than the heap's top, else the top is
// blended and replaced by the new splat; at the end the heap is emptied nearest first. A push or a
// pop costs O(log K) LDS moves instead of the O(K) of a sorted insertion, and the heap needs no
// registers per slot: one wave (64 threads) renders one 8x8 quadrant of a tile, and each of its
// pixels keeps K slots of 12 B (ordered depth, list position, gaussian) in LDS, slot-major so that
// the lanes' accesses to one slot index are conflict-free. Alpha and colour of a slot are fetched
// again from its gaussian when it is blended (the same expression on the same inputs).
__device__ __forceinline__ uint32_t depthOrder(float d) {
    const uint32_t u = __float_as_uint(d + 0.0f);  // -0 -> +0 (they compare equal)
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

template <int FORMAT, int RS>
__global__ __launch_bounds__(64) void kRenderStpHeap(const float4* __restrict__ splats,
                                                     const uint32_t* __restrict__ packed,
                                                     const float4* __restrict__ rayDepth,
                                                     const uint32_t* __restrict__ values,
                                                     const uint2* __restrict__ ranges, uint32_t width,
                                                     uint32_t height, uint32_t tilesX, uint32_t tileCount,
                                                     uint32_t gaussianCount, float focalX, float focalY, uint32_t K,
                                                     void* __restrict__ out) {
    extern __shared__ uint32_t sHeap[];  // [3][K][64]: ordered depth, list position, gaussian
    __shared__ float4 sS0[64];           // cx, cy, a, b
    __shared__ float2 sS1[64];           // c, o
    __shared__ float4 sColor[64];
    __shared__ float4 sRayA[64], sRayB[64];
    __shared__ float sRayC[64];
    __shared__ uint32_t sG[64];
    uint32_t* hKey = sHeap;
    uint32_t* hSeq = sHeap + K * 64u;
    uint32_t* hG = sHeap + 2u * K * 64u;

    const uint32_t tile = blockIdx.x >> 2, quadrant = blockIdx.x & 3u;
    if (tile >= tileCount) return;  // uniform
    const uint32_t lane = threadIdx.x;
    const uint32_t px = (tile % tilesX) * 16 + (quadrant & 1u) * 8 + (lane & 7u);
    const uint32_t py = (tile / tilesX) * 16 + (quadrant >> 1) * 8 + (lane >> 3);
    const float pxf = float(px), pyf = float(py);
    const float rx = (pxf - 0.5f * (float(width) - 1.0f)) / focalX;
    const float ry = (pyf - 0.5f * (float(height) - 1.0f)) / focalY;
    bool done = !(px < width && py < height);
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    uint32_t size = 0;

    // alpha and colour of gaussian g at this pixel (the push-time expression).
    auto alphaOf = [&](float cx, float cy, float4 co) {
        const float dx = cx - pxf, dy = cy - pyf;
        const float power = -0.5f * (co.x * dx * dx + co.z * dy * dy) - co.y * dx * dy;
        if (power > 0.0f) return 0.0f;
        const float ex = co.w * expf(power);
        const float a = kMaxAlpha < ex ? kMaxAlpha : ex;
        return a < kMinAlpha ? 0.0f : a;
    };
    auto blendG = [&](uint32_t g) {
        float cx, cy;
        float4 co, col;
        fetchSplat<RS>(splats, packed, g, cx, cy, co, col);
        stpBlend<1>(alphaOf(cx, cy, co), col.x, col.y, col.z, T, C0, C1, C2, done);
    };
    auto less = [](uint32_t ka, uint32_t sa, uint32_t kb, uint32_t sb) { return ka < kb || (ka == kb && sa < sb); };
    auto put = [&](uint32_t i, uint32_t k, uint32_t q, uint32_t g) {
        hKey[i * 64u + lane] = k;
        hSeq[i * 64u + lane] = q;
        hG[i * 64u + lane] = g;
    };
    // Sift (k, q, g) down from the root of a heap of `size` slots.
    auto siftDown = [&](uint32_t k, uint32_t q, uint32_t g) {
        uint32_t i = 0;
        for (;;) {
            uint32_t c = 2u * i + 1u;
            if (c >= size) break;
            uint32_t ck = hKey[c * 64u + lane], cq = hSeq[c * 64u + lane];
            if (c + 1u < size) {
                const uint32_t rk = hKey[(c + 1u) * 64u + lane], rq = hSeq[(c + 1u) * 64u + lane];
                if (less(rk, rq, ck, cq)) {
                    ++c;
                    ck = rk;
                    cq = rq;
                }
            }
            if (!less(ck, cq, k, q)) break;
            put(i, ck, cq, hG[c * 64u + lane]);
            i = c;
        }
        put(i, k, q, g);
    };

    const uint2 range = ranges[tile];
    for (uint32_t batch = range.x; batch < range.y; batch += 64u) {
        if (__builtin_amdgcn_ballot_w64(!done) == 0) break;  // uniform: the quadrant is done
        const uint32_t e = batch + lane;
        __syncthreads();  // the previous batch is no longer read
        if (e < range.y) {
            uint32_t g = values[e];
            g = g < gaussianCount ? g : 0u;  // list values are gaussian indices; never read past the scene
            float cx, cy;
            float4 co, col;
            fetchSplat<RS>(splats, packed, g, cx, cy, co, col);
            sS0[lane] = make_float4(cx, cy, co.x, co.y);
            sS1[lane] = make_float2(co.z, co.w);
            sColor[lane] = col;
            const float4* rd = rayDepth + 3 * size_t(g);
            sRayA[lane] = rd[0];
            sRayB[lane] = rd[1];
            sRayC[lane] = rd[2].x;
            sG[lane] = g;
        }
        __syncthreads();
        const uint32_t n = range.y - batch < 64u ? range.y - batch : 64u;
        for (uint32_t j = 0; !done && j < n; ++j) {
            const float4 s0 = sS0[j];
            const float2 s1 = sS1[j];
            const float a = alphaOf(s0.x, s0.y, make_float4(s0.z, s0.w, s1.x, s1.y));
            if (a == 0.0f) continue;
            const float4 pa = sRayA[j], pb = sRayB[j];
            const float vPv = pa.x * rx * rx + 2.0f * pa.y * rx * ry + 2.0f * pa.z * rx + pa.w * ry * ry +
                              2.0f * pb.x * ry + pb.y;
            const float vb = pb.z * rx + pb.w * ry + sRayC[j];
            float depth = vb / vPv;
            if (!(depth < 3.0e38f)) depth = 3.0e38f;  // finite (NaN/Inf from a degenerate splat)
            const uint32_t key = depthOrder(depth), seq = batch + j, g = sG[j];
            if (size < K) {  // not full: push (sift up)
                uint32_t i = size++;
                while (i > 0u) {
                    const uint32_t p = (i - 1u) >> 1;
                    const uint32_t pk = hKey[p * 64u + lane], pq = hSeq[p * 64u + lane];
                    if (!less(key, seq, pk, pq)) break;
                    put(i, pk, pq, hG[p * 64u + lane]);
                    i = p;
                }
                put(i, key, seq, g);
                continue;
            }
            if (less(key, seq, hKey[lane], hSeq[lane])) {  // nearer than the whole window
                const float4 col = sColor[j];
                stpBlend<1>(a, col.x, col.y, col.z, T, C0, C1, C2, done);
                continue;
            }
            blendG(hG[lane]);  // the window's nearest leaves, the new splat takes its place
            siftDown(key, seq, g);
        }
    }
    while (size > 0u && !done) {  // the rest, nearest first
        blendG(hG[lane]);
        --size;
        siftDown(hKey[size * 64u + lane], hSeq[size * 64u + lane], hG[size * 64u + lane]);
    }
    if (px < width && py < height) {
        const size_t p = size_t(py) * width + px;
        if (FORMAT == GSR_TARGET_RGBA8) {
            const uint32_t rgba = quantize8(C0) | (quantize8(C1) << 8) | (quantize8(C2) << 16) |
                                  (quantize8(1.0f - T) << 24);
            static_cast<uint32_t*>(out)[p] = rgba;
        } else {
            static_cast<float4*>(out)[p] = make_float4(C0, C1, C2, 1.0f - T);
        }
    }
}

template <int FORMAT, int RS, int K>
void launchStpK(const float4* splats, const uint32_t* packed, const float4* rayDepth, const uint32_t* values,
                const uint2* ranges, const DevCamera& cam, uint32_t gaussianCount, void* out, hipStream_t stream) {
    const uint32_t tileCount = cam.tilesX * cam.tilesY;
    const uint32_t rowsPerXcd = (cam.tilesY + kXcds - 1) / kXcds;
    hipLaunchKernelGGL((kRenderStp<FORMAT, RS, K>), dim3(rowsPerXcd * cam.tilesX * kXcds), dim3(256), 0, stream, splats,
                       packed, rayDepth, values, ranges, cam.width, cam.height, cam.tilesX, tileCount, gaussianCount,
                       cam.focalX, cam.focalY, out);
    markKernel("kRenderStp", stream);
}

template <int FORMAT, int RS, int K = 1>
void launchStp(int window, const float4* splats, const uint32_t* packed, const float4* rayDepth,
               const uint32_t* values, const uint2* ranges, const DevCamera& cam, uint32_t gaussianCount, void* out,
               hipStream_t stream) {
    if (K == 1 && window > kMaxSortWindow) {  // the LDS heap form
        const uint32_t tileCount = cam.tilesX * cam.tilesY;
        const size_t lds = size_t(window) * 64u * 12u;
        hipLaunchKernelGGL((kRenderStpHeap<FORMAT, RS>), dim3(tileCount * 4u), dim3(64), lds, stream, splats, packed,
                           rayDepth, values, ranges, cam.width, cam.height, cam.tilesX, tileCount, gaussianCount,
                           cam.focalX, cam.focalY, uint32_t(window), out);
        markKernel("kRenderStpHeap", stream);
        return;
    }
    if constexpr (K <= kMaxSortWindow) {
        if (window == K)
            launchStpK<FORMAT, RS, K>(splats, packed, rayDepth, values, ranges, cam, gaussianCount, out, stream);
        else
            launchStp<FORMAT, RS, K + 1>(window, splats, packed, rayDepth, values, ranges, cam, gaussianCount, out,
                                         stream);
    }
}

template <int FORMAT, int RS>
void launchCulled(const float4* splats, const uint32_t* packed, const uint32_t* values, const uint2* ranges,
                  const DevCamera& cam, uint32_t gaussianCount, void* out, uint32_t* tileStats, int batch,
                  hipStream_t stream, const FusedSort* fused) {
    const uint32_t tileCount = cam.tilesX * cam.tilesY;
    const uint32_t rowsPerXcd = (cam.tilesY + kXcds - 1) / kXcds;
    const dim3 grid(rowsPerXcd * cam.tilesX * kXcds);
    TileJob job{splats, packed, values, ranges, cam.width, cam.height, cam.tilesX, gaussianCount, out, tileStats};
    if (fused) {
        job.sortValues = fused->values;
        job.depthBits = fused->depthBits;
        job.sortScratch = fused->scratch;
    }
    if (tileStats)  // the statistics build: the default batch only
        hipLaunchKernelGGL((kRenderCulled<FORMAT, RS, true, kBatch>), grid, dim3(256), 0, stream, job, tileCount);
    else if (batch == 128)
        hipLaunchKernelGGL((kRenderCulled<FORMAT, RS, false, 128>), grid, dim3(256), 0, stream, job, tileCount);
    else if (batch == 512)
        hipLaunchKernelGGL((kRenderCulled<FORMAT, RS, false, 512>), grid, dim3(256), 0, stream, job, tileCount);
    else
        hipLaunchKernelGGL((kRenderCulled<FORMAT, RS, false, kBatch>), grid, dim3(256), 0, stream, job, tileCount);
    markKernel("kRenderCulled", stream);
}

template <int FORMAT, bool FAST, int PPT, bool PRED, int RS>
void launchOne(const float4* splats, const uint32_t* packed, const uint32_t* values, const uint2* ranges,
               const DevCamera& cam, uint32_t gaussianCount, void* out, hipStream_t stream) {
    const uint32_t tileCount = cam.tilesX * cam.tilesY;
    const uint32_t rowsPerXcd = (cam.tilesY + kXcds - 1) / kXcds;
    const uint32_t perXcd = rowsPerXcd * cam.tilesX;
    hipLaunchKernelGGL((kRender<FORMAT, FAST, PPT, PRED, RS>), dim3(perXcd * kXcds), dim3(256 / PPT), 0, stream,
                       splats, packed, values, ranges, cam.width, cam.height, cam.tilesX, tileCount, gaussianCount,
                       out);
    markKernel("kRender", stream);
}

template <int FORMAT, int RS>
void launchFormat(int ppt, bool exact, int loop, const float4* splats, const uint32_t* packed, const uint32_t* values,
                  const uint2* ranges, const DevCamera& cam, uint32_t gaussianCount, void* out, uint32_t* tileStats,
                  int batch, hipStream_t stream, const FusedSort* fused) {
    if (exact) {
        if (ppt == 4)
            launchOne<FORMAT, false, 4, false, RS>(splats, packed, values, ranges, cam, gaussianCount, out, stream);
        else if (ppt == 2)
            launchOne<FORMAT, false, 2, false, RS>(splats, packed, values, ranges, cam, gaussianCount, out, stream);
        else
            launchOne<FORMAT, false, 1, false, RS>(splats, packed, values, ranges, cam, gaussianCount, out, stream);
    } else if (ppt == 4) {
        launchOne<FORMAT, true, 4, false, RS>(splats, packed, values, ranges, cam, gaussianCount, out, stream);
    } else if (ppt == 2) {
        launchOne<FORMAT, true, 2, false, RS>(splats, packed, values, ranges, cam, gaussianCount, out, stream);
    } else if (loop == GSR_RENDER_LOOP_DIVERGENT) {
        launchOne<FORMAT, true, 1, false, RS>(splats, packed, values, ranges, cam, gaussianCount, out, stream);
    } else if (loop == GSR_RENDER_LOOP_PREDICATED) {
        launchOne<FORMAT, true, 1, true, RS>(splats, packed, values, ranges, cam, gaussianCount, out, stream);
    } else {
        launchCulled<FORMAT, RS>(splats, packed, values, ranges, cam, gaussianCount, out, tileStats, batch, stream,
                                 fused);
    }
}

template <int FORMAT>
void launchSplat(int renderSplat, int ppt, bool exact, int loop, int sortWindow, const float4* rayDepth,
                 const float4* splats, const uint32_t* packed, const uint32_t* values, const uint2* ranges,
                 const DevCamera& cam, uint32_t gaussianCount, void* out, uint32_t* tileStats, int batch,
                 hipStream_t stream, const FusedSort* fused) {
    if (sortWindow > 0) {
        if (renderSplat == GSR_RENDER_SPLAT_P32)
            launchStp<FORMAT, GSR_RENDER_SPLAT_P32>(sortWindow, splats, packed, rayDepth, values, ranges, cam,
                                                    gaussianCount, out, stream);
        else if (renderSplat == GSR_RENDER_SPLAT_P24)
            launchStp<FORMAT, GSR_RENDER_SPLAT_P24>(sortWindow, splats, packed, rayDepth, values, ranges, cam,
                                                    gaussianCount, out, stream);
        else
            launchStp<FORMAT, GSR_RENDER_SPLAT_FULL>(sortWindow, splats, packed, rayDepth, values, ranges, cam,
                                                     gaussianCount, out, stream);
        return;
    }
    if (renderSplat == GSR_RENDER_SPLAT_P32)
        launchFormat<FORMAT, GSR_RENDER_SPLAT_P32>(ppt, exact, loop, splats, packed, values, ranges, cam, gaussianCount,
                                                   out, tileStats, batch, stream, fused);
    else if (renderSplat == GSR_RENDER_SPLAT_P24)
        launchFormat<FORMAT, GSR_RENDER_SPLAT_P24>(ppt, exact, loop, splats, packed, values, ranges, cam, gaussianCount,
                                                   out, tileStats, batch, stream, fused);
    else
        launchFormat<FORMAT, GSR_RENDER_SPLAT_FULL>(ppt, exact, loop, splats, packed, values, ranges, cam,
                                                    gaussianCount, out, tileStats, batch, stream, fused);
}

}  // namespace

void launchRenderViews(const TileJob* jobs, const uint32_t* tileStart, uint32_t views, uint32_t workgroups,
                       uint32_t* counter, gsr_target_format format, int renderSplat, hipStream_t stream) {
    if (views == 0 || workgroups == 0) return;
    hipLaunchKernelGGL(kZeroCounter, dim3(1), dim3(1), 0, stream, counter);
    markKernel("kZeroCounter", stream);
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(workgroups), dim3(256), 0, stream, jobs, tileStart, views, counter);
    };
    const bool f8 = format == GSR_TARGET_RGBA8;
    if (renderSplat == GSR_RENDER_SPLAT_P32)
        f8 ? go(kRenderViews<GSR_TARGET_RGBA8, GSR_RENDER_SPLAT_P32>) : go(kRenderViews<GSR_TARGET_RGBA32F, GSR_RENDER_SPLAT_P32>);
    else if (renderSplat == GSR_RENDER_SPLAT_P24)
        f8 ? go(kRenderViews<GSR_TARGET_RGBA8, GSR_RENDER_SPLAT_P24>) : go(kRenderViews<GSR_TARGET_RGBA32F, GSR_RENDER_SPLAT_P24>);
    else
        f8 ? go(kRenderViews<GSR_TARGET_RGBA8, GSR_RENDER_SPLAT_FULL>)
           : go(kRenderViews<GSR_TARGET_RGBA32F, GSR_RENDER_SPLAT_FULL>);
    markKernel("kRenderViews", stream);
}

void launchRender(const float4* splats, const uint32_t* packed, int renderSplat, const uint32_t* values,
                  const uint2* ranges, const DevCamera& cam, uint32_t gaussianCount, gsr_target_format format,
                  int pixelsPerThread, bool exactAlpha, int loop, int sortWindow, const float4* rayDepth, void* out,
                  uint32_t* tileStats, int batch, hipStream_t stream, const FusedSort* fused) {
    if (cam.tilesX * cam.tilesY == 0) return;
    if (cam.binSize == 8) {  // 8x8 bins: the culled loop only (the renderer refuses the other forms)
        const uint32_t bins = cam.binsX * cam.binsY;
        const dim3 grid((bins + 3) / 4);
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, dim3(256), 0, stream, splats, packed, values, ranges, cam.width,
                               cam.height, cam.binsX, bins, gaussianCount, out);
        };
        auto pick = [&](auto rs) {
            constexpr int RS = decltype(rs)::value;
            const bool f8 = format == GSR_TARGET_RGBA8;
            if (batch == 32) f8 ? go(kRenderBins8<GSR_TARGET_RGBA8, RS, 32>) : go(kRenderBins8<GSR_TARGET_RGBA32F, RS, 32>);
            else if (batch == 128)
                f8 ? go(kRenderBins8<GSR_TARGET_RGBA8, RS, 128>) : go(kRenderBins8<GSR_TARGET_RGBA32F, RS, 128>);
            else f8 ? go(kRenderBins8<GSR_TARGET_RGBA8, RS, 64>) : go(kRenderBins8<GSR_TARGET_RGBA32F, RS, 64>);
        };
        if (renderSplat == GSR_RENDER_SPLAT_P32) pick(std::integral_constant<int, GSR_RENDER_SPLAT_P32>{});
        else if (renderSplat == GSR_RENDER_SPLAT_P24) pick(std::integral_constant<int, GSR_RENDER_SPLAT_P24>{});
        else pick(std::integral_constant<int, GSR_RENDER_SPLAT_FULL>{});
        markKernel("kRenderBins8", stream);
        return;
    }
    if (format == GSR_TARGET_RGBA8)
        launchSplat<GSR_TARGET_RGBA8>(renderSplat, pixelsPerThread, exactAlpha, loop, sortWindow, rayDepth, splats,
                                      packed, values, ranges, cam, gaussianCount, out, tileStats, batch, stream, fused);
    else
        launchSplat<GSR_TARGET_RGBA32F>(renderSplat, pixelsPerThread, exactAlpha, loop, sortWindow, rayDepth, splats,
                                        packed, values, ranges, cam, gaussianCount, out, tileStats, batch, stream,
                                        fused);
}

}  // namespace gsr
