// K1 splat_preprocess (reference shaders/splat_preprocess.comp:273-355 + 411-450, the GSR_TILES
// build; CPU twin splat/splat_math.cpp:153-193): one thread per gaussian.
//
//   - near cull !(z > 0.2) (NaN-safe), opacity cull o < 1/255, EWA covariance, conic, radius,
//     centre, tile rectangle (off-screen cull), SH colour;
//   - the 48-byte Splat of every gaussian (culled ones carry depth + reason) and its tile
//     count: the rectangle's area (rect) or the tiles of the alpha >= 1/255 ellipse (tight);
//   - fused reduce of the reduce-then-scan: each 256-thread block writes one BlockStats
//     (sum of its tile counts, non-zero counts, per-reason cull counts), so there are no
//     global atomics and the scan reads 32 B per block instead of 4 B per gaussian;
//   - zeroes the per-tile ranges for the ranges pass later in the frame.
#include "gsr_device.hpp"

namespace gsr {

namespace {

template <int FMT>
__global__ __launch_bounds__(kPreprocessThreads) void kPreprocess(SceneRef scene, DevCamera cam, uint32_t bounds,
                                                                  float4* __restrict__ splats,
                                                                  uint32_t* __restrict__ tileCounts,
                                                                  BlockStats* __restrict__ blockStats,
                                                                  BlockStats* __restrict__ superStats,
                                                                  gsr_control* __restrict__ control,
                                                                  uint32_t* __restrict__ spanInfo,
                                                                  uint4* __restrict__ spanPlanes,
                                                                  uint32_t* __restrict__ depthBits,
                                                                  uint32_t* __restrict__ rangesWords,
                                                                  uint32_t rangeWordCount, uint32_t renderSplat,
                                                                  uint32_t* __restrict__ packed,
                                                                  float4* __restrict__ rayDepth,
                                                                  const uint8_t* __restrict__ useMlp) {
    const uint32_t i = blockIdx.x * kPreprocessThreads + threadIdx.x;
    if (i < rangeWordCount) rangesWords[i] = 0u;
    if (superStats && i == 0) {  // the counters the frame's passes accumulate into
        control->coverage_mismatch = 0u;
        control->sort_timeouts = 0u;
    }

    uint32_t reason = kNoGaussian;
    uint32_t touched = 0;
    if (i < scene.count) {
        float px, py, pz, alpha;
        uint4 g0, g1;
        // f32 scene: the scales and rotation are loaded with the position (one memory round trip
        // for the geometry); the SH quads only once the splat is on screen.
        float4 scF, rotF;
        if (FMT == GSR_SCENE_F32_SOA) {
            scF = scene.scales[i];
            rotF = scene.rotations[i];
            const float4 pa = scene.posAlpha[i];
            px = pa.x;
            py = pa.y;
            pz = pa.z;
            alpha = pa.w;
        } else {
            g0 = scene.geometry[i];
            g1 = scene.geometry[size_t(scene.count) + i];
            px = __uint_as_float(g0.x);
            py = __uint_as_float(g0.y);
            pz = __uint_as_float(g0.z);
            alpha = halfBitsToFloat(g0.w & 0xFFFFu);
        }
        const float* v = cam.view;
        // transformPoint (splat_math.h:65-69), per component, no contraction.
        const float t0 = v[0] * px + v[4] * py + v[8] * pz + v[12];
        const float t1 = v[1] * px + v[5] * py + v[9] * pz + v[13];
        const float t2 = v[2] * px + v[6] * py + v[10] * pz + v[14];

        float4 o0 = make_float4(0.0f, 0.0f, t2, 0.0f);
        float4 o1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float4 o2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (!(t2 > kNearCullDepth)) {
            reason = GSR_CULL_NEAR_PLANE;
        } else if (alpha < kMinAlpha) {
            reason = GSR_CULL_LOW_OPACITY;
        } else {
            float sx, sy, sz, qw, qx, qy, qz;
            if (FMT == GSR_SCENE_F32_SOA) {
                const float4 sc = scF;
                const float4 q = rotF;
                sx = sc.x;
                sy = sc.y;
                sz = sc.z;
                qw = q.x;
                qx = q.y;
                qy = q.z;
                qz = q.w;
            } else {
                sx = halfBitsToFloat(g0.w >> 16);
                sy = halfBitsToFloat(g1.x & 0xFFFFu);
                sz = halfBitsToFloat(g1.x >> 16);
                const float hw = halfBitsToFloat(g1.y & 0xFFFFu), hx = halfBitsToFloat(g1.y >> 16);
                const float hy = halfBitsToFloat(g1.z & 0xFFFFu), hz = halfBitsToFloat(g1.z >> 16);
                // decodeGaussian (scene_compression.cpp:408-415): q * (1 / sqrt(dot(q, q))).
                const float len2 = hw * hw + hx * hx + hy * hy + hz * hz;
                if (len2 > 0.0f) {
                    const float inv = 1.0f / sqrtf(len2);
                    qw = hw * inv;
                    qx = hx * inv;
                    qy = hy * inv;
                    qz = hz * inv;
                } else {
                    qw = 1.0f;
                    qx = qy = qz = 0.0f;
                }
            }
            // quaternionToMatrix + computeCov3D (splat_math.cpp:44-63), R[col][row].
            const float R00 = 1.0f - 2.0f * (qy * qy + qz * qz), R01 = 2.0f * (qx * qy + qw * qz),
                        R02 = 2.0f * (qx * qz - qw * qy);
            const float R10 = 2.0f * (qx * qy - qw * qz), R11 = 1.0f - 2.0f * (qx * qx + qz * qz),
                        R12 = 2.0f * (qy * qz + qw * qx);
            const float R20 = 2.0f * (qx * qz + qw * qy), R21 = 2.0f * (qy * qz - qw * qx),
                        R22 = 1.0f - 2.0f * (qx * qx + qy * qy);
            const float M[3][3] = {{R00 * sx, R01 * sx, R02 * sx}, {R10 * sy, R11 * sy, R12 * sy},
                                   {R20 * sz, R21 * sz, R22 * sz}};
            float S[3][3];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k) S[j][k] = M[0][k] * M[0][j] + M[1][k] * M[1][j] + M[2][k] * M[2][j];
            // computeCov2D (splat_math.cpp:65-91).
            const float limX = kFrustumClampFactor * cam.tanX;
            const float limY = kFrustumClampFactor * cam.tanY;
            const float txtz = t0 / t2;
            const float tytz = t1 / t2;
            const float cx = txtz < -limX ? -limX : txtz;
            const float cy = tytz < -limY ? -limY : tytz;
            const float tx = (limX < cx ? limX : cx) * t2;
            const float ty = (limY < cy ? limY : cy) * t2;
            const float tz = t2;
            const float j00 = cam.focalX / tz;
            const float j20 = -(cam.focalX * tx) / (tz * tz);
            const float j11 = cam.focalY / tz;
            const float j21 = -(cam.focalY * ty) / (tz * tz);
            float T[3][2];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                T[j][0] = j00 * v[j * 4 + 0] + j20 * v[j * 4 + 2];
                T[j][1] = j11 * v[j * 4 + 1] + j21 * v[j * 4 + 2];
            }
            float U[3][2];
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 2; ++k) U[j][k] = T[0][k] * S[j][0] + T[1][k] * S[j][1] + T[2][k] * S[j][2];
            const float c00 = U[0][0] * T[0][0] + U[1][0] * T[1][0] + U[2][0] * T[2][0];
            const float c01 = U[0][1] * T[0][0] + U[1][1] * T[1][0] + U[2][1] * T[2][0];
            const float c11 = U[0][1] * T[0][1] + U[1][1] * T[1][1] + U[2][1] * T[2][1];
            const float ca = c00 + kLowPassFilter, cb = c01, cc = c11 + kLowPassFilter;
            // conicAndRadius (splat_math.cpp:102-114).
            const float det = ca * cc - cb * cb;
            if (det == 0.0f) {
                reason = GSR_CULL_SINGULAR_COV;
            } else {
                const float detInv = 1.0f / det;
                const float ka = cc * detInv, kb = -cb * detInv, kc = ca * detInv;
                const float mid = 0.5f * (ca + cc);
                const float disc = mid * mid - det;
                const float root = sqrtf(disc > 0.1f ? disc : 0.1f);
                const float l1 = mid + root, l2 = mid - root;
                const float radius = ceilf(3.0f * sqrtf(l1 > l2 ? l1 : l2));
                // projectToPixel (splat_math.cpp:93-100).
                const float ndcX = t0 / (t2 * cam.tanX);
                const float ndcY = t1 / (t2 * cam.tanY);
                const float ctrX = ((ndcX + 1.0f) * float(cam.width) - 1.0f) * 0.5f;
                const float ctrY = ((ndcY + 1.0f) * float(cam.height) - 1.0f) * 0.5f;
                const Rect rect = tileRect(ctrX, ctrY, radius, cam.tilesX, cam.tilesY);
                o0.x = ctrX;
                o0.y = ctrY;
                if (rect.x1 <= rect.x0 || rect.y1 <= rect.y0) {
                    reason = GSR_CULL_OFF_SCREEN;
                } else {
                    if (useMlp && useMlp[i]) {
                        o2.x = o2.y = o2.z = 0.0f;  // the learned colour pass writes it
                    } else {
                        // SH colour, dir = normalize(mean - camPos) in data space.
                        const float dx = px - cam.pos[0], dy = py - cam.pos[1], dz = pz - cam.pos[2];
                        const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
                        float rgb[3];
                        evalShStreamed<FMT>(scene, i, scene.degree, dx * inv, dy * inv, dz * inv, rgb);
                        o2.x = rgb[0];
                        o2.y = rgb[1];
                        o2.z = rgb[2];
                    }
                    o2.w = radius;
                    // p24 keeps the conic and opacity in fp16: the tiles are counted, duplicated
                    // and rendered with the values the record gives back.
                    float ra = ka, rb = kb, rc = kc, ro = alpha;
                    if (renderSplat == GSR_RENDER_SPLAT_P24) {
                        ra = halfBitsToFloat(halfBits(ka * kConicScaleP24)) * kInvConicScaleP24;
                        rb = halfBitsToFloat(halfBits(kb * kConicScaleP24)) * kInvConicScaleP24;
                        rc = halfBitsToFloat(halfBits(kc * kConicScaleP24)) * kInvConicScaleP24;
                        ro = halfBitsToFloat(halfBits(alpha));
                    }
                    o1 = make_float4(ra, rb, rc, ro);
                    reason = GSR_CULL_VISIBLE;
                    // The bins the duplicate pass will write (splat_preprocess.comp:291-308): INRIA's
                    // rectangle count on 16-px bins with rect bounds, else the coverage count.
                    SpanRecord spans;
                    if (bounds == GSR_BOUNDS_RECT && cam.binSize == uint32_t(kTileSize)) {
                        touched = uint32_t((rect.x1 - rect.x0) * (rect.y1 - rect.y0));
                        spans.info = uint32_t(rect.y0) |
                                     (uint32_t(uint16_t(int16_t(-(rect.y1 - rect.y0) - 1))) << 16);
#pragma unroll
                        for (int k = 0; k < kSpanRows; ++k) spans.span[k] = 0u;
                        spans.span[0] = uint32_t(rect.x0) | (uint32_t(rect.x1 - rect.x0) << 16);
                    } else {
                        const Coverage cv = makeCoverage(bounds, rect, int(cam.binSize), int(cam.binsX),
                                                         int(cam.binsY), ctrX, ctrY, ra, rb, rc, ro);
                        touched = coverageCountSpans(cv, int(cam.binSize), spans);
                    }
                    if (touched != 0u) {  // the duplicate pass reads them for its listed gaussians
                        if (depthBits) depthBits[i] = __float_as_uint(o0.z);  // the list's depth keys
                        spanInfo[i] = spans.info;
                        spanPlanes[i] = make_uint4(spans.span[0], spans.span[1], spans.span[2], spans.span[3]);
                        if (int(int16_t(uint16_t(spans.info >> 16))) > kSpanRows / 2)  // the second plane if used
                            spanPlanes[size_t(scene.count) + i] =
                                make_uint4(spans.span[4], spans.span[5], spans.span[6], spans.span[7]);
                    }
                    if (rayDepth) {
                        // StopThePop (computeRayDepthParams, splat_math.cpp:195-205): P = Sigma^-1 in
                        // camera space = sum_k w_k M_k M_k^T with M = mat3(view) R, w = 1 / s^2,
                        // s = max(sigma, max(1e-4 smax, 1e-12)); b = P t.
                        const float Rm[3][3] = {{R00, R01, R02}, {R10, R11, R12}, {R20, R21, R22}};
                        float Mc[3][3];
#pragma unroll
                        for (int c = 0; c < 3; ++c)
#pragma unroll
                            for (int r = 0; r < 3; ++r)
                                Mc[c][r] = v[0 * 4 + r] * Rm[c][0] + v[1 * 4 + r] * Rm[c][1] + v[2 * 4 + r] * Rm[c][2];
                        const float smax = selMax(sx, selMax(sy, sz));
                        const float floorS = selMax(1e-4f * smax, 1e-12f);
                        const float sig[3] = {sx, sy, sz};
                        float w[3];
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const float sk = sig[k] < floorS ? floorS : sig[k];
                            w[k] = 1.0f / (sk * sk);
                        }
                        float P[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
#pragma unroll
                        for (int k = 0; k < 3; ++k)
#pragma unroll
                            for (int a = 0; a < 3; ++a)
#pragma unroll
                                for (int b = 0; b < 3; ++b) P[a][b] = P[a][b] + w[k] * (Mc[k][b] * Mc[k][a]);
                        const float tc[3] = {t0, t1, t2};
                        float bv[3];
#pragma unroll
                        for (int r = 0; r < 3; ++r) bv[r] = P[0][r] * tc[0] + P[1][r] * tc[1] + P[2][r] * tc[2];
                        float4* rd = rayDepth + 3 * size_t(i);
                        rd[0] = make_float4(P[0][0], P[0][1], P[0][2], P[1][1]);
                        rd[1] = make_float4(P[1][2], P[2][2], bv[0], bv[1]);
                        rd[2] = make_float4(bv[2], 0.0f, 0.0f, 0.0f);
                    }
                    if (renderSplat != GSR_RENDER_SPLAT_FULL) {
                        const uint32_t rad = uint32_t(radius < 65535.0f ? radius : 65535.0f);
                        const uint32_t rg = halfBits(o2.x) | (halfBits(o2.y) << 16);
                        const uint32_t br = halfBits(o2.z) | (rad << 16);
                        if (renderSplat == GSR_RENDER_SPLAT_P32) {
                            uint4* w = reinterpret_cast<uint4*>(packed) + 2 * size_t(i);
                            w[0] = make_uint4(__float_as_uint(ctrX), __float_as_uint(ctrY), __float_as_uint(ka),
                                              __float_as_uint(kb));
                            w[1] = make_uint4(__float_as_uint(kc), __float_as_uint(alpha), rg, br);
                        } else {
                            uint2* w = reinterpret_cast<uint2*>(packed) + 3 * size_t(i);
                            w[0] = make_uint2(__float_as_uint(ctrX), __float_as_uint(ctrY));
                            w[1] = make_uint2(halfBits(ka * kConicScaleP24) | (halfBits(kb * kConicScaleP24) << 16),
                                              halfBits(kc * kConicScaleP24) | (halfBits(alpha) << 16));
                            w[2] = make_uint2(rg, br);
                        }
                    }
                }
            }
        }
        o0.w = __uint_as_float(reason);
        float4* dst = splats + 3 * size_t(i);
        dst[0] = o0;
        dst[1] = o1;
        dst[2] = o2;
        tileCounts[i] = touched;
    }

    // Block reduction: wave ballots for the cull reasons, shuffles for the tile sum.
    __shared__ uint32_t sStats[kPreprocessThreads / 64][8];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t sum = touched;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    const uint32_t nz = uint32_t(__popcll(__ballot(touched != 0u)));
    uint32_t cull[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) cull[r] = uint32_t(__popcll(__ballot(reason == uint32_t(r))));
    if (lane == 0) {
        sStats[wave][0] = sum;
        sStats[wave][1] = nz;
#pragma unroll
        for (int r = 0; r < 5; ++r) sStats[wave][2 + r] = cull[r];
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < kPreprocessThreads / 64; ++w) acc += sStats[w][threadIdx.x];
        reinterpret_cast<uint32_t*>(&blockStats[blockIdx.x])[threadIdx.x] = acc;
        if (superStats && acc != 0u)  // global mode: the superblock's sums (the duplicate's block offsets)
            atomicAdd(reinterpret_cast<uint32_t*>(&superStats[blockIdx.x / kSuperBlocks]) + threadIdx.x, acc);
    }
    if (threadIdx.x == 7) blockStats[blockIdx.x].pad = 0u;
}

}  // namespace

void launchPreprocess(const SceneRef& scene, const DevCamera& cam, uint32_t bounds, float4* splats,
                      uint32_t* tileCounts, BlockStats* blockStats, BlockStats* superStats, gsr_control* control,
                      uint32_t* spanInfo, uint4* spanPlanes, uint32_t* depthBits, uint32_t* rangesWords,
                      uint32_t rangeWordCount, uint32_t renderSplat, uint32_t* packed, float4* rayDepth,
                      const uint8_t* useMlp, hipStream_t stream) {
    uint32_t n = scene.count > rangeWordCount ? scene.count : rangeWordCount;
    const uint32_t blocks = (n + kPreprocessThreads - 1) / kPreprocessThreads;
    if (blocks == 0) return;
    if (scene.format == GSR_SCENE_F32_SOA)
        hipLaunchKernelGGL(kPreprocess<GSR_SCENE_F32_SOA>, dim3(blocks), dim3(kPreprocessThreads), 0, stream, scene,
                           cam, bounds, splats, tileCounts, blockStats, superStats, control, spanInfo, spanPlanes,
                           depthBits, rangesWords, rangeWordCount, renderSplat, packed, rayDepth, useMlp);
    else
        hipLaunchKernelGGL(kPreprocess<GSR_SCENE_F16>, dim3(blocks), dim3(kPreprocessThreads), 0, stream, scene, cam,
                           bounds, splats, tileCounts, blockStats, superStats, control, spanInfo, spanPlanes,
                           depthBits, rangesWords, rangeWordCount, renderSplat, packed, rayDepth, useMlp);
    markKernel("kPreprocess", stream);
}

}  // namespace gsr
