// C entry points of the CPU oracle (ctypes, tests and the CPU baseline). TEST
// INFRASTRUCTURE ONLY (see gsr_oracle.hpp).
//
//   oracle_render        renderReference (reference_renderer.cpp:41-234): preprocess,
//                        DepthSorter, per-tile counting-sort binning, front-to-back raster
//   oracle_preprocess    preprocessGaussian over a scene (+ the GPU culls and tile counts,
//                        splat_preprocess.comp:411-450)
//   oracle_tile_lists    the (tile, splat) list contract of the GPU path
//                        (test_tile_raster.cpp:332-382, cpuTileList) incl. capacity clamping
#include "gsr_oracle.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <numeric>
#include <thread>

using namespace oracle;

namespace {

using Clock = std::chrono::steady_clock;
double msBetween(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// parallelFor (reference_renderer.cpp:23-37): chunks handed out from an atomic counter.
void parallelFor(uint32_t count, unsigned threads, uint32_t chunk, const std::function<void(uint32_t)>& fn) {
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const uint32_t begin = next.fetch_add(chunk);
            if (begin >= count) return;
            const uint32_t end = std::min(count, begin + chunk);
            for (uint32_t i = begin; i < end; ++i) fn(i);
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; ++t) pool.emplace_back(worker);
    worker();
    for (std::thread& t : pool) t.join();
}

unsigned resolveThreads(int threads) {
    return threads > 0 ? unsigned(threads) : std::max(1u, std::thread::hardware_concurrency());
}

struct SceneView {
    uint32_t n;
    const float* posAlpha;  // 4 per gaussian
    const float* scales;    // 4 per gaussian (sigma, 0)
    const float* rot;       // 4 per gaussian (w, x, y, z)
    const float* sh;        // K * 3 per gaussian
    uint32_t K;
};

}  // namespace

extern "C" {

struct oracle_render_options {
    int32_t maxShDegree;       // evaluated bands = min(scene degree, this)
    int32_t earlyTermination;  // INRIA's T < 1e-4 stop
    uint32_t bound;            // 0 = tile (INRIA), 1 = radius square
    uint32_t bounds;           // 0 rect, 1 tight, 2 cutoff-only
    int32_t threads;           // 0 = hardware_concurrency
    float background[3];
    uint32_t renderSplat;      // 0 full, 1 p32, 2 p24: render what the packed record holds
    uint32_t sortWindow;       // StopThePop window K (0 = INRIA order; 0xFFFFFFFF = exact per-pixel sort)
};

struct oracle_stats {  // ReferenceStats (reference_renderer.h:51-62)
    uint32_t gaussians, visible, culledNear, culledSingular, culledOffScreen, threads;
    uint64_t tileInstances, evaluatedFragments, blendedFragments, terminatedPixels;
    double preprocessMs, sortMs, binMs, rasterMs, totalMs;
};

void oracle_camera_from_viewer(const float view[16], const float proj[16], const float model[16], uint32_t w,
                               uint32_t h, float viewInria[16], float* tanX, float* tanY) {
    cameraFromViewer(view, proj, model, w, h, viewInria, tanX, tanY);
}

// makeSplatCamera's derived fields.
void oracle_make_camera(const float viewInria[16], uint32_t w, uint32_t h, float tanX, float tanY, float pos[3],
                        float focal[2], uint32_t tiles[2]) {
    const Camera<float> c = makeCamera<float>(viewInria, w, h, tanX, tanY);
    for (int i = 0; i < 3; ++i) pos[i] = c.pos[i];
    focal[0] = c.focalX;
    focal[1] = c.focalY;
    tiles[0] = c.tilesX;
    tiles[1] = c.tilesY;
}

float oracle_det_logf(float x) { return det_logf(x); }
uint16_t oracle_float_to_half(float v) { return floatToHalf(v); }
float oracle_half_to_float(uint16_t h) { return halfToFloat(h); }

// The individual per-gaussian steps on a camera-space point (the known answers of
// test_splat_math.cpp). Outputs in float; *_f64 the same in double.
void oracle_splat_math(const float t[3], const float sigma[3], const float q[4], const float viewInria[16], uint32_t w,
                       uint32_t h, float tanX, float tanY, float cov3D[9], float cov2D[3], float conic[3],
                       float* radius, int32_t* invertible, float center[2]) {
    const Camera<float> cam = makeCamera<float>(viewInria, w, h, tanX, tanY);
    const M3<float> S = computeCov3D<float>(sigma, q);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) cov3D[c * 3 + r] = S.m[c][r];
    computeCov2D<float>(t, S, cam, cov2D);
    *invertible = conicAndRadius<float>(cov2D, conic, radius) ? 1 : 0;
    projectToPixel<float>(t, cam, center);
}

void oracle_splat_math_f64(const double t[3], const double sigma[3], const double q[4], const float viewInria[16],
                           uint32_t w, uint32_t h, float tanX, float tanY, double cov3D[9], double cov2D[3],
                           double conic[3], double* radius, int32_t* invertible, double center[2]) {
    const Camera<double> cam = makeCamera<double>(viewInria, w, h, tanX, tanY);
    const M3<double> S = computeCov3D<double>(sigma, q);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) cov3D[c * 3 + r] = S.m[c][r];
    computeCov2D<double>(t, S, cam, cov2D);
    *invertible = conicAndRadius<double>(cov2D, conic, radius) ? 1 : 0;
    projectToPixel<double>(t, cam, center);
}

void oracle_tile_rect(float cx, float cy, float radius, uint32_t tilesX, uint32_t tilesY, int32_t out[4]) {
    const TileRect r = computeTileRect(cx, cy, radius, tilesX, tilesY);
    out[0] = r.minX;
    out[1] = r.minY;
    out[2] = r.maxX;
    out[3] = r.maxY;
}

void oracle_eval_sh(int32_t degree, const float* sh, const float dir[3], float out[3]) {
    evalSH<float>(degree, sh, dir, out);
}

float oracle_splat_alpha_at(const void* splat, float px, float py) {
    return splatAlphaAt(*static_cast<const Splat*>(splat), px, py);
}

// preprocessGaussian over a whole scene. out: 48-byte splats; counts: the bin tiles of each
// drawn splat under `bounds` (0 if culled), as the GPU's tileCounts[].
// binSize S = 16 or 8: the counts are of bins of S pixels (TileSettings::tileSize; the grid
// ceil(w / S) x ceil(h / S)), from getRect's 16-px rectangle (splat_preprocess.comp:291-308).
void oracle_preprocess_bins(uint32_t n, const float* posAlpha, const float* scales, const float* rot, const float* sh,
                            uint32_t K, int32_t degree, const float viewInria[16], uint32_t w, uint32_t h, float tanX,
                            float tanY, int32_t gpuCull, uint32_t bounds, uint32_t renderSplat, int32_t binSize,
                            void* outSplats, uint32_t* outCounts, int32_t threads) {
    const Camera<float> cam = makeCamera<float>(viewInria, w, h, tanX, tanY);
    const int S = binSize;
    const int gx = int((w + uint32_t(S) - 1) / uint32_t(S)), gy = int((h + uint32_t(S) - 1) / uint32_t(S));
    Splat* out = static_cast<Splat*>(outSplats);
    parallelFor(n, resolveThreads(threads), 4096, [&](uint32_t i) {
        const SplatT<float> s = preprocessGaussian<float>(posAlpha + size_t(i) * 4, scales + size_t(i) * 4,
                                                          rot + size_t(i) * 4, sh + size_t(i) * K * 3, degree, cam,
                                                          gpuCull != 0);
        out[i] = toSplat(s);
        // p24: the conic and opacity as the packed record holds them (splat_preprocess.comp:296-305).
        if (s.cullReason == kVisible) applyRenderSplat(renderSplat, out[i], false);
        if (outCounts) {
            uint32_t c = 0;
            if (s.cullReason == kVisible) {
                const Coverage cv = splatCoverage(bounds, out[i], cam.tilesX, cam.tilesY, S, gx, gy);
                c = coverageCount(cv, S);
            }
            outCounts[i] = c;
        }
    });
}

void oracle_preprocess(uint32_t n, const float* posAlpha, const float* scales, const float* rot, const float* sh,
                       uint32_t K, int32_t degree, const float viewInria[16], uint32_t w, uint32_t h, float tanX,
                       float tanY, int32_t gpuCull, uint32_t bounds, uint32_t renderSplat, void* outSplats,
                       uint32_t* outCounts, int32_t threads) {
    oracle_preprocess_bins(n, posAlpha, scales, rot, sh, K, degree, viewInria, w, h, tanX, tanY, gpuCull, bounds,
                           renderSplat, kTileSize, outSplats, outCounts, threads);
}

// FP64 twin: 12 doubles per gaussian: cx, cy, depth, cull, conic a b c, opacity, r g b, radius.
void oracle_preprocess_f64(uint32_t n, const float* posAlpha, const float* scales, const float* rot, const float* sh,
                           uint32_t K, int32_t degree, const float viewInria[16], uint32_t w, uint32_t h, float tanX,
                           float tanY, int32_t gpuCull, double* out, int32_t threads) {
    const Camera<double> cam = makeCamera<double>(viewInria, w, h, tanX, tanY);
    parallelFor(n, resolveThreads(threads), 4096, [&](uint32_t i) {
        const SplatT<double> s = preprocessGaussian<double>(posAlpha + size_t(i) * 4, scales + size_t(i) * 4,
                                                            rot + size_t(i) * 4, sh + size_t(i) * K * 3, degree, cam,
                                                            gpuCull != 0);
        double* o = out + size_t(i) * 12;
        o[0] = s.center[0];
        o[1] = s.center[1];
        o[2] = s.depth;
        o[3] = double(s.cullReason);
        o[4] = s.conic[0];
        o[5] = s.conic[1];
        o[6] = s.conic[2];
        o[7] = s.opacity;
        o[8] = s.color[0];
        o[9] = s.color[1];
        o[10] = s.color[2];
        o[11] = s.radius;
    });
}

// The bin tiles of one projected splat, enumerated row by row (what the duplicate pass
// writes). Returns the count; fills up to maxTiles tile ids; *borderline per
// coverageIsBorderline(eps = 2e-3).
uint32_t oracle_coverage_tiles(const void* splat, uint32_t mode, int32_t S, uint32_t tilesX16, uint32_t tilesY16,
                               int32_t gridX, int32_t gridY, uint32_t* outTiles, uint32_t maxTiles,
                               int32_t* borderline) {
    const Splat& s = *static_cast<const Splat*>(splat);
    const Coverage cv = splatCoverage(mode, s, tilesX16, tilesY16, S, gridX, gridY);
    int r0, r1;
    coverageRows(cv, S, &r0, &r1);
    uint32_t k = 0;
    for (int r = r0; r < r1; ++r) {
        int x0, x1;
        coverageRowSpan(cv, r, S, &x0, &x1);
        for (int x = x0; x < x1; ++x, ++k)
            if (outTiles && k < maxTiles) outTiles[k] = uint32_t(r * gridX + x);
    }
    if (borderline) *borderline = coverageIsBorderline(cv, S, 2e-3f) ? 1 : 0;
    return coverageCount(cv, S);
}

// The GPU path's sorted (tile, splat) list from projected splats and their counts
// (cpuTileList, test_tile_raster.cpp:332-382). mode 0 = global (duplicate in index order,
// one stable sort by (tile, depth bits)), 1 = binning (duplicate in (depth, index) order,
// stable sort by tile). Entries past `capacity` are dropped (overflow). Outputs (may be
// null): keysHi = tile, keysLo = depth bits, values = gaussian index (capacity each);
// ranges = [start, end) per tile (2 * tiles). Returns entries; *duplicates = M.
// The same with bins of S = 16 or 8 pixels on a grid of gridX x gridY bins (tilesX / tilesY: the
// 16-px tiles of getRect); tile ids and ranges are bin ids.
uint32_t oracle_tile_lists_bins(const void* splats, const uint32_t* counts, uint32_t n, uint32_t tilesX,
                                uint32_t tilesY, int32_t S, uint32_t gridX, uint32_t gridY, uint32_t bounds,
                                uint32_t capacity, uint32_t mode, uint32_t* keysHi, uint32_t* keysLo, uint32_t* values,
                                uint32_t* ranges, uint64_t* duplicates, uint32_t* coverageMismatch);

uint32_t oracle_tile_lists(const void* splats, const uint32_t* counts, uint32_t n, uint32_t tilesX, uint32_t tilesY,
                           uint32_t bounds, uint32_t capacity, uint32_t mode, uint32_t* keysHi, uint32_t* keysLo,
                           uint32_t* values, uint32_t* ranges, uint64_t* duplicates, uint32_t* coverageMismatch) {
    return oracle_tile_lists_bins(splats, counts, n, tilesX, tilesY, kTileSize, tilesX, tilesY, bounds, capacity, mode,
                                  keysHi, keysLo, values, ranges, duplicates, coverageMismatch);
}

uint32_t oracle_tile_lists_bins(const void* splats, const uint32_t* counts, uint32_t n, uint32_t tilesX,
                                uint32_t tilesY, int32_t S, uint32_t gridX, uint32_t gridY, uint32_t bounds,
                                uint32_t capacity, uint32_t mode, uint32_t* keysHi, uint32_t* keysLo, uint32_t* values,
                                uint32_t* ranges, uint64_t* duplicates, uint32_t* coverageMismatch) {
    const Splat* sp = static_cast<const Splat*>(splats);
    std::vector<uint32_t> visible;
    for (uint32_t i = 0; i < n; ++i)
        if (counts[i] != 0) visible.push_back(i);
    if (mode == 1)
        std::stable_sort(visible.begin(), visible.end(),
                         [&](uint32_t a, uint32_t b) { return fbits(sp[a].depth) < fbits(sp[b].depth); });
    std::vector<uint32_t> tiles, vals;
    uint64_t dup = 0;
    uint32_t mismatch = 0;
    const int gx = int(gridX), gy = int(gridY);
    for (uint32_t g : visible) {
        const Coverage cv = splatCoverage(bounds, sp[g], tilesX, tilesY, S, gx, gy);
        int r0, r1;
        coverageRows(cv, S, &r0, &r1);
        uint32_t written = 0;
        for (int r = r0; r < r1; ++r) {
            int x0, x1;
            coverageRowSpan(cv, r, S, &x0, &x1);
            for (int x = x0; x < x1; ++x) {
                ++dup;
                ++written;
                if (tiles.size() < capacity) {
                    tiles.push_back(uint32_t(r * gx + x));
                    vals.push_back(g);
                }
            }
        }
        mismatch += written != counts[g];
    }
    std::vector<uint32_t> order(tiles.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        if (tiles[a] != tiles[b]) return tiles[a] < tiles[b];
        if (mode == 0) return fbits(sp[vals[a]].depth) < fbits(sp[vals[b]].depth);
        return false;
    });
    const uint32_t entries = uint32_t(order.size());
    for (uint32_t k = 0; k < entries; ++k) {
        const uint32_t o = order[k];
        if (keysHi) keysHi[k] = tiles[o];
        if (keysLo) keysLo[k] = fbits(sp[vals[o]].depth);
        if (values) values[k] = vals[o];
    }
    if (ranges) {
        const size_t T = size_t(gridX) * gridY;
        std::fill(ranges, ranges + 2 * T, 0u);
        for (uint32_t k = 0; k < entries; ++k) {
            const uint32_t t = tiles[order[k]];
            if (k == 0 || tiles[order[k - 1]] != t) ranges[2 * t] = k;
            ranges[2 * t + 1] = k + 1;
        }
    }
    if (duplicates) *duplicates = dup;
    if (coverageMismatch) *coverageMismatch = mismatch;
    return entries;
}

uint32_t oracle_radix_sort_pairs(uint32_t* keys, uint32_t* values, uint32_t n) {
    std::vector<uint32_t> k(keys, keys + n), v(values, values + n);
    const uint32_t passes = radixSortPairs(k, v);
    std::copy(k.begin(), k.end(), keys);
    std::copy(v.begin(), v.end(), values);
    return passes;
}

// f16 records (scene_compression.cpp:34-49 encode, 382-415 + 432-435 decode): geometry
// 8 words per gaussian, colour K*3 halves padded to a multiple of 8 halves.
void oracle_encode_f16(uint32_t n, const float* posAlpha, const float* scales, const float* rot, const float* sh,
                       uint32_t K, uint32_t* geom, uint16_t* color) {
    const uint32_t colorHalves = (K * 3 + 7) / 8 * 8;
    for (uint32_t i = 0; i < n; ++i) {
        const float* pa = posAlpha + size_t(i) * 4;
        const float* sc = scales + size_t(i) * 4;
        const float* q = rot + size_t(i) * 4;
        uint32_t* w = geom + size_t(i) * 8;
        w[0] = fbits(pa[0]);
        w[1] = fbits(pa[1]);
        w[2] = fbits(pa[2]);
        w[3] = uint32_t(floatToHalf(pa[3])) | (uint32_t(floatToHalf(sc[0])) << 16);
        w[4] = uint32_t(floatToHalf(sc[1])) | (uint32_t(floatToHalf(sc[2])) << 16);
        w[5] = uint32_t(floatToHalf(q[0])) | (uint32_t(floatToHalf(q[1])) << 16);
        w[6] = uint32_t(floatToHalf(q[2])) | (uint32_t(floatToHalf(q[3])) << 16);
        w[7] = 0;
        uint16_t* c = color + size_t(i) * colorHalves;
        for (uint32_t j = 0; j < colorHalves; ++j) c[j] = j < K * 3 ? floatToHalf(sh[size_t(i) * K * 3 + j]) : 0;
    }
}

void oracle_decode_f16(uint32_t n, const uint32_t* geom, const uint16_t* color, uint32_t K, float* posAlpha,
                       float* scales, float* rot, float* sh) {
    const uint32_t colorHalves = (K * 3 + 7) / 8 * 8;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* g = geom + size_t(i) * 8;
        float* pa = posAlpha + size_t(i) * 4;
        float* sc = scales + size_t(i) * 4;
        float* q = rot + size_t(i) * 4;
        pa[0] = bitsf(g[0]);
        pa[1] = bitsf(g[1]);
        pa[2] = bitsf(g[2]);
        pa[3] = halfToFloat(uint16_t(g[3] & 0xFFFFu));
        sc[0] = halfToFloat(uint16_t(g[3] >> 16));
        sc[1] = halfToFloat(uint16_t(g[4] & 0xFFFFu));
        sc[2] = halfToFloat(uint16_t(g[4] >> 16));
        sc[3] = 0.0f;
        const float qw = halfToFloat(uint16_t(g[5] & 0xFFFFu)), qx = halfToFloat(uint16_t(g[5] >> 16));
        const float qy = halfToFloat(uint16_t(g[6] & 0xFFFFu)), qz = halfToFloat(uint16_t(g[6] >> 16));
        const float len2 = qw * qw + qx * qx + qy * qy + qz * qz;
        if (len2 > 0.0f) {
            const float inv = 1.0f / std::sqrt(len2);
            q[0] = qw * inv;
            q[1] = qx * inv;
            q[2] = qy * inv;
            q[3] = qz * inv;
        } else {
            q[0] = 1.0f;
            q[1] = q[2] = q[3] = 0.0f;
        }
        const uint16_t* c = color + size_t(i) * colorHalves;
        for (uint32_t j = 0; j < K * 3; ++j) sh[size_t(i) * K * 3 + j] = halfToFloat(c[j]);
    }
}

// renderReference (reference_renderer.cpp:41-234), without the StopThePop window. rgb:
// W*H*3 floats, row 0 = top, unclamped. tileStart/tileLists (optional): the per-tile lists
// in depth order (tiles + 1 and tileInstances entries; pass null to skip).
// computeRayDepthParams for one gaussian: out = p00, p01, p02, p11, p12, p22, b0, b1, b2.
void oracle_ray_depth(const float sigma[3], const float q[4], const float tCam[3], const float viewInria[16],
                      float out[9]) {
    float view[4][4];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) view[c][r] = viewInria[c * 4 + r];
    const RayDepth d = computeRayDepth(sigma, q, tCam, view);
    const float v[9] = {d.p00, d.p01, d.p02, d.p11, d.p12, d.p22, d.b0, d.b1, d.b2};
    for (int i = 0; i < 9; ++i) out[i] = v[i];
}

// encodeRenderSplat over n splats: 8 (p32) or 6 (p24) words each.
void oracle_encode_render_splats(uint32_t n, const void* splats, uint32_t mode, uint32_t* words) {
    const Splat* sp = static_cast<const Splat*>(splats);
    const uint32_t stride = mode == kRenderSplatP24 ? 6u : 8u;
    for (uint32_t i = 0; i < n; ++i) encodeRenderSplat(mode, sp[i], words + size_t(i) * stride);
}

void oracle_render(uint32_t n, const float* posAlpha, const float* scales, const float* rot, const float* sh,
                   uint32_t K, int32_t shDegree, const float viewInria[16], uint32_t w, uint32_t h, float tanX,
                   float tanY, const oracle_render_options* opt, float* rgb, oracle_stats* statsOut,
                   uint32_t* tileStartOut, uint32_t* tileListsOut, uint64_t tileListsCapacity) {
    const auto tStart = Clock::now();
    const Camera<float> cam = makeCamera<float>(viewInria, w, h, tanX, tanY);
    oracle_stats stats{};
    stats.gaussians = n;
    stats.threads = resolveThreads(opt->threads);
    const unsigned threads = stats.threads;
    std::fill(rgb, rgb + size_t(w) * h * 3, 0.0f);

    // 1) Preprocess.
    const int degree = std::min(shDegree, int(opt->maxShDegree));
    std::vector<Splat> splats(n);
    parallelFor(n, threads, 4096, [&](uint32_t i) {
        splats[i] = toSplat(preprocessGaussian<float>(posAlpha + size_t(i) * 4, scales + size_t(i) * 4,
                                                      rot + size_t(i) * 4, sh + size_t(i) * K * 3, degree, cam, false));
        if (splats[i].cullReason == kVisible) applyRenderSplat(opt->renderSplat, splats[i], true);
    });
    for (const Splat& s : splats) {
        switch (s.cullReason) {
            case kVisible: stats.visible++; break;
            case kNearPlane: stats.culledNear++; break;
            case kSingularCov: stats.culledSingular++; break;
            case kOffScreen: stats.culledOffScreen++; break;
            default: break;
        }
    }
    // StopThePop: ray-depth parameters of every visible gaussian (reference_renderer.cpp:61-70).
    std::vector<RayDepth> rayDepth;
    if (opt->sortWindow > 0) {
        rayDepth.resize(n);
        parallelFor(n, threads, 4096, [&](uint32_t i) {
            if (splats[i].cullReason != kVisible) return;
            const float* m = posAlpha + size_t(i) * 4;
            float t[3];
            for (int r = 0; r < 3; ++r)
                t[r] = cam.view[0][r] * m[0] + cam.view[1][r] * m[1] + cam.view[2][r] * m[2] + cam.view[3][r];
            rayDepth[i] = computeRayDepth(scales + size_t(i) * 4, rot + size_t(i) * 4, t, cam.view);
        });
    }
    const auto tPre = Clock::now();

    // 2) Depth order (DepthSorter::sort, depth_sort.cpp:17-42).
    std::vector<uint32_t> keys, order;
    keys.reserve(n);
    order.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = posAlpha + size_t(i) * 4;
        const float z = cam.view[0][2] * p[0] + cam.view[1][2] * p[1] + cam.view[2][2] * p[2] + cam.view[3][2];
        if (!(z > kNearCullDepth)) continue;
        keys.push_back(fbits(z));
        order.push_back(i);
    }
    radixSortPairs(keys, order);
    const auto tSort = Clock::now();

    // 3) Per-tile lists in depth order (counting sort by tile).
    const uint32_t tileCount = cam.tilesX * cam.tilesY;
    std::vector<uint32_t> tileStart(tileCount + 1, 0);
    const int gx = int(cam.tilesX), gy = int(cam.tilesY);
    auto forEachTile = [&](const Splat& s, auto&& fn) {
        const Coverage cv = splatCoverage(opt->bounds, s, cam.tilesX, cam.tilesY, kTileSize, gx, gy);
        int r0, r1;
        coverageRows(cv, kTileSize, &r0, &r1);
        for (int ty = r0; ty < r1; ++ty) {
            int x0, x1;
            coverageRowSpan(cv, ty, kTileSize, &x0, &x1);
            for (int tx = x0; tx < x1; ++tx) fn(uint32_t(ty * gx + tx));
        }
    };
    for (uint32_t idx : order) {
        const Splat& s = splats[idx];
        if (s.cullReason != kVisible) continue;
        forEachTile(s, [&](uint32_t t) { tileStart[t + 1]++; });
    }
    for (uint32_t t = 0; t < tileCount; ++t) tileStart[t + 1] += tileStart[t];
    stats.tileInstances = tileStart[tileCount];
    std::vector<uint32_t> tileLists(tileStart[tileCount]);
    {
        std::vector<uint32_t> cursor(tileStart.begin(), tileStart.end() - 1);
        for (uint32_t idx : order) {
            const Splat& s = splats[idx];
            if (s.cullReason != kVisible) continue;
            forEachTile(s, [&](uint32_t t) { tileLists[cursor[t]++] = idx; });
        }
    }
    const auto tBin = Clock::now();

    // 4) Rasterize, one tile per task (reference_renderer.cpp:118-222).
    std::atomic<uint64_t> evaluated{0}, blended{0}, terminated{0};
    const bool radiusBound = opt->bound == kBoundRadius;
    parallelFor(tileCount, threads, 1, [&](uint32_t tile) {
        const uint32_t tx = tile % cam.tilesX, ty = tile / cam.tilesX;
        std::vector<Splat> local;
        local.reserve(tileStart[tile + 1] - tileStart[tile]);
        for (uint32_t j = tileStart[tile]; j < tileStart[tile + 1]; ++j) local.push_back(splats[tileLists[j]]);
        std::vector<RayDepth> localDepth;
        if (opt->sortWindow > 0)
            for (uint32_t j = tileStart[tile]; j < tileStart[tile + 1]; ++j) localDepth.push_back(rayDepth[tileLists[j]]);
        struct Pending {
            float depth, alpha, color[3];
        };
        std::vector<Pending> window;
        uint64_t tEval = 0, tBlend = 0, tTerm = 0;
        const uint32_t x0 = tx * kTileSize, y0 = ty * kTileSize;
        const uint32_t x1 = std::min(cam.width, x0 + kTileSize), y1 = std::min(cam.height, y0 + kTileSize);
        for (uint32_t y = y0; y < y1; ++y) {
            for (uint32_t x = x0; x < x1; ++x) {
                const float px = float(x), py = float(y);
                float T = 1.0f;
                float C[3] = {0.0f, 0.0f, 0.0f};
                if (opt->sortWindow > 0) {
                    // StopThePop (reference_renderer.cpp:143-196): the same blend rule, in the
                    // order of a K-slot window sorted by the depth of each splat's maximum along
                    // the pixel's ray; the nearest of (window + new) leaves the window.
                    const float vx = (px - 0.5f * (float(cam.width) - 1.0f)) / cam.focalX;
                    const float vy = (py - 0.5f * (float(cam.height) - 1.0f)) / cam.focalY;
                    bool done = false;
                    auto blend = [&](float alpha, const float* col) {
                        const float testT = T * (1.0f - alpha);
                        if (testT < kMinTransmittance) {
                            done = true;
                            ++tTerm;
                            return;
                        }
                        for (int c = 0; c < 3; ++c) C[c] += col[c] * alpha * T;
                        T = testT;
                        ++tBlend;
                    };
                    const bool exact = opt->sortWindow == 0xFFFFFFFFu;
                    const size_t K = opt->sortWindow;
                    window.clear();
                    for (size_t j = 0; j < local.size() && !done; ++j) {
                        const Splat& s = local[j];
                        if (radiusBound) {
                            const float fx = px + 0.5f, fy = py + 0.5f;
                            const float cx = s.center[0] + 0.5f, cy = s.center[1] + 0.5f;
                            if (!(fx >= cx - s.radius && fx < cx + s.radius && fy >= cy - s.radius && fy < cy + s.radius))
                                continue;
                        }
                        ++tEval;
                        const float alpha = splatAlphaAt(s, px, py);
                        if (alpha == 0.0f) continue;
                        const Pending e{localDepth[j].depthAt(vx, vy), alpha, {s.color[0], s.color[1], s.color[2]}};
                        if (exact) {
                            window.push_back(e);
                        } else if (window.size() == K) {
                            if (e.depth < window[0].depth) {
                                blend(e.alpha, e.color);
                                continue;
                            }
                            blend(window[0].alpha, window[0].color);
                            size_t i = 0;
                            for (; i + 1 < K && window[i + 1].depth <= e.depth; ++i) window[i] = window[i + 1];
                            window[i] = e;
                        } else {
                            size_t i = window.size();
                            window.push_back(e);
                            for (; i > 0 && window[i - 1].depth > e.depth; --i) window[i] = window[i - 1];
                            window[i] = e;
                        }
                    }
                    if (exact)
                        std::stable_sort(window.begin(), window.end(),
                                         [](const Pending& a, const Pending& b) { return a.depth < b.depth; });
                    for (size_t i = 0; i < window.size() && !done; ++i) blend(window[i].alpha, window[i].color);
                    float* dst = &rgb[(size_t(y) * cam.width + x) * 3];
                    for (int c = 0; c < 3; ++c) dst[c] = C[c] + T * opt->background[c];
                    continue;
                }
                for (const Splat& s : local) {
                    if (radiusBound) {
                        // insideRadiusSquare (splat_math.h:166-170)
                        const float fx = px + 0.5f, fy = py + 0.5f;
                        const float cx = s.center[0] + 0.5f, cy = s.center[1] + 0.5f;
                        if (!(fx >= cx - s.radius && fx < cx + s.radius && fy >= cy - s.radius && fy < cy + s.radius))
                            continue;
                    }
                    ++tEval;
                    const float alpha = splatAlphaAt(s, px, py);
                    if (alpha == 0.0f) continue;
                    const float testT = T * (1.0f - alpha);
                    if (opt->earlyTermination && testT < kMinTransmittance) {
                        ++tTerm;
                        break;
                    }
                    for (int c = 0; c < 3; ++c) C[c] += s.color[c] * alpha * T;
                    T = testT;
                    ++tBlend;
                }
                float* dst = &rgb[(size_t(y) * cam.width + x) * 3];
                for (int c = 0; c < 3; ++c) dst[c] = C[c] + T * opt->background[c];
            }
        }
        evaluated += tEval;
        blended += tBlend;
        terminated += tTerm;
    });
    stats.evaluatedFragments = evaluated;
    stats.blendedFragments = blended;
    stats.terminatedPixels = terminated;
    const auto tEnd = Clock::now();
    stats.preprocessMs = msBetween(tStart, tPre);
    stats.sortMs = msBetween(tPre, tSort);
    stats.binMs = msBetween(tSort, tBin);
    stats.rasterMs = msBetween(tBin, tEnd);
    stats.totalMs = msBetween(tStart, tEnd);
    if (statsOut) *statsOut = stats;
    if (tileStartOut) std::copy(tileStart.begin(), tileStart.end(), tileStartOut);
    if (tileListsOut)
        std::copy(tileLists.begin(), tileLists.begin() + std::min<uint64_t>(tileListsCapacity, tileLists.size()),
                  tileListsOut);
}

}  // extern "C"
