// Test infrastructure only: the FP64 forward of the differentiable render and its central finite
// differences, the reference for the GPU backward (no counterpart in the reference, which renders
// forward only). Parameters are the raw ones the gradients are taken in: means, log-scales
// (sigma = exp), quaternions (w, x, y, z, used unnormalised as the f32 scene path does), opacity
// logits (opacity = sigmoid) and SH coefficients. The per-gaussian math is the oracle's templated
// FP64 twin (gsr_oracle.hpp: computeCov3D, computeCov2D, conicAndRadius, projectToPixel,
// computeTileRect; preprocessGaussian's culls in the GPU form, splat_math.cpp:153-193); the
// composite is renderReference's per-pixel loop (reference_renderer.cpp:118-196) in double, with
// rect bounds (a splat reaches the pixels of the tiles of its tile rect), black background, and
// splats in (depth, index) order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "gsr_oracle.hpp"

using namespace oracle;

namespace {

struct Params {
    uint32_t n;
    const double* means;
    const double* logScales;
    const double* quats;
    const double* logits;
    const double* sh;
    uint32_t K;
    int32_t degree;
};

struct DSplat {
    double cx, cy, depth, a, b, c, o, col[3];
    TileRect rect;
    uint32_t index;
};

// evalSH (splat_math.cpp:129-151) with double coefficients, clamped at 0.
void evalShD(int degree, const double* sh, const double dir[3], double out[3]) {
    for (int ch = 0; ch < 3; ++ch) {
        auto c = [&](int k) { return sh[k * 3 + ch]; };
        double r = double(kShC0) * c(0);
        if (degree > 0) {
            const double x = dir[0], y = dir[1], z = dir[2], C1 = double(kShC1);
            r = r - C1 * y * c(1) + C1 * z * c(2) - C1 * x * c(3);
            if (degree > 1) {
                const double xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
                r = r + double(kShC2[0]) * xy * c(4) + double(kShC2[1]) * yz * c(5) +
                    double(kShC2[2]) * (2 * zz - xx - yy) * c(6) + double(kShC2[3]) * xz * c(7) +
                    double(kShC2[4]) * (xx - yy) * c(8);
                if (degree > 2) {
                    r = r + double(kShC3[0]) * y * (3 * xx - yy) * c(9) + double(kShC3[1]) * xy * z * c(10) +
                        double(kShC3[2]) * y * (4 * zz - xx - yy) * c(11) +
                        double(kShC3[3]) * z * (2 * zz - 3 * xx - 3 * yy) * c(12) +
                        double(kShC3[4]) * x * (4 * zz - xx - yy) * c(13) + double(kShC3[5]) * z * (xx - yy) * c(14) +
                        double(kShC3[6]) * x * (xx - 3 * yy) * c(15);
                }
            }
        }
        r += 0.5;
        out[ch] = r < 0 ? 0 : r;
    }
}

// The GPU preprocess's culls and splat, in double, from raw parameters.
bool splatOf(const Params& p, uint32_t i, const Camera<double>& cam, DSplat* s) {
    const double* m = p.means + 3 * size_t(i);
    double t[3];
    for (int r = 0; r < 3; ++r)
        t[r] = cam.view[0][r] * m[0] + cam.view[1][r] * m[1] + cam.view[2][r] * m[2] + cam.view[3][r];
    if (!(t[2] > double(kNearCullDepth))) return false;
    const double o = 1.0 / (1.0 + std::exp(-p.logits[i]));
    if (o < double(kMinAlpha)) return false;
    const double sigma[3] = {std::exp(p.logScales[3 * size_t(i)]), std::exp(p.logScales[3 * size_t(i) + 1]),
                             std::exp(p.logScales[3 * size_t(i) + 2])};
    const double* q = p.quats + 4 * size_t(i);
    const M3<double> cov3 = computeCov3D(sigma, q);
    double cov2[3], conic[3], radius = 0, center[2];
    computeCov2D(t, cov3, cam, cov2);
    if (!conicAndRadius(cov2, conic, &radius)) return false;
    projectToPixel(t, cam, center);
    const TileRect rect = computeTileRect(float(center[0]), float(center[1]), float(radius), cam.tilesX, cam.tilesY);
    if (rect.empty()) return false;
    const double v[3] = {m[0] - cam.pos[0], m[1] - cam.pos[1], m[2] - cam.pos[2]};
    const double inv = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double dir[3] = {v[0] * inv, v[1] * inv, v[2] * inv};
    evalShD(p.degree, p.sh + size_t(i) * p.K * 3, dir, s->col);
    s->cx = center[0];
    s->cy = center[1];
    s->depth = t[2];
    s->a = conic[0];
    s->b = conic[1];
    s->c = conic[2];
    s->o = o;
    s->rect = rect;
    s->index = i;
    return true;
}

// L1 loss of the render against target (H x W x 3), and the image if asked for.
double lossOf(const Params& p, const Camera<double>& cam, const float* target, double* image) {
    std::vector<DSplat> splats;
    splats.reserve(p.n);
    for (uint32_t i = 0; i < p.n; ++i) {
        DSplat s;
        if (splatOf(p, i, cam, &s)) splats.push_back(s);
    }
    std::stable_sort(splats.begin(), splats.end(), [](const DSplat& a, const DSplat& b) { return a.depth < b.depth; });
    double loss = 0;
    for (uint32_t y = 0; y < cam.height; ++y) {
        const int ty = int(y / kTileSize);
        for (uint32_t x = 0; x < cam.width; ++x) {
            const int tx = int(x / kTileSize);
            double T = 1, C[3] = {0, 0, 0};
            for (const DSplat& s : splats) {
                if (tx < s.rect.minX || tx >= s.rect.maxX || ty < s.rect.minY || ty >= s.rect.maxY) continue;
                const double dx = s.cx - double(x), dy = s.cy - double(y);
                const double power = -0.5 * (s.a * dx * dx + s.c * dy * dy) - s.b * dx * dy;
                if (power > 0) continue;
                const double e = s.o * std::exp(power);
                const double alpha = e < double(kMaxAlpha) ? e : double(kMaxAlpha);
                if (alpha < double(kMinAlpha)) continue;
                const double testT = T * (1 - alpha);
                if (testT < double(kMinTransmittance)) break;
                for (int c = 0; c < 3; ++c) C[c] += s.col[c] * alpha * T;
                T = testT;
            }
            const size_t px = (size_t(y) * cam.width + x) * 3;
            for (int c = 0; c < 3; ++c) {
                if (image) image[px + c] = C[c];
                if (target) loss += std::fabs(C[c] - double(target[px + c]));
            }
        }
    }
    return loss / (3.0 * double(cam.width) * double(cam.height));
}

}  // namespace

extern "C" {

// The FP64 image (H x W x 3) and L1 loss against target (may be null: loss 0).
double oracle_diff_render_f64(uint32_t n, const double* means, const double* logScales, const double* quats,
                              const double* logits, const double* sh, uint32_t K, int32_t degree,
                              const float viewInria[16], uint32_t width, uint32_t height, float tanX, float tanY,
                              const float* target, double* image) {
    const Params p{n, means, logScales, quats, logits, sh, K, degree};
    const Camera<double> cam = makeCamera<double>(viewInria, width, height, tanX, tanY);
    return lossOf(p, cam, target, image);
}

// Central differences (L(x + h) - L(x - h)) / 2h of the L1 loss for the gaussians `which`, in
// the order mean (3), log-scale (3), quaternion (4), logit (1), SH (K x 3): count x (11 + 3K).
void oracle_diff_fd_f64(uint32_t n, const double* means, const double* logScales, const double* quats,
                        const double* logits, const double* sh, uint32_t K, int32_t degree, const float viewInria[16],
                        uint32_t width, uint32_t height, float tanX, float tanY, const float* target, uint32_t count,
                        const uint32_t* which, double h, double* out, int32_t threads) {
    const Camera<double> cam = makeCamera<double>(viewInria, width, height, tanX, tanY);
    const uint32_t P = 11 + 3 * K;
    const uint64_t tasks = uint64_t(count) * P;
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        // Each worker perturbs its own copies.
        std::vector<double> m(means, means + 3 * size_t(n)), ls(logScales, logScales + 3 * size_t(n)),
            q(quats, quats + 4 * size_t(n)), lg(logits, logits + n), s(sh, sh + size_t(n) * K * 3);
        const Params p{n, m.data(), ls.data(), q.data(), lg.data(), s.data(), K, degree};
        for (uint64_t t = next++; t < tasks; t = next++) {
            const uint32_t g = which[t / P], k = uint32_t(t % P);
            double* x;
            if (k < 3) x = &m[3 * size_t(g) + k];
            else if (k < 6) x = &ls[3 * size_t(g) + k - 3];
            else if (k < 10) x = &q[4 * size_t(g) + k - 6];
            else if (k < 11) x = &lg[g];
            else x = &s[size_t(g) * K * 3 + (k - 11)];
            const double x0 = *x;
            *x = x0 + h;
            const double lp = lossOf(p, cam, target, nullptr);
            *x = x0 - h;
            const double lm = lossOf(p, cam, target, nullptr);
            *x = x0;
            out[t] = (lp - lm) / (2 * h);
        }
    };
    const unsigned nt = threads > 0 ? unsigned(threads) : std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < nt; ++i) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
}

}  // extern "C"
