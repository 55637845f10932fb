// Test infrastructure only: the FP64 forward of the differentiable render, its central finite
// differences, and its exact gradient (a vector-Jacobian product), the references for the GPU
// backward (no counterpart in the reference, which renders forward only). Parameters are the raw ones the gradients are taken in: means, log-scales
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

// evalSH (splat_math.cpp:129-151) with double (or dual) coefficients, clamped at 0.
template <typename Real>
void evalShD(int degree, const Real* sh, const Real dir[3], Real out[3]) {
    for (int ch = 0; ch < 3; ++ch) {
        auto c = [&](int k) { return sh[k * 3 + ch]; };
        Real r = Real(kShC0) * c(0);
        if (degree > 0) {
            const Real x = dir[0], y = dir[1], z = dir[2], C1 = Real(kShC1), two = 2, three = 3, four = 4;
            r = r - C1 * y * c(1) + C1 * z * c(2) - C1 * x * c(3);
            if (degree > 1) {
                const Real xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
                r = r + Real(kShC2[0]) * xy * c(4) + Real(kShC2[1]) * yz * c(5) +
                    Real(kShC2[2]) * (two * zz - xx - yy) * c(6) + Real(kShC2[3]) * xz * c(7) +
                    Real(kShC2[4]) * (xx - yy) * c(8);
                if (degree > 2) {
                    r = r + Real(kShC3[0]) * y * (three * xx - yy) * c(9) + Real(kShC3[1]) * xy * z * c(10) +
                        Real(kShC3[2]) * y * (four * zz - xx - yy) * c(11) +
                        Real(kShC3[3]) * z * (two * zz - three * xx - three * yy) * c(12) +
                        Real(kShC3[4]) * x * (four * zz - xx - yy) * c(13) + Real(kShC3[5]) * z * (xx - yy) * c(14) +
                        Real(kShC3[6]) * x * (xx - three * yy) * c(15);
                }
            }
        }
        r = r + Real(0.5);
        out[ch] = r < Real(0) ? Real(0) : r;
    }
}

// A forward-mode dual number: a value and its derivatives with respect to the 27 raw parameters
// of one gaussian (mean 3, log-scale 3, quaternion 4, logit 1, then SH coefficient k of all
// three channels at once for k < 16: channel ch's colour depends on channel ch's coefficients
// only, so derivative 11 + k of colour channel ch is d colour_ch / d sh[k][ch]). The oracle's
// templated per-gaussian math run on it gives the exact Jacobian of the splat (centre, conic,
// opacity, colour) in the raw parameters, by the chain rule of every operation as written.
struct Dual {
    static constexpr int N = 27;
    double v = 0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}  // NOLINT: constants convert implicitly
    static Dual seed(double x, int k) {
        Dual r(x);
        r.d[k] = 1;
        return r;
    }
};

Dual operator+(const Dual& a, const Dual& b) {
    Dual r(a.v + b.v);
    for (int k = 0; k < Dual::N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
Dual operator-(const Dual& a, const Dual& b) {
    Dual r(a.v - b.v);
    for (int k = 0; k < Dual::N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
Dual operator-(const Dual& a) {
    Dual r(-a.v);
    for (int k = 0; k < Dual::N; ++k) r.d[k] = -a.d[k];
    return r;
}
Dual operator*(const Dual& a, const Dual& b) {
    Dual r(a.v * b.v);
    for (int k = 0; k < Dual::N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
Dual operator/(const Dual& a, const Dual& b) {
    Dual r(a.v / b.v);
    for (int k = 0; k < Dual::N; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
Dual exp(const Dual& a) {
    Dual r(std::exp(a.v));
    for (int k = 0; k < Dual::N; ++k) r.d[k] = r.v * a.d[k];
    return r;
}
Dual sqrt(const Dual& a) {
    Dual r(std::sqrt(a.v));
    for (int k = 0; k < Dual::N; ++k) r.d[k] = a.d[k] / (2 * r.v);
    return r;
}

// The splat quantities the composite reads, as duals: centre, conic (a, b, c), opacity, colour.
struct GSplat {
    Dual cx, cy, a, b, c, o, col[3];
};

// splatOf's arithmetic (the preprocess: view transform, covariance, EWA projection, inverse,
// projection to pixels, view direction, SH colour) on duals seeded with gaussian i's parameters.
GSplat dualSplat(const Params& p, uint32_t i, const Camera<Dual>& cam) {
    Dual m[3], sigma[3], q[4];
    for (int k = 0; k < 3; ++k) {
        m[k] = Dual::seed(p.means[3 * size_t(i) + k], k);
        sigma[k] = exp(Dual::seed(p.logScales[3 * size_t(i) + k], 3 + k));
    }
    for (int k = 0; k < 4; ++k) q[k] = Dual::seed(p.quats[4 * size_t(i) + k], 6 + k);
    const Dual o = Dual(1.0) / (Dual(1.0) + exp(-Dual::seed(p.logits[i], 10)));
    Dual t[3];
    for (int r = 0; r < 3; ++r)
        t[r] = cam.view[0][r] * m[0] + cam.view[1][r] * m[1] + cam.view[2][r] * m[2] + cam.view[3][r];
    const M3<Dual> cov3 = computeCov3D(sigma, q);
    Dual cov2[3], center[2];
    computeCov2D(t, cov3, cam, cov2);
    const Dual det = cov2[0] * cov2[2] - cov2[1] * cov2[1];
    const Dual detInv = Dual(1.0) / det;
    projectToPixel(t, cam, center);
    const Dual v[3] = {m[0] - cam.pos[0], m[1] - cam.pos[1], m[2] - cam.pos[2]};
    const Dual inv = Dual(1.0) / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const Dual dir[3] = {v[0] * inv, v[1] * inv, v[2] * inv};
    std::vector<Dual> sh(size_t(p.K) * 3);
    for (uint32_t k = 0; k < p.K; ++k)
        for (int ch = 0; ch < 3; ++ch) sh[k * 3 + ch] = Dual::seed(p.sh[(size_t(i) * p.K + k) * 3 + ch], 11 + int(k));
    GSplat s;
    evalShD(p.degree, sh.data(), dir, s.col);
    s.cx = center[0];
    s.cy = center[1];
    s.a = cov2[2] * detInv;
    s.b = -cov2[1] * detInv;
    s.c = cov2[0] * detInv;
    s.o = o;
    return s;
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

// The vector-Jacobian product of the render for dL/dimage = G (H x W x 3): gradients in the order
// of oracle_diff_fd_f64, n x (11 + 3K). The forward is lossOf's (same culls, rect bounds, order,
// alpha cutoffs and termination); per pixel the composite's derivatives are taken in reverse:
// with T_i the transmittance before contributor i and B the colour of the contributors behind it,
// dC/dcol_i = alpha_i T_i and dC/dalpha_i = col_i T_i - B / (1 - alpha_i); alpha = o exp(power)
// unless clamped at 0.99 (then constant), power = -(a dx^2 + c dy^2) / 2 - b dx dy with
// dx = cx - x, dy = cy - y. The splat quantities' gradients are then carried to the raw
// parameters through the dual Jacobian of dualSplat.
void gradOf(const Params& p, const Camera<double>& cam, const float* G, double* out) {
    std::vector<DSplat> splats;
    splats.reserve(p.n);
    for (uint32_t i = 0; i < p.n; ++i) {
        DSplat s;
        if (splatOf(p, i, cam, &s)) splats.push_back(s);
    }
    std::stable_sort(splats.begin(), splats.end(), [](const DSplat& a, const DSplat& b) { return a.depth < b.depth; });
    // Per splat (in depth order): dL/d (cx, cy, a, b, c, o, col0, col1, col2).
    std::vector<double> gs(splats.size() * 9, 0.0);
    struct Hit {
        uint32_t s;
        double alpha, T, e, dx, dy;
        bool clamped;
    };
    std::vector<Hit> hits;
    for (uint32_t y = 0; y < cam.height; ++y) {
        const int ty = int(y / kTileSize);
        for (uint32_t x = 0; x < cam.width; ++x) {
            const int tx = int(x / kTileSize);
            hits.clear();
            double T = 1;
            for (uint32_t si = 0; si < splats.size(); ++si) {
                const DSplat& s = splats[si];
                if (tx < s.rect.minX || tx >= s.rect.maxX || ty < s.rect.minY || ty >= s.rect.maxY) continue;
                const double dx = s.cx - double(x), dy = s.cy - double(y);
                const double power = -0.5 * (s.a * dx * dx + s.c * dy * dy) - s.b * dx * dy;
                if (power > 0) continue;
                const double e = std::exp(power);
                const double oe = s.o * e;
                const double alpha = oe < double(kMaxAlpha) ? oe : double(kMaxAlpha);
                if (alpha < double(kMinAlpha)) continue;
                const double testT = T * (1 - alpha);
                if (testT < double(kMinTransmittance)) break;
                hits.push_back({si, alpha, T, e, dx, dy, !(oe < double(kMaxAlpha))});
                T = testT;
            }
            const float* g = G + (size_t(y) * cam.width + x) * 3;
            double B[3] = {0, 0, 0};
            for (size_t h = hits.size(); h-- > 0;) {
                const Hit& hit = hits[h];
                const DSplat& s = splats[hit.s];
                double* gsp = &gs[size_t(hit.s) * 9];
                double dAlpha = 0;
                for (int c = 0; c < 3; ++c) {
                    gsp[6 + c] += double(g[c]) * hit.alpha * hit.T;
                    dAlpha += double(g[c]) * (s.col[c] * hit.T - B[c] / (1 - hit.alpha));
                    B[c] += s.col[c] * hit.alpha * hit.T;
                }
                if (hit.clamped) continue;
                const double dPower = dAlpha * hit.alpha;
                gsp[0] += -dPower * (s.a * hit.dx + s.b * hit.dy);
                gsp[1] += -dPower * (s.c * hit.dy + s.b * hit.dx);
                gsp[2] += -0.5 * dPower * hit.dx * hit.dx;
                gsp[3] += -dPower * hit.dx * hit.dy;
                gsp[4] += -0.5 * dPower * hit.dy * hit.dy;
                gsp[5] += dAlpha * hit.e;
            }
        }
    }
    const uint32_t P = 11 + 3 * p.K;
    std::fill(out, out + size_t(p.n) * P, 0.0);
    Camera<Dual> camD;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) camD.view[c][r] = cam.view[c][r];
    for (int i = 0; i < 3; ++i) camD.pos[i] = cam.pos[i];
    camD.width = cam.width;
    camD.height = cam.height;
    camD.tanHalfFovX = cam.tanHalfFovX;
    camD.tanHalfFovY = cam.tanHalfFovY;
    camD.focalX = cam.focalX;
    camD.focalY = cam.focalY;
    camD.tilesX = cam.tilesX;
    camD.tilesY = cam.tilesY;
    for (size_t si = 0; si < splats.size(); ++si) {
        const double* gsp = &gs[si * 9];
        const GSplat d = dualSplat(p, splats[si].index, camD);
        const Dual* q[6] = {&d.cx, &d.cy, &d.a, &d.b, &d.c, &d.o};
        double* o = out + size_t(splats[si].index) * P;
        for (int k = 0; k < 11; ++k) {
            double v = 0;
            for (int j = 0; j < 6; ++j) v += gsp[j] * q[j]->d[k];
            for (int c = 0; c < 3; ++c) v += gsp[6 + c] * d.col[c].d[k];
            o[k] = v;
        }
        for (uint32_t k = 0; k < p.K; ++k)
            for (int c = 0; c < 3; ++c) o[11 + k * 3 + c] = gsp[6 + c] * d.col[c].d[11 + k];
    }
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

// The exact gradient of L for dL/dimage = gradImage (H x W x 3): n x (11 + 3K), the order of
// oracle_diff_fd_f64; culled gaussians get zeros.
void oracle_diff_grad_f64(uint32_t n, const double* means, const double* logScales, const double* quats,
                          const double* logits, const double* sh, uint32_t K, int32_t degree, const float viewInria[16],
                          uint32_t width, uint32_t height, float tanX, float tanY, const float* gradImage,
                          double* out) {
    const Params p{n, means, logScales, quats, logits, sh, K, degree};
    const Camera<double> cam = makeCamera<double>(viewInria, width, height, tanX, tanY);
    gradOf(p, cam, gradImage, out);
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
