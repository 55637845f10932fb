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

// --- Tiled form (large scenes) ------------------------------------------------------------------
// The same forward, composite and gradient as lossOf / gradOf, but each pixel walks only the
// splats whose tile rect holds its tile (per-tile lists in (depth, index) order), so the cost is
// that of a render instead of pixels x gaussians. The central differences re-render only the
// tiles the moved gaussian reaches before or after the move: every other pixel keeps its value,
// so the change of the loss is the change over those tiles.

struct Tiled {
    std::vector<DSplat> splats;    // visible, stable-sorted by depth (index order within a depth)
    std::vector<uint32_t> start;   // tiles + 1
    std::vector<uint32_t> list;    // positions in splats, per tile in depth order
};

Tiled buildTiled(const Params& p, const Camera<double>& cam) {
    Tiled t;
    t.splats.reserve(p.n);
    for (uint32_t i = 0; i < p.n; ++i) {
        DSplat s;
        if (splatOf(p, i, cam, &s)) t.splats.push_back(s);
    }
    std::stable_sort(t.splats.begin(), t.splats.end(),
                     [](const DSplat& a, const DSplat& b) { return a.depth < b.depth; });
    const uint32_t tiles = cam.tilesX * cam.tilesY;
    t.start.assign(tiles + 1, 0);
    for (const DSplat& s : t.splats)
        for (int ty = s.rect.minY; ty < s.rect.maxY; ++ty)
            for (int tx = s.rect.minX; tx < s.rect.maxX; ++tx) t.start[size_t(ty) * cam.tilesX + tx + 1]++;
    for (uint32_t i = 0; i < tiles; ++i) t.start[i + 1] += t.start[i];
    t.list.resize(t.start[tiles]);
    std::vector<uint32_t> cursor(t.start.begin(), t.start.end() - 1);
    for (uint32_t k = 0; k < t.splats.size(); ++k) {
        const DSplat& s = t.splats[k];
        for (int ty = s.rect.minY; ty < s.rect.maxY; ++ty)
            for (int tx = s.rect.minX; tx < s.rect.maxX; ++tx) t.list[cursor[size_t(ty) * cam.tilesX + tx]++] = k;
    }
    return t;
}

// The composite of one pixel over splats in order (lossOf's loop).
void shadePixel(const DSplat* const* seq, size_t count, uint32_t x, uint32_t y, double C[3]) {
    double T = 1;
    C[0] = C[1] = C[2] = 0;
    for (size_t j = 0; j < count; ++j) {
        const DSplat& s = *seq[j];
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
}

// Sum over the pixels of tile `tile` of |C - target| (image: written if not null).
double tileLoss(const Camera<double>& cam, uint32_t tile, const DSplat* const* seq, size_t count, const float* target,
                double* image) {
    const uint32_t tx = tile % cam.tilesX, ty = tile / cam.tilesX;
    const uint32_t x0 = tx * kTileSize, y0 = ty * kTileSize;
    const uint32_t x1 = std::min(cam.width, x0 + kTileSize), y1 = std::min(cam.height, y0 + kTileSize);
    double sum = 0;
    for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = x0; x < x1; ++x) {
            double C[3];
            shadePixel(seq, count, x, y, C);
            const size_t px = (size_t(y) * cam.width + x) * 3;
            for (int c = 0; c < 3; ++c) {
                if (image) image[px + c] = C[c];
                if (target) sum += std::fabs(C[c] - double(target[px + c]));
            }
        }
    return sum;
}

template <typename Fn>
void parallelTasks(uint64_t tasks, unsigned threads, Fn&& fn) {
    std::atomic<uint64_t> next{0};
    auto work = [&](unsigned worker) {
        for (uint64_t t = next++; t < tasks; t = next++) fn(t, worker);
    };
    std::vector<std::thread> pool;
    for (unsigned i = 1; i < threads; ++i) pool.emplace_back(work, i);
    work(0);
    for (std::thread& th : pool) th.join();
}

// The tile's splats as pointers, with gaussian g's splat left out and `moved` (if not null and
// its rect holds the tile) put in its (depth, index) place.
void tileSeq(const Tiled& t, const Camera<double>& cam, uint32_t tile, uint32_t g, const DSplat* moved,
             std::vector<const DSplat*>& out) {
    out.clear();
    const uint32_t tx = tile % cam.tilesX, ty = tile / cam.tilesX;
    const bool in = moved && int(tx) >= moved->rect.minX && int(tx) < moved->rect.maxX && int(ty) >= moved->rect.minY &&
                    int(ty) < moved->rect.maxY;
    bool placed = !in;
    for (uint32_t j = t.start[tile]; j < t.start[tile + 1]; ++j) {
        const DSplat& s = t.splats[t.list[j]];
        if (s.index == g) continue;
        if (!placed && (s.depth > moved->depth || (s.depth == moved->depth && s.index > g))) {
            out.push_back(moved);
            placed = true;
        }
        out.push_back(&s);
    }
    if (!placed) out.push_back(moved);
}

}  // namespace

extern "C" {

// The tiled FP64 references for large scenes (same forward as oracle_diff_render_f64, each pixel
// walking only its tile's list). image (H x W x 3) and *loss (L1 against target) if not null;
// for the `count` gaussians `which`: grad (count x (11 + 3K)) = rows of oracle_diff_grad_f64 for
// dL/dimage = gradImage, and fd (count x (11 + 3K)) = oracle_diff_fd_f64's central differences
// with step h, each from a re-render of the tiles the gaussian reaches at x, x + h or x - h.
void oracle_diff_local_f64(uint32_t n, const double* means, const double* logScales, const double* quats,
                           const double* logits, const double* sh, uint32_t K, int32_t degree,
                           const float viewInria[16], uint32_t width, uint32_t height, float tanX, float tanY,
                           const float* target, const float* gradImage, uint32_t count, const uint32_t* which,
                           double h, double* image, double* lossOut, double* grad, double* fd, int32_t threadsArg) {
    const Camera<double> cam = makeCamera<double>(viewInria, width, height, tanX, tanY);
    const Params p{n, means, logScales, quats, logits, sh, K, degree};
    const unsigned threads = threadsArg > 0 ? unsigned(threadsArg) : std::max(1u, std::thread::hardware_concurrency());
    const Tiled t = buildTiled(p, cam);
    const uint32_t tiles = cam.tilesX * cam.tilesY;
    const uint32_t P = 11 + 3 * K;
    auto seqOf = [&](uint32_t tile, std::vector<const DSplat*>& out) {
        out.clear();
        for (uint32_t j = t.start[tile]; j < t.start[tile + 1]; ++j) out.push_back(&t.splats[t.list[j]]);
    };
    if (image || lossOut) {
        std::vector<double> partial(tiles, 0.0);
        parallelTasks(tiles, threads, [&](uint64_t tile, unsigned) {
            std::vector<const DSplat*> seq;
            seqOf(uint32_t(tile), seq);
            partial[tile] = tileLoss(cam, uint32_t(tile), seq.data(), seq.size(), target, image);
        });
        double sum = 0;
        for (double v : partial) sum += v;
        if (lossOut) *lossOut = target ? sum / (3.0 * double(width) * double(height)) : 0.0;
    }
    std::vector<int32_t> slot(n, -1);
    for (uint32_t k = 0; k < count; ++k) slot[which[k]] = int32_t(k);
    if (grad && gradImage) {
        // gradOf's per-pixel reverse pass over the tile lists, accumulated only for `which`.
        std::vector<std::vector<double>> local(threads, std::vector<double>(size_t(count) * 9, 0.0));
        parallelTasks(tiles, threads, [&](uint64_t tileIdx, unsigned worker) {
            const uint32_t tile = uint32_t(tileIdx);
            std::vector<double>& gs = local[worker];
            struct Hit {
                const DSplat* s;
                double alpha, T, e, dx, dy;
                bool clamped;
            };
            std::vector<Hit> hits;
            const uint32_t tx = tile % cam.tilesX, ty = tile / cam.tilesX;
            const uint32_t x0 = tx * kTileSize, y0 = ty * kTileSize;
            const uint32_t x1 = std::min(cam.width, x0 + kTileSize), y1 = std::min(cam.height, y0 + kTileSize);
            for (uint32_t y = y0; y < y1; ++y)
                for (uint32_t x = x0; x < x1; ++x) {
                    hits.clear();
                    double T = 1;
                    for (uint32_t j = t.start[tile]; j < t.start[tile + 1]; ++j) {
                        const DSplat& s = t.splats[t.list[j]];
                        const double dx = s.cx - double(x), dy = s.cy - double(y);
                        const double power = -0.5 * (s.a * dx * dx + s.c * dy * dy) - s.b * dx * dy;
                        if (power > 0) continue;
                        const double e = std::exp(power);
                        const double oe = s.o * e;
                        const double alpha = oe < double(kMaxAlpha) ? oe : double(kMaxAlpha);
                        if (alpha < double(kMinAlpha)) continue;
                        const double testT = T * (1 - alpha);
                        if (testT < double(kMinTransmittance)) break;
                        hits.push_back({&s, alpha, T, e, dx, dy, !(oe < double(kMaxAlpha))});
                        T = testT;
                    }
                    const float* gi = gradImage + (size_t(y) * cam.width + x) * 3;
                    double B[3] = {0, 0, 0};
                    for (size_t hh = hits.size(); hh-- > 0;) {
                        const Hit& hit = hits[hh];
                        const DSplat& s = *hit.s;
                        const int32_t k = slot[s.index];
                        double dAlpha = 0;
                        for (int c = 0; c < 3; ++c) {
                            if (k >= 0) gs[size_t(k) * 9 + 6 + c] += double(gi[c]) * hit.alpha * hit.T;
                            dAlpha += double(gi[c]) * (s.col[c] * hit.T - B[c] / (1 - hit.alpha));
                            B[c] += s.col[c] * hit.alpha * hit.T;
                        }
                        if (k < 0 || hit.clamped) continue;
                        double* gsp = &gs[size_t(k) * 9];
                        const double dPower = dAlpha * hit.alpha;
                        gsp[0] += -dPower * (s.a * hit.dx + s.b * hit.dy);
                        gsp[1] += -dPower * (s.c * hit.dy + s.b * hit.dx);
                        gsp[2] += -0.5 * dPower * hit.dx * hit.dx;
                        gsp[3] += -dPower * hit.dx * hit.dy;
                        gsp[4] += -0.5 * dPower * hit.dy * hit.dy;
                        gsp[5] += dAlpha * hit.e;
                    }
                }
        });
        std::vector<double> gs(size_t(count) * 9, 0.0);
        for (const auto& l : local)
            for (size_t i = 0; i < gs.size(); ++i) gs[i] += l[i];
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
        std::fill(grad, grad + size_t(count) * P, 0.0);
        for (uint32_t k = 0; k < count; ++k) {
            DSplat probe;
            if (!splatOf(p, which[k], cam, &probe)) continue;  // culled: zeros
            const double* gsp = &gs[size_t(k) * 9];
            const GSplat d = dualSplat(p, which[k], camD);
            const Dual* q[6] = {&d.cx, &d.cy, &d.a, &d.b, &d.c, &d.o};
            double* o = grad + size_t(k) * P;
            for (int j = 0; j < 11; ++j) {
                double v = 0;
                for (int m = 0; m < 6; ++m) v += gsp[m] * q[m]->d[j];
                for (int c = 0; c < 3; ++c) v += gsp[6 + c] * d.col[c].d[j];
                o[j] = v;
            }
            for (uint32_t kk = 0; kk < K; ++kk)
                for (int c = 0; c < 3; ++c) o[11 + kk * 3 + c] = gsp[6 + c] * d.col[c].d[11 + kk];
        }
    }
    if (fd && target) {
        const double norm = 3.0 * double(width) * double(height);
        parallelTasks(uint64_t(count) * P, threads, [&](uint64_t task, unsigned) {
            const uint32_t g = which[task / P], k = uint32_t(task % P);
            std::vector<double> m(means + 3 * size_t(g), means + 3 * size_t(g) + 3);
            std::vector<double> ls(logScales + 3 * size_t(g), logScales + 3 * size_t(g) + 3);
            std::vector<double> q(quats + 4 * size_t(g), quats + 4 * size_t(g) + 4);
            double lg = logits[g];
            std::vector<double> shg(sh + size_t(g) * K * 3, sh + size_t(g + 1) * K * 3);
            // A one-gaussian parameter set addressing gaussian g's own (moved) values.
            auto splatAt = [&](double delta, DSplat* out) {
                double* x;
                if (k < 3) x = &m[k];
                else if (k < 6) x = &ls[k - 3];
                else if (k < 10) x = &q[k - 6];
                else if (k < 11) x = &lg;
                else x = &shg[k - 11];
                const double x0 = *x;
                *x = x0 + delta;
                const Params one{1, m.data(), ls.data(), q.data(), &lg, shg.data(), K, degree};
                const bool vis = splatOf(one, 0, cam, out);
                *x = x0;
                if (vis) out->index = g;
                return vis;
            };
            DSplat plus, minus, base;
            const bool vp = splatAt(h, &plus), vm = splatAt(-h, &minus), v0 = splatAt(0.0, &base);
            // The tiles any of the three reaches.
            std::vector<uint32_t> touched;
            for (const DSplat* s : {v0 ? &base : nullptr, vp ? &plus : nullptr, vm ? &minus : nullptr}) {
                if (!s) continue;
                for (int ty = s->rect.minY; ty < s->rect.maxY; ++ty)
                    for (int tx = s->rect.minX; tx < s->rect.maxX; ++tx) touched.push_back(uint32_t(ty) * cam.tilesX + tx);
            }
            std::sort(touched.begin(), touched.end());
            touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
            std::vector<const DSplat*> seq;
            double lp = 0, lm = 0;
            for (uint32_t tile : touched) {
                tileSeq(t, cam, tile, g, vp ? &plus : nullptr, seq);
                lp += tileLoss(cam, tile, seq.data(), seq.size(), target, nullptr);
                tileSeq(t, cam, tile, g, vm ? &minus : nullptr, seq);
                lm += tileLoss(cam, tile, seq.data(), seq.size(), target, nullptr);
            }
            fd[task] = (lp - lm) / norm / (2 * h);
        });
    }
}

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
