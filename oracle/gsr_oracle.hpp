// CPU oracle for the splat-tiles forward path. TEST INFRASTRUCTURE ONLY: tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, as the checker
// or as the timed CPU baseline, never as part of the product.
//
// A from-scratch C++20 restatement (no glm) of the reference's CPU INRIA forward pass:
//   - per-gaussian math      src/splat/splat_math.{h,cpp}
//   - tight tile coverage    src/splat/tile_coverage.{h,cpp}
//   - depth sort             src/splat/depth_sort.{h,cpp}
//   - reference renderer     src/splat/reference_renderer.{h,cpp}
//   - fp16 scene decode      src/scene/half.h, src/scene/scene_compression.cpp:34-49, 382-460
//   - the GPU list contract  tests/test_tile_raster.cpp:332-382 (cpuTileList)
// in the reference's written FP32 operation order (every matrix product is spelled out,
// summed left to right; build with -ffp-contract=off). The per-gaussian math is templated
// on the float type: Real = double is the FP64 twin used to state the FP32 error.
//
// Deviations from the reference, each deliberate:
//   - the natural log of the tight-bounds threshold is det_logf() below (a double-precision
//     atanh series rounded once to float) instead of the platform libm logf, so that the
//     HIP kernels can reproduce it bit for bit; it is the correctly rounded logf except
//     within ~1e-16 relative of a rounding midpoint (pinned against libm in the tests);
//   - matrix products with a structurally zero operand (the Jacobian's two zeros) omit
//     those terms; glm's summation order for the rest is restated left to right.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {

// --- INRIA constants (splat_math.h:25-40, gaussian_scene.h:103) ---------------------------
inline constexpr float kNearCullDepth = 0.2f;
inline constexpr float kLowPassFilter = 0.3f;
inline constexpr float kFrustumClampFactor = 1.3f;
inline constexpr float kMaxAlpha = 0.99f;
inline constexpr float kMinAlpha = 1.0f / 255.0f;
inline constexpr float kMinTransmittance = 1e-4f;
inline constexpr int kTileSize = 16;
inline constexpr float kShC0 = 0.28209479177387814f;
inline constexpr float kShC1 = 0.4886025119029199f;
inline constexpr float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                                   -1.0925484305920792f, 0.5462742152960396f};
inline constexpr float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                                   0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                                   -0.5900435899266435f};

enum CullReason : uint32_t { kVisible = 0, kNearPlane = 1, kSingularCov = 2, kOffScreen = 3, kLowOpacity = 4 };
enum Bounds : uint32_t { kRect = 0, kTight = 1, kCutoffOnly = 2 };
enum CoverageBound : uint32_t { kBoundTile = 0, kBoundRadius = 1 };

inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// SplatCamera (splat_math.h:43-50). view[c][r] is column-major, as glm::mat4.
template <typename Real>
struct Camera {
    Real view[4][4];
    Real pos[3];
    uint32_t width = 0, height = 0;
    Real tanHalfFovX = 1, tanHalfFovY = 1, focalX = 1, focalY = 1;
    uint32_t tilesX = 0, tilesY = 0;
};

// makeSplatCamera (splat_math.cpp:9-27): position = -(R^T t), focal = size / (2 tan).
template <typename Real>
Camera<Real> makeCamera(const float viewInria[16], uint32_t width, uint32_t height, float tanX, float tanY) {
    Camera<Real> c;
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) c.view[col][row] = Real(viewInria[col * 4 + row]);
    // (R^T t)_i = sum_k R[i][k] t_k with R[col][row] = view[col][row]: row i of R^T is column i of R.
    for (int i = 0; i < 3; ++i) {
        const Real s = c.view[i][0] * c.view[3][0] + c.view[i][1] * c.view[3][1] + c.view[i][2] * c.view[3][2];
        c.pos[i] = -s;
    }
    c.width = width;
    c.height = height;
    c.tanHalfFovX = Real(tanX);
    c.tanHalfFovY = Real(tanY);
    c.focalX = Real(width) / (Real(2) * Real(tanX));
    c.focalY = Real(height) / (Real(2) * Real(tanY));
    c.tilesX = (width + kTileSize - 1) / kTileSize;
    c.tilesY = (height + kTileSize - 1) / kTileSize;
    return c;
}

// splatCameraFromViewer (splat_math.cpp:29-42): viewInria = diag(1,-1,-1,1) * view * model;
// tan(fov/2) from proj[0][0] and |proj[1][1]|. glm's mat4 product, summed left to right.
void cameraFromViewer(const float view[16], const float proj[16], const float model[16], uint32_t width,
                      uint32_t height, float viewInriaOut[16], float* tanX, float* tanY);

// --- The per-gaussian math (splat_math.cpp:44-193) ---------------------------------------
template <typename Real>
struct M3 {
    Real m[3][3];  // [col][row]
};

// quaternionToMatrix (splat_math.cpp:44-53), q = (w, x, y, z).
template <typename Real>
M3<Real> quaternionToMatrix(const Real q[4]) {
    const Real r = q[0], x = q[1], y = q[2], z = q[3];
    const Real one = 1, two = 2;
    M3<Real> R;
    R.m[0][0] = one - two * (y * y + z * z);
    R.m[0][1] = two * (x * y + r * z);
    R.m[0][2] = two * (x * z - r * y);
    R.m[1][0] = two * (x * y - r * z);
    R.m[1][1] = one - two * (x * x + z * z);
    R.m[1][2] = two * (y * z + r * x);
    R.m[2][0] = two * (x * z + r * y);
    R.m[2][1] = two * (y * z - r * x);
    R.m[2][2] = one - two * (x * x + y * y);
    return R;
}

// computeCov3D (splat_math.cpp:55-63): M = R diag(sigma), Sigma = M M^T,
// Sigma[j][i] = M[0][i] M[0][j] + M[1][i] M[1][j] + M[2][i] M[2][j].
template <typename Real>
M3<Real> computeCov3D(const Real sigma[3], const Real q[4]) {
    const M3<Real> R = quaternionToMatrix(q);
    M3<Real> M;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) M.m[c][r] = R.m[c][r] * sigma[c];
    M3<Real> S;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) S.m[j][i] = M.m[0][i] * M.m[0][j] + M.m[1][i] * M.m[1][j] + M.m[2][i] * M.m[2][j];
    return S;
}

// computeCov2D (splat_math.cpp:65-91): J clamped at 1.3 tan(fov/2), T = J W, cov = T Sigma T^T + 0.3 I.
// Returns (a, b, c) of [[a b][b c]].
template <typename Real>
void computeCov2D(const Real t[3], const M3<Real>& cov3D, const Camera<Real>& cam, Real out[3]) {
    const Real limX = Real(kFrustumClampFactor) * cam.tanHalfFovX;
    const Real limY = Real(kFrustumClampFactor) * cam.tanHalfFovY;
    const Real txtz = t[0] / t[2];
    const Real tytz = t[1] / t[2];
    const Real cx = txtz < -limX ? -limX : txtz;  // max(-lim, v)
    const Real cy = tytz < -limY ? -limY : tytz;
    const Real tx = (limX < cx ? limX : cx) * t[2];  // min(lim, .)
    const Real ty = (limY < cy ? limY : cy) * t[2];
    const Real tz = t[2];
    const Real j00 = cam.focalX / tz;
    const Real j20 = -(cam.focalX * tx) / (tz * tz);
    const Real j11 = cam.focalY / tz;
    const Real j21 = -(cam.focalY * ty) / (tz * tz);
    // T = J W (W = the rotation block of viewInria, W[col][row] = view[col][row]).
    // T[j][0] = J00 W[j][0] + J20 W[j][2]; T[j][1] = J11 W[j][1] + J21 W[j][2].
    Real T[3][2];
    for (int j = 0; j < 3; ++j) {
        T[j][0] = j00 * cam.view[j][0] + j20 * cam.view[j][2];
        T[j][1] = j11 * cam.view[j][1] + j21 * cam.view[j][2];
    }
    // U = T Sigma: U[j][i] = T[0][i] S[j][0] + T[1][i] S[j][1] + T[2][i] S[j][2].
    Real U[3][2];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 2; ++i)
            U[j][i] = T[0][i] * cov3D.m[j][0] + T[1][i] * cov3D.m[j][1] + T[2][i] * cov3D.m[j][2];
    // cov = U T^T: cov[j][i] = U[0][i] T[0][j] + U[1][i] T[1][j] + U[2][i] T[2][j].
    const Real c00 = U[0][0] * T[0][0] + U[1][0] * T[1][0] + U[2][0] * T[2][0];
    const Real c01 = U[0][1] * T[0][0] + U[1][1] * T[1][0] + U[2][1] * T[2][0];
    const Real c11 = U[0][1] * T[0][1] + U[1][1] * T[1][1] + U[2][1] * T[2][1];
    out[0] = c00 + Real(kLowPassFilter);
    out[1] = c01;
    out[2] = c11 + Real(kLowPassFilter);
}

// projectToPixel (splat_math.cpp:93-100).
template <typename Real>
void projectToPixel(const Real t[3], const Camera<Real>& cam, Real out[2]) {
    const Real ndcX = t[0] / (t[2] * cam.tanHalfFovX);
    const Real ndcY = t[1] / (t[2] * cam.tanHalfFovY);
    out[0] = ((ndcX + Real(1)) * Real(cam.width) - Real(1)) * Real(0.5);
    out[1] = ((ndcY + Real(1)) * Real(cam.height) - Real(1)) * Real(0.5);
}

// conicAndRadius (splat_math.cpp:102-114). false for det == 0.
template <typename Real>
bool conicAndRadius(const Real cov[3], Real conic[3], Real* radius) {
    const Real det = cov[0] * cov[2] - cov[1] * cov[1];
    if (det == Real(0)) return false;
    const Real detInv = Real(1) / det;
    conic[0] = cov[2] * detInv;
    conic[1] = -cov[1] * detInv;
    conic[2] = cov[0] * detInv;
    const Real mid = Real(0.5) * (cov[0] + cov[2]);
    const Real disc = mid * mid - det;
    const Real root = std::sqrt(disc > Real(0.1f) ? disc : Real(0.1f));
    const Real lambda1 = mid + root;
    const Real lambda2 = mid - root;
    *radius = std::ceil(Real(3) * std::sqrt(lambda1 > lambda2 ? lambda1 : lambda2));
    return true;
}

struct TileRect {
    int minX = 0, minY = 0, maxX = 0, maxY = 0;
    bool empty() const { return maxX <= minX || maxY <= minY; }
};

// computeTileRect (splat_math.cpp:116-127): INRIA getRect with int() truncation.
inline TileRect computeTileRect(float px, float py, float radius, uint32_t tilesX, uint32_t tilesY) {
    const int gx = int(tilesX), gy = int(tilesY);
    const int r = int(radius);
    const float S = float(kTileSize);
    auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
    TileRect t;
    t.minX = clampi(int((px - float(r)) / S), gx);
    t.minY = clampi(int((py - float(r)) / S), gy);
    t.maxX = clampi(int((px + float(r) + float(kTileSize - 1)) / S), gx);
    t.maxY = clampi(int((py + float(r) + float(kTileSize - 1)) / S), gy);
    return t;
}

// evalSH (splat_math.cpp:129-151), one colour channel at a time (glm's vec3 ops are
// per component). sh = K RGB triplets.
template <typename Real>
void evalSH(int degree, const float* sh, const Real dir[3], Real out[3]) {
    for (int ch = 0; ch < 3; ++ch) {
        auto c = [&](int k) { return Real(sh[k * 3 + ch]); };
        Real result = Real(kShC0) * c(0);
        if (degree > 0) {
            const Real x = dir[0], y = dir[1], z = dir[2];
            const Real C1 = Real(kShC1);
            result = result - C1 * y * c(1) + C1 * z * c(2) - C1 * x * c(3);
            if (degree > 1) {
                const Real xx = x * x, yy = y * y, zz = z * z;
                const Real xy = x * y, yz = y * z, xz = x * z;
                result = result + Real(kShC2[0]) * xy * c(4) + Real(kShC2[1]) * yz * c(5) +
                         Real(kShC2[2]) * (Real(2) * zz - xx - yy) * c(6) + Real(kShC2[3]) * xz * c(7) +
                         Real(kShC2[4]) * (xx - yy) * c(8);
                if (degree > 2) {
                    result = result + Real(kShC3[0]) * y * (Real(3) * xx - yy) * c(9) +
                             Real(kShC3[1]) * xy * z * c(10) + Real(kShC3[2]) * y * (Real(4) * zz - xx - yy) * c(11) +
                             Real(kShC3[3]) * z * (Real(2) * zz - Real(3) * xx - Real(3) * yy) * c(12) +
                             Real(kShC3[4]) * x * (Real(4) * zz - xx - yy) * c(13) +
                             Real(kShC3[5]) * z * (xx - yy) * c(14) + Real(kShC3[6]) * x * (xx - Real(3) * yy) * c(15);
                }
            }
        }
        result = result + Real(0.5);
        out[ch] = result < Real(0) ? Real(0) : result;  // glm::max(result, 0)
    }
}

// ProjectedSplat (splat_math.h:115-123), 48 bytes.
struct Splat {
    float center[2];
    float depth;
    uint32_t cullReason;
    float conicOpacity[4];
    float color[3];
    float radius;
};
static_assert(sizeof(Splat) == 48, "ProjectedSplat is 48 bytes");

template <typename Real>
struct SplatT {
    Real center[2] = {0, 0};
    Real depth = 0;
    uint32_t cullReason = 0;
    Real conic[3] = {0, 0, 0};
    Real opacity = 0;
    Real color[3] = {0, 0, 0};
    Real radius = 0;
};

// preprocessGaussian (splat_math.cpp:153-193). gpuCull = the GPU twin's extra culls
// (splat_preprocess.comp:294-309): !(z > 0.2) instead of z <= 0.2 (NaN-safe) and the
// low-opacity cull (o < 1/255) before the covariance.
template <typename Real>
SplatT<Real> preprocessGaussian(const float pa[4], const float sigmaF[3], const float qF[4], const float* sh,
                                int degree, const Camera<Real>& cam, bool gpuCull) {
    SplatT<Real> s;
    const Real mean[3] = {Real(pa[0]), Real(pa[1]), Real(pa[2])};
    // transformPoint (splat_math.h:65-69), written out per component.
    Real t[3];
    for (int r = 0; r < 3; ++r)
        t[r] = cam.view[0][r] * mean[0] + cam.view[1][r] * mean[1] + cam.view[2][r] * mean[2] + cam.view[3][r];
    s.depth = t[2];
    const bool nearCulled = gpuCull ? !(t[2] > Real(kNearCullDepth)) : (t[2] <= Real(kNearCullDepth));
    if (nearCulled) {
        s.cullReason = kNearPlane;
        return s;
    }
    if (gpuCull && Real(pa[3]) < Real(kMinAlpha)) {
        s.cullReason = kLowOpacity;
        return s;
    }
    const Real sigma[3] = {Real(sigmaF[0]), Real(sigmaF[1]), Real(sigmaF[2])};
    const Real q[4] = {Real(qF[0]), Real(qF[1]), Real(qF[2]), Real(qF[3])};
    const M3<Real> cov3D = computeCov3D(sigma, q);
    Real cov2D[3];
    computeCov2D(t, cov3D, cam, cov2D);
    Real conic[3], radius = 0;
    if (!conicAndRadius(cov2D, conic, &radius)) {
        s.cullReason = kSingularCov;
        return s;
    }
    Real center[2];
    projectToPixel(t, cam, center);
    if (computeTileRect(float(center[0]), float(center[1]), float(radius), cam.tilesX, cam.tilesY).empty()) {
        s.center[0] = center[0];
        s.center[1] = center[1];
        s.cullReason = kOffScreen;
        return s;
    }
    // dir = normalize(mean - camPos) = v * (1 / sqrt(dot(v, v))) (glm::normalize).
    const Real v[3] = {mean[0] - cam.pos[0], mean[1] - cam.pos[1], mean[2] - cam.pos[2]};
    const Real inv = Real(1) / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const Real dir[3] = {v[0] * inv, v[1] * inv, v[2] * inv};
    evalSH(degree, sh, dir, s.color);
    s.center[0] = center[0];
    s.center[1] = center[1];
    s.conic[0] = conic[0];
    s.conic[1] = conic[1];
    s.conic[2] = conic[2];
    s.opacity = Real(pa[3]);
    s.radius = radius;
    s.cullReason = kVisible;
    return s;
}

inline Splat toSplat(const SplatT<float>& s) {
    Splat o;
    o.center[0] = s.center[0];
    o.center[1] = s.center[1];
    o.depth = s.depth;
    o.cullReason = s.cullReason;
    o.conicOpacity[0] = s.conic[0];
    o.conicOpacity[1] = s.conic[1];
    o.conicOpacity[2] = s.conic[2];
    o.conicOpacity[3] = s.opacity;
    o.color[0] = s.color[0];
    o.color[1] = s.color[1];
    o.color[2] = s.color[2];
    o.radius = s.radius;
    return o;
}

// Opacity at a pixel (splat_math.h:147-156), or 0 where INRIA skips.
inline float splatAlphaAt(const Splat& s, float px, float py) {
    const float dx = s.center[0] - px;
    const float dy = s.center[1] - py;
    const float* co = s.conicOpacity;
    const float power = -0.5f * (co[0] * dx * dx + co[2] * dy * dy) - co[1] * dx * dy;
    if (power > 0.0f) return 0.0f;
    const float e = co[3] * std::exp(power);
    const float alpha = kMaxAlpha < e ? kMaxAlpha : e;  // fmin for non-NaN
    if (alpha < kMinAlpha) return 0.0f;
    return alpha;
}

// --- Deterministic float log --------------------------------------------------------------
// log(x) = e ln2 + 2 atanh(s), s = (m - 1) / (m + 1), m in [sqrt(1/2), sqrt(2)), atanh by
// its Taylor series to s^21 (|s| <= 0.1716: truncation < 1e-17 relative), all in double
// with + - * / only, then one rounding to float.
float det_logf(float x);

// --- Tight tile coverage (tile_coverage.cpp:89-209) ----------------------------------------
struct CutoffEllipse {
    float cx = 0, cy = 0;
    float A = 0, B = 0, C = 0, D = 0;
    float t = -1.0f;
    float exPad = 0, eyPad = 0;
    bool valid = false;
};
struct BinRect {
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
};
struct Coverage {
    BinRect bins;
    CutoffEllipse e;
    bool useEllipse = false;
};
struct RowBounds {
    float x0 = 0, x1 = 0;
    bool empty = true;
};

CutoffEllipse makeCutoffEllipse(float cx, float cy, const float conic[3], float opacity);
BinRect rectToBins(const TileRect& r, int S, int gridX, int gridY);
Coverage makeCoverage(uint32_t mode, const TileRect& rect16, int S, int gridX, int gridY, float cx, float cy,
                      const float conic[3], float opacity);
void coverageRows(const Coverage& cv, int S, int* r0, int* r1);
RowBounds ellipseRowBounds(const CutoffEllipse& e, int row, int S);
void coverageRowSpan(const Coverage& cv, int row, int S, int* x0, int* x1);
uint32_t coverageCount(const Coverage& cv, int S);
Coverage splatCoverage(uint32_t mode, const Splat& s, uint32_t tilesX16, uint32_t tilesY16, int S, int gridX,
                       int gridY);
bool coverageIsBorderline(const Coverage& cv, int S, float eps);

// --- Depth sort (depth_sort.cpp:17-75) -----------------------------------------------------
// Stable LSD radix sort of (key, value) pairs, 4 x 8 bits, trivial passes skipped. Returns
// the number of passes that moved data.
uint32_t radixSortPairs(std::vector<uint32_t>& keys, std::vector<uint32_t>& values);

// --- fp16 (half.h:32-79) and the f16 record decode (scene_compression.cpp:382-415, 432-435) --
uint16_t floatToHalf(float value);
float halfToFloat(uint16_t h);

// --- StopThePop ray depth (splat_math.h:134-143, splat_math.cpp:195-205) --------------------
// P = Sigma^-1 in INRIA camera space = sum_k w_k M_k M_k^T, M = mat3(view) R (columns: the
// gaussian's axes), w = 1 / s^2 with s = max(sigma, max(1e-4 smax, 1e-12)); b = P t. The depth of
// the gaussian's maximum along the ray t (vx, vy, 1) is (v . b) / (v^T P v).
struct RayDepth {
    float p00 = 0, p01 = 0, p02 = 0, p11 = 0, p12 = 0, p22 = 0, b0 = 0, b1 = 0, b2 = 0;
    float depthAt(float vx, float vy) const {
        const float vPv = p00 * vx * vx + 2.0f * p01 * vx * vy + 2.0f * p02 * vx + p11 * vy * vy + 2.0f * p12 * vy + p22;
        const float vb = b0 * vx + b1 * vy + b2;
        return vb / vPv;
    }
};
RayDepth computeRayDepth(const float sigma[3], const float q[4], const float tCam[3], const float view[4][4]);

// --- Packed render splats (render_splat.h:1-69, render_splat.glsl:1-75) --------------------
// p32 (8 words): cx, cy, a, b | c, opacity, h(r) | h(g) << 16, h(b) | radius << 16.
// p24 (6 words): cx, cy | h(a S) | h(b S) << 16, h(c S) | h(opacity) << 16 | colour as p32,
// S = 2^12; radius = min(radius, 65535) as an integer.
constexpr uint32_t kRenderSplatFull = 0, kRenderSplatP32 = 1, kRenderSplatP24 = 2;
constexpr float kConicScaleP24 = 4096.0f;
void encodeRenderSplat(uint32_t mode, const Splat& s, uint32_t* words);
// What the packed record gives back: p24's conic and opacity (the preprocess counts tiles and the
// duplicate and render passes work with these), and with colour = false, the colour untouched.
void applyRenderSplat(uint32_t mode, Splat& s, bool colour);

}  // namespace oracle
