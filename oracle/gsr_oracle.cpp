// CPU oracle, non-template parts. TEST INFRASTRUCTURE ONLY (see gsr_oracle.hpp).
#include "gsr_oracle.hpp"

#include <algorithm>
#include <array>
#include <cstring>
#include <limits>

namespace oracle {

void cameraFromViewer(const float view[16], const float proj[16], const float model[16], uint32_t width,
                      uint32_t height, float out[16], float* tanX, float* tanY) {
    (void)width;
    (void)height;
    // glm mat4 product: R[j][i] = A[0][i] B[j][0] + A[1][i] B[j][1] + A[2][i] B[j][2] + A[3][i] B[j][3].
    auto mul = [](const float* A, const float* B, float* R) {
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i)
                R[j * 4 + i] = A[0 * 4 + i] * B[j * 4 + 0] + A[1 * 4 + i] * B[j * 4 + 1] + A[2 * 4 + i] * B[j * 4 + 2] +
                               A[3 * 4 + i] * B[j * 4 + 3];
    };
    float flip[16] = {1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1};
    float fv[16];
    mul(flip, view, fv);
    mul(fv, model, out);
    *tanX = 1.0f / proj[0];
    *tanY = 1.0f / std::fabs(proj[5]);
}

float det_logf(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -std::numeric_limits<float>::infinity() : std::numeric_limits<float>::quiet_NaN();
    if (std::isinf(x)) return x;
    uint32_t b = fbits(x);
    int e = 0;
    if ((b >> 23) == 0) {  // subnormal: scale by 2^24 (exact)
        b = fbits(x * 16777216.0f);
        e = -24;
    }
    e += int(b >> 23) - 127;
    double m = double(bitsf((b & 0x7FFFFFu) | 0x3F800000u));  // [1, 2)
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    const double r = 2.0 * s * p;
    return float(double(e) * 0.6931471805599453 + r);
}

// --- Tight coverage ------------------------------------------------------------------------

CutoffEllipse makeCutoffEllipse(float cx, float cy, const float conic[3], float opacity) {
    CutoffEllipse e;
    e.cx = cx;
    e.cy = cy;
    const float A = conic[0], B = conic[1], C = conic[2];
    const float D = A * C - B * B;
    e.A = A;
    e.B = B;
    e.C = C;
    e.D = D;
    e.valid = A > 0.0f && C > 0.0f && D > 0.0f;
    if (!e.valid) return e;
    if (!(opacity >= kMinAlpha)) return e;
    const float qmax = 2.0f * det_logf(255.0f * opacity);
    const float rho = std::fabs(B) / std::sqrt(A * C);
    const float omr = 1.0f - rho;
    const float K = 2.0f / (omr < 1e-30f ? 1e-30f : omr);  // max(1 - rho, 1e-30)
    const float qm1 = qmax < 1.0f ? 1.0f : qmax;             // max(qmax, 1)
    const float t = qmax + (1e-3f + 2e-5f * K) * qm1 + 2e-3f;
    if (!(t < 1e30f)) {
        e.valid = false;
        return e;
    }
    e.t = t;
    const float ex = std::sqrt(t * C / D);
    const float ey = std::sqrt(t * A / D);
    e.exPad = ex + (0.02f + 2e-3f * ex);
    e.eyPad = ey + (0.02f + 2e-3f * ey);
    return e;
}

namespace {

float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
int floorDivClamped(float v, int S) { return int(std::floor(clampf(v, -65536.0f, 1048576.0f) / float(S))); }

void ellipseRows(const CutoffEllipse& e, int S, const BinRect& b, int* r0, int* r1) {
    if (e.t < 0.0f) {
        *r0 = 0;
        *r1 = 0;
        return;
    }
    const float y0 = std::ceil(e.cy - e.eyPad);
    const float y1 = std::floor(e.cy + e.eyPad);
    const int lo = std::max(b.y0, floorDivClamped(y0, S));
    const int hi = std::min(b.y1, floorDivClamped(y1, S) + 1);
    *r0 = lo;
    *r1 = std::max(lo, hi);
}

bool nearInteger(float v, float eps) { return std::fabs(v - std::round(v)) < eps; }

}  // namespace

RowBounds ellipseRowBounds(const CutoffEllipse& e, int row, int S) {
    RowBounds rb;
    const float ey0 = float(row * S) - e.cy;
    const float ey1 = float(row * S + S - 1) - e.cy;
    const float nep = -e.eyPad;
    const float lo = ey0 < nep ? nep : ey0;        // max(ey0, -eyPad)
    const float hi = e.eyPad < ey1 ? e.eyPad : ey1;  // min(ey1, eyPad)
    if (!(lo <= hi)) return rb;
    const float exTrue = std::sqrt(e.t * e.C / e.D);
    const float eyRight = clampf(-e.B * exTrue / e.C, lo, hi);
    const float eyLeft = clampf(e.B * exTrue / e.C, lo, hi);
    const float At = e.A * e.t;
    const float dr = At - e.D * eyRight * eyRight;
    const float dl = At - e.D * eyLeft * eyLeft;
    const float xRight = (-e.B * eyRight + std::sqrt(0.0f < dr ? dr : 0.0f)) / e.A;
    const float xLeft = (-e.B * eyLeft - std::sqrt(0.0f < dl ? dl : 0.0f)) / e.A;
    const float pad = 0.02f + 2e-3f * exTrue;
    rb.x0 = e.cx + xLeft - pad;
    rb.x1 = e.cx + xRight + pad;
    rb.empty = false;
    return rb;
}

BinRect rectToBins(const TileRect& r, int S, int gridX, int gridY) {
    const int f = 16 / S;
    return {r.minX * f, r.minY * f, std::min(r.maxX * f, gridX), std::min(r.maxY * f, gridY)};
}

Coverage makeCoverage(uint32_t mode, const TileRect& rect16, int S, int gridX, int gridY, float cx, float cy,
                      const float conic[3], float opacity) {
    Coverage cv;
    cv.bins = mode == kCutoffOnly ? BinRect{0, 0, gridX, gridY} : rectToBins(rect16, S, gridX, gridY);
    if (mode != kRect) {
        cv.e = makeCutoffEllipse(cx, cy, conic, opacity);
        cv.useEllipse = cv.e.valid;
        if (!cv.e.valid && mode == kCutoffOnly) cv.bins = rectToBins(rect16, S, gridX, gridY);
    }
    return cv;
}

void coverageRows(const Coverage& cv, int S, int* r0, int* r1) {
    if (!cv.useEllipse) {
        *r0 = cv.bins.y0;
        *r1 = std::max(cv.bins.y0, cv.bins.y1);
        return;
    }
    ellipseRows(cv.e, S, cv.bins, r0, r1);
}

void coverageRowSpan(const Coverage& cv, int row, int S, int* x0, int* x1) {
    if (!cv.useEllipse) {
        *x0 = cv.bins.x0;
        *x1 = std::max(cv.bins.x0, cv.bins.x1);
        return;
    }
    const RowBounds rb = ellipseRowBounds(cv.e, row, S);
    if (rb.empty) {
        *x0 = cv.bins.x0;
        *x1 = cv.bins.x0;
        return;
    }
    const int a = std::max(cv.bins.x0, floorDivClamped(std::ceil(rb.x0), S));
    const int b = std::min(cv.bins.x1, floorDivClamped(std::floor(rb.x1), S) + 1);
    *x0 = a;
    *x1 = std::max(a, b);
}

uint32_t coverageCount(const Coverage& cv, int S) {
    int r0, r1;
    coverageRows(cv, S, &r0, &r1);
    if (!cv.useEllipse) return uint32_t(std::max(0, cv.bins.x1 - cv.bins.x0) * (r1 - r0));
    uint32_t n = 0;
    for (int r = r0; r < r1; ++r) {
        int x0, x1;
        coverageRowSpan(cv, r, S, &x0, &x1);
        n += uint32_t(x1 - x0);
    }
    return n;
}

Coverage splatCoverage(uint32_t mode, const Splat& s, uint32_t tilesX16, uint32_t tilesY16, int S, int gridX,
                       int gridY) {
    const TileRect rect = computeTileRect(s.center[0], s.center[1], s.radius, tilesX16, tilesY16);
    return makeCoverage(mode, rect, S, gridX, gridY, s.center[0], s.center[1], s.conicOpacity, s.conicOpacity[3]);
}

bool coverageIsBorderline(const Coverage& cv, int S, float eps) {
    if (!cv.useEllipse || cv.e.t < 0.0f) return false;
    const float rho = std::fabs(cv.e.B) / std::sqrt(cv.e.A * cv.e.C);
    const float K = 2.0f / std::max(1.0f - rho, 1e-30f);
    eps += (1e-7f * K + 1e-12f * K * K) * (cv.e.exPad + cv.e.eyPad);
    if (nearInteger(cv.e.cy - cv.e.eyPad, eps) || nearInteger(cv.e.cy + cv.e.eyPad, eps)) return true;
    int r0, r1;
    coverageRows(cv, S, &r0, &r1);
    for (int r = r0; r < r1; ++r) {
        const RowBounds rb = ellipseRowBounds(cv.e, r, S);
        if (rb.empty) continue;
        if (nearInteger(rb.x0, eps) || nearInteger(rb.x1, eps)) return true;
        const float ey0 = float(r * S) - cv.e.cy, ey1 = float(r * S + S - 1) - cv.e.cy;
        if (std::fabs(ey1 + cv.e.eyPad) < eps || std::fabs(ey0 - cv.e.eyPad) < eps) return true;
    }
    return false;
}

// --- Depth sort ------------------------------------------------------------------------------

uint32_t radixSortPairs(std::vector<uint32_t>& keys, std::vector<uint32_t>& values) {
    const uint32_t n = uint32_t(keys.size());
    std::vector<uint32_t> keysTmp(n), valuesTmp(n);
    uint32_t passes = 0;
    for (uint32_t shift = 0; shift < 32; shift += 8) {
        std::array<uint32_t, 256> histogram{};
        for (uint32_t i = 0; i < n; ++i) histogram[(keys[i] >> shift) & 0xFFu]++;
        if (n == 0 || histogram[(keys[0] >> shift) & 0xFFu] == n) continue;
        uint32_t sum = 0;
        for (uint32_t& h : histogram) {
            const uint32_t c = h;
            h = sum;
            sum += c;
        }
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t dst = histogram[(keys[i] >> shift) & 0xFFu]++;
            keysTmp[dst] = keys[i];
            valuesTmp[dst] = values[i];
        }
        keys.swap(keysTmp);
        values.swap(valuesTmp);
        ++passes;
    }
    return passes;
}

// --- half ------------------------------------------------------------------------------------

uint16_t floatToHalf(float value) {
    const uint32_t f = fbits(value);
    const uint32_t sign = (f >> 16) & 0x8000u;
    const uint32_t a = f & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return uint16_t(sign | 0x7C00u | (a > 0x7F800000u ? 0x200u : 0u));
    if (a >= 0x477FF000u) return uint16_t(sign | 0x7C00u);
    if (a < 0x38800000u) {
        if (a < 0x33000000u) return uint16_t(sign);
        const uint32_t exponent = a >> 23;
        const uint32_t mant = (a & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - exponent;
        uint32_t h = mant >> shift;
        const uint32_t rest = mant & ((1u << shift) - 1u);
        const uint32_t halfway = 1u << (shift - 1u);
        if (rest > halfway || (rest == halfway && (h & 1u))) ++h;
        return uint16_t(sign | h);
    }
    uint32_t h = (a - 0x38000000u) >> 13;
    const uint32_t rest = a & 0x1FFFu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) ++h;
    return uint16_t(sign | h);
}

float halfToFloat(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t exponent = (h >> 10) & 0x1Fu;
    const uint32_t mant = h & 0x3FFu;
    if (exponent == 0) {
        const float v = float(mant) * (1.0f / 16777216.0f);
        return sign ? -v : v;
    }
    if (exponent == 31) return bitsf(sign | 0x7F800000u | (mant << 13));
    return bitsf(sign | ((exponent + 112u) << 23) | (mant << 13));
}

// --- Packed render splats (render_splat.h:29-66) ---------------------------------------------

namespace {
uint32_t bitsOf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
uint32_t halfPair(float lo, float hi) { return uint32_t(floatToHalf(lo)) | (uint32_t(floatToHalf(hi)) << 16); }
uint32_t colorRadius(float b, float radius) {
    return uint32_t(floatToHalf(b)) | (uint32_t(radius < 65535.0f ? radius : 65535.0f) << 16);
}
float viaHalf(float v) { return halfToFloat(floatToHalf(v)); }
}  // namespace

void encodeRenderSplat(uint32_t mode, const Splat& s, uint32_t* w) {
    const float* co = s.conicOpacity;
    if (mode == kRenderSplatP32) {
        w[0] = bitsOf(s.center[0]);
        w[1] = bitsOf(s.center[1]);
        w[2] = bitsOf(co[0]);
        w[3] = bitsOf(co[1]);
        w[4] = bitsOf(co[2]);
        w[5] = bitsOf(co[3]);
        w[6] = halfPair(s.color[0], s.color[1]);
        w[7] = colorRadius(s.color[2], s.radius);
    } else {
        w[0] = bitsOf(s.center[0]);
        w[1] = bitsOf(s.center[1]);
        w[2] = halfPair(co[0] * kConicScaleP24, co[1] * kConicScaleP24);
        w[3] = halfPair(co[2] * kConicScaleP24, co[3]);
        w[4] = halfPair(s.color[0], s.color[1]);
        w[5] = colorRadius(s.color[2], s.radius);
    }
}

void applyRenderSplat(uint32_t mode, Splat& s, bool colour) {
    if (mode == kRenderSplatFull) return;
    if (mode == kRenderSplatP24) {
        const float inv = 1.0f / kConicScaleP24;
        for (int a = 0; a < 3; ++a) s.conicOpacity[a] = viaHalf(s.conicOpacity[a] * kConicScaleP24) * inv;
        s.conicOpacity[3] = viaHalf(s.conicOpacity[3]);
    }
    if (colour)
        for (int c = 0; c < 3; ++c) s.color[c] = viaHalf(s.color[c]);
}

// --- StopThePop ray depth (splat_math.cpp:195-205) ----------------------------------------------

RayDepth computeRayDepth(const float sigma[3], const float q[4], const float tCam[3], const float view[4][4]) {
    const M3<float> R = quaternionToMatrix<float>(q);
    float M[3][3];  // M[c][r]: column c = the gaussian's axis c in camera space
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            M[c][r] = view[0][r] * R.m[c][0] + view[1][r] * R.m[c][1] + view[2][r] * R.m[c][2];
    const float smax = std::max(sigma[0], std::max(sigma[1], sigma[2]));
    const float floorS = std::max(1e-4f * smax, 1e-12f);
    float w[3];
    for (int k = 0; k < 3; ++k) {
        const float sk = sigma[k] < floorS ? floorS : sigma[k];
        w[k] = 1.0f / (sk * sk);
    }
    float P[3][3] = {};
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) P[i][j] = P[i][j] + w[k] * (M[k][j] * M[k][i]);
    float b[3];
    for (int r = 0; r < 3; ++r) b[r] = P[0][r] * tCam[0] + P[1][r] * tCam[1] + P[2][r] * tCam[2];
    RayDepth d;
    d.p00 = P[0][0];
    d.p01 = P[0][1];
    d.p02 = P[0][2];
    d.p11 = P[1][1];
    d.p12 = P[1][2];
    d.p22 = P[2][2];
    d.b0 = b[0];
    d.b1 = b[1];
    d.b2 = b[2];
    return d;
}

}  // namespace oracle
