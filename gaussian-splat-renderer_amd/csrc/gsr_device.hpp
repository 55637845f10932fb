// Device-side per-gaussian math for gfx950: the INRIA preprocess and the tight tile
// coverage. Every expression follows the reference's written FP32 order
// (splat/splat_math.cpp:44-193, splat/tile_coverage.cpp:89-209) and is compiled with
// -ffp-contract=off and IEEE division/sqrt, so the count pass (preprocess) and the write
// pass (duplicate) compute the same bits, and both equal the CPU restatement.
#pragma once

#include "gsr_internal.hpp"

namespace gsr {

// fp16 bits -> float, exactly (integer decode, also for subnormals, Inf and NaN).
__device__ __forceinline__ float halfBitsToFloat(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t exponent = (h >> 10) & 0x1Fu;
    const uint32_t mant = h & 0x3FFu;
    if (exponent == 0) {
        const float v = float(mant) * (1.0f / 16777216.0f);
        return sign ? -v : v;
    }
    if (exponent == 31) return __uint_as_float(sign | 0x7F800000u | (mant << 13));
    return __uint_as_float(sign | ((exponent + 112u) << 23) | (mant << 13));
}

// float -> fp16 bits, IEEE round to nearest even (v_cvt_f16_f32).
__device__ __forceinline__ uint32_t halfBits(float v) {
    return uint32_t(__builtin_bit_cast(uint16_t, static_cast<_Float16>(v)));
}

// p24 render splats: the conic is stored in fp16 pre-multiplied by 2^12.
constexpr float kConicScaleP24 = 4096.0f;
constexpr float kInvConicScaleP24 = 1.0f / 4096.0f;

__device__ __forceinline__ float selMin(float a, float b) { return b < a ? b : a; }   // std::min(a, b)
__device__ __forceinline__ float selMax(float a, float b) { return a < b ? b : a; }   // std::max(a, b)
__device__ __forceinline__ float clampSel(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

// Natural log as the CPU oracle computes it: x = 2^e m, m in [sqrt(1/2), sqrt(2)),
// log = e ln2 + 2 atanh((m-1)/(m+1)), atanh by its series to s^21, in double with + - * /
// only (IEEE on both sides), rounded once to float.
__device__ __forceinline__ float detLogf(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (__builtin_isinf(x)) return x;
    uint32_t b = __float_as_uint(x);
    int e = 0;
    if ((b >> 23) == 0) {
        b = __float_as_uint(x * 16777216.0f);
        e = -24;
    }
    e += int(b >> 23) - 127;
    double m = double(__uint_as_float((b & 0x7FFFFFu) | 0x3F800000u));
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

// --- Tile rectangle (getRect, splat_math.cpp:116-127) -----------------------------------------
struct Rect {
    int x0, y0, x1, y1;
};
__device__ __forceinline__ int clampTile(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }
__device__ __forceinline__ Rect tileRect(float px, float py, float radius, uint32_t tilesX, uint32_t tilesY) {
    const int gx = int(tilesX), gy = int(tilesY);
    const int r = int(radius);
    const float S = float(kTileSize);
    Rect t;
    t.x0 = clampTile(int((px - float(r)) / S), gx);
    t.y0 = clampTile(int((py - float(r)) / S), gy);
    t.x1 = clampTile(int((px + float(r) + float(kTileSize - 1)) / S), gx);
    t.y1 = clampTile(int((py + float(r) + float(kTileSize - 1)) / S), gy);
    return t;
}

// --- Cutoff ellipse (makeCutoffEllipse, tile_coverage.cpp:89-115) ------------------------------
struct Ellipse {
    float cx, cy, A, B, C, D, t, exPad, eyPad;
    bool valid;
};

__device__ __forceinline__ Ellipse makeEllipse(float cx, float cy, float A, float B, float C, float opacity) {
    Ellipse e;
    e.cx = cx;
    e.cy = cy;
    e.A = A;
    e.B = B;
    e.C = C;
    e.D = A * C - B * B;
    e.t = -1.0f;
    e.exPad = 0.0f;
    e.eyPad = 0.0f;
    e.valid = A > 0.0f && C > 0.0f && e.D > 0.0f;
    if (!e.valid) return e;
    if (!(opacity >= kMinAlpha)) return e;
    const float qmax = 2.0f * detLogf(255.0f * opacity);
    const float rho = fabsf(B) / sqrtf(A * C);
    const float omr = 1.0f - rho;
    const float K = 2.0f / (omr < 1e-30f ? 1e-30f : omr);
    const float qm1 = qmax < 1.0f ? 1.0f : qmax;
    const float t = qmax + (1e-3f + 2e-5f * K) * qm1 + 2e-3f;
    if (!(t < 1e30f)) {
        e.valid = false;
        return e;
    }
    e.t = t;
    const float ex = sqrtf(t * C / e.D);
    const float ey = sqrtf(t * A / e.D);
    e.exPad = ex + (0.02f + 2e-3f * ex);
    e.eyPad = ey + (0.02f + 2e-3f * ey);
    return e;
}

// floor(v / S) for the bin sizes S = 8 and 16, v clamped: the division by a power of two is the
// exact scaling ldexp(v, -log2 S).
__device__ __forceinline__ int floorDivClamped(float v, int S) {
    return int(floorf(ldexpf(clampSel(v, -65536.0f, 1048576.0f), S == 8 ? -3 : -4)));
}

// Coverage of one splat in bins of S = 16 or 8 pixels (tile_coverage.cpp:24-164): either the
// getRect rectangle (on 16-px tiles) in bins (useEllipse = false) or that rectangle intersected
// with the cutoff ellipse, one contiguous run per bin row.
struct Coverage {
    Rect bins;
    Ellipse e;
    bool useEllipse;
};

// rectToBins: the 16-px rectangle in bins of S pixels, clipped to the bin grid.
__device__ __forceinline__ Rect rectToBins(const Rect& r, int S, int gridX, int gridY) {
    const int f = 16 / S;
    const int x1 = r.x1 * f, y1 = r.y1 * f;
    return Rect{r.x0 * f, r.y0 * f, x1 < gridX ? x1 : gridX, y1 < gridY ? y1 : gridY};
}

__device__ __forceinline__ Coverage makeCoverage(uint32_t mode, const Rect& rect, int S, int gridX, int gridY,
                                                 float cx, float cy, float A, float B, float C, float opacity) {
    Coverage cv;
    if (mode == GSR_BOUNDS_CUTOFF_ONLY) {
        cv.bins = Rect{0, 0, gridX, gridY};
    } else {
        cv.bins = rectToBins(rect, S, gridX, gridY);
    }
    cv.useEllipse = false;
    cv.e.valid = false;
    cv.e.t = -1.0f;
    if (mode != GSR_BOUNDS_RECT) {
        cv.e = makeEllipse(cx, cy, A, B, C, opacity);
        cv.useEllipse = cv.e.valid;
        if (!cv.e.valid && mode == GSR_BOUNDS_CUTOFF_ONLY) cv.bins = rectToBins(rect, S, gridX, gridY);
    }
    return cv;
}

__device__ __forceinline__ void coverageRows(const Coverage& cv, int S, int* r0, int* r1) {
    if (!cv.useEllipse) {
        *r0 = cv.bins.y0;
        *r1 = cv.bins.y1 > cv.bins.y0 ? cv.bins.y1 : cv.bins.y0;
        return;
    }
    if (cv.e.t < 0.0f) {
        *r0 = 0;
        *r1 = 0;
        return;
    }
    const float y0 = ceilf(cv.e.cy - cv.e.eyPad);
    const float y1 = floorf(cv.e.cy + cv.e.eyPad);
    const int a = floorDivClamped(y0, S);
    const int b = floorDivClamped(y1, S) + 1;
    const int lo = cv.bins.y0 > a ? cv.bins.y0 : a;
    const int hi = cv.bins.y1 < b ? cv.bins.y1 : b;
    *r0 = lo;
    *r1 = hi > lo ? hi : lo;
}

// ellipseRowSpan (tile_coverage.cpp:135-161): the run [x0, x1) of tile row `row`.
__device__ __forceinline__ void coverageRowSpan(const Coverage& cv, int row, int S, int* x0, int* x1) {
    if (!cv.useEllipse) {
        *x0 = cv.bins.x0;
        *x1 = cv.bins.x1 > cv.bins.x0 ? cv.bins.x1 : cv.bins.x0;
        return;
    }
    const Ellipse& e = cv.e;
    const float ey0 = float(row * S) - e.cy;
    const float ey1 = float(row * S + S - 1) - e.cy;
    const float nep = -e.eyPad;
    const float lo = ey0 < nep ? nep : ey0;
    const float hi = e.eyPad < ey1 ? e.eyPad : ey1;
    if (!(lo <= hi)) {
        *x0 = cv.bins.x0;
        *x1 = cv.bins.x0;
        return;
    }
    const float exTrue = sqrtf(e.t * e.C / e.D);
    const float eyRight = clampSel(-e.B * exTrue / e.C, lo, hi);
    const float eyLeft = clampSel(e.B * exTrue / e.C, lo, hi);
    const float At = e.A * e.t;
    const float dr = At - e.D * eyRight * eyRight;
    const float dl = At - e.D * eyLeft * eyLeft;
    const float xRight = (-e.B * eyRight + sqrtf(0.0f < dr ? dr : 0.0f)) / e.A;
    const float xLeft = (-e.B * eyLeft - sqrtf(0.0f < dl ? dl : 0.0f)) / e.A;
    const float pad = 0.02f + 2e-3f * exTrue;
    const float bx0 = e.cx + xLeft - pad;
    const float bx1 = e.cx + xRight + pad;
    const int a = floorDivClamped(ceilf(bx0), S);
    const int b = floorDivClamped(floorf(bx1), S) + 1;
    const int lo2 = cv.bins.x0 > a ? cv.bins.x0 : a;
    const int hi2 = cv.bins.x1 < b ? cv.bins.x1 : b;
    *x0 = lo2;
    *x1 = hi2 > lo2 ? hi2 : lo2;
}

// The splat's coverage as the duplicate pass walks it: the first bin row, the row count (negative,
// -(rows) - 1, for a rectangle: one span for every row) and the first kSpanRows row spans
// (x0 | width << 16; a rectangle: span 0). Written by the preprocess with the tile count, so the
// duplicate pass need not set the coverage up again.
constexpr int kSpanRows = 8;
struct SpanRecord {
    uint32_t info;  // first row (low 16 bits) | int16 row code (high 16 bits)
    uint32_t span[kSpanRows];
};
__device__ __forceinline__ uint32_t coverageCountSpans(const Coverage& cv, int S, SpanRecord& rec) {
    int r0, r1;
    coverageRows(cv, S, &r0, &r1);
#pragma unroll
    for (int k = 0; k < kSpanRows; ++k) rec.span[k] = 0u;
    if (!cv.useEllipse) {
        const int w = cv.bins.x1 - cv.bins.x0;
        rec.info = uint32_t(r0 & 0xFFFF) | (uint32_t(uint16_t(int16_t(-(r1 - r0) - 1))) << 16);
        rec.span[0] = uint32_t(cv.bins.x0) | (uint32_t(w > 0 ? w : 0) << 16);
        return uint32_t((w > 0 ? w : 0) * (r1 - r0));
    }
    rec.info = uint32_t(r0 & 0xFFFF) | (uint32_t(uint16_t(int16_t(r1 - r0))) << 16);
    uint32_t n = 0;
#pragma unroll
    for (int k = 0; k < kSpanRows / 2; ++k) {  // the first rows straight into their slots
        if (r0 + k < r1) {
            int x0, x1;
            coverageRowSpan(cv, r0 + k, S, &x0, &x1);
            rec.span[k] = uint32_t(x0) | (uint32_t(x1 - x0) << 16);
            n += uint32_t(x1 - x0);
        }
    }
    for (int r = r0 + kSpanRows / 2; r < r1; ++r) {  // taller splats
        int x0, x1;
        coverageRowSpan(cv, r, S, &x0, &x1);
        const uint32_t sp = uint32_t(x0) | (uint32_t(x1 - x0) << 16);
#pragma unroll
        for (int k = kSpanRows / 2; k < kSpanRows; ++k)
            if (r - r0 == k) rec.span[k] = sp;
        n += uint32_t(x1 - x0);
    }
    return n;
}

__device__ __forceinline__ uint32_t coverageCount(const Coverage& cv, int S) {
    int r0, r1;
    coverageRows(cv, S, &r0, &r1);
    if (!cv.useEllipse) {
        const int w = cv.bins.x1 - cv.bins.x0;
        return uint32_t((w > 0 ? w : 0) * (r1 - r0));
    }
    uint32_t n = 0;
    for (int r = r0; r < r1; ++r) {
        int x0, x1;
        coverageRowSpan(cv, r, S, &x0, &x1);
        n += uint32_t(x1 - x0);
    }
    return n;
}

// --- SH colour (shared by the preprocess and the learned-colour training targets) ------------

constexpr float kC0 = 0.28209479177387814f;
constexpr float kC1 = 0.4886025119029199f;
constexpr float kC2_0 = 1.0925484305920792f, kC2_1 = -1.0925484305920792f, kC2_2 = 0.31539156525252005f,
                kC2_3 = -1.0925484305920792f, kC2_4 = 0.5462742152960396f;
constexpr float kC3_0 = -0.5900435899266435f, kC3_1 = 2.890611442640554f, kC3_2 = -0.4570457994644658f,
                kC3_3 = 0.3731763325901154f, kC3_4 = -0.4570457994644658f, kC3_5 = 1.445305721320277f,
                kC3_6 = -0.5900435899266435f;

// evalSH (splat_math.cpp:129-151), one channel; c = the 16 coefficients of that channel.
// evalSH before its final clamp: the sum + 0.5.
__device__ __forceinline__ float evalShRaw(int degree, const float* c, float x, float y, float z) {
    float result = kC0 * c[0];
    if (degree > 0) {
        result = result - kC1 * y * c[1] + kC1 * z * c[2] - kC1 * x * c[3];
        if (degree > 1) {
            const float xx = x * x, yy = y * y, zz = z * z;
            const float xy = x * y, yz = y * z, xz = x * z;
            result = result + kC2_0 * xy * c[4] + kC2_1 * yz * c[5] + kC2_2 * (2.0f * zz - xx - yy) * c[6] +
                     kC2_3 * xz * c[7] + kC2_4 * (xx - yy) * c[8];
            if (degree > 2) {
                result = result + kC3_0 * y * (3.0f * xx - yy) * c[9] + kC3_1 * xy * z * c[10] +
                         kC3_2 * y * (4.0f * zz - xx - yy) * c[11] +
                         kC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy) * c[12] +
                         kC3_4 * x * (4.0f * zz - xx - yy) * c[13] + kC3_5 * z * (xx - yy) * c[14] +
                         kC3_6 * x * (xx - 3.0f * yy) * c[15];
            }
        }
    }
    return result + 0.5f;
}

__device__ __forceinline__ float evalShChannel(int degree, const float* c, float x, float y, float z) {
    const float result = evalShRaw(degree, c, x, y, z);
    return result < 0.0f ? 0.0f : result;
}

// evalSH's basis values for the evaluated bands, each formed as evalShRaw forms it (term k of a
// channel is b[k] * c[k]; terms 1 and 3 are subtracted, the rest added, in k order).
__device__ __forceinline__ void shBasis(int degree, float x, float y, float z, float b[16]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) b[k] = 0.0f;
    b[0] = kC0;
    if (degree > 0) {
        b[1] = kC1 * y;
        b[2] = kC1 * z;
        b[3] = kC1 * x;
        if (degree > 1) {
            const float xx = x * x, yy = y * y, zz = z * z;
            const float xy = x * y, yz = y * z, xz = x * z;
            b[4] = kC2_0 * xy;
            b[5] = kC2_1 * yz;
            b[6] = kC2_2 * (2.0f * zz - xx - yy);
            b[7] = kC2_3 * xz;
            b[8] = kC2_4 * (xx - yy);
            if (degree > 2) {
                b[9] = kC3_0 * y * (3.0f * xx - yy);
                b[10] = kC3_1 * xy * z;
                b[11] = kC3_2 * y * (4.0f * zz - xx - yy);
                b[12] = kC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
                b[13] = kC3_4 * x * (4.0f * zz - xx - yy);
                b[14] = kC3_5 * z * (xx - yy);
                b[15] = kC3_6 * x * (xx - 3.0f * yy);
            }
        }
    }
}

__device__ __forceinline__ float shTerm(int k, float acc, float b, float c) {
    return k == 0 ? b * c : (k == 1 || k == 3) ? acc - b * c : acc + b * c;
}

// evalShChannel for the three channels, reading the coefficient quads in record order (word j =
// channel j % 3 of coefficient j / 3) and adding each word to its channel's sum as it arrives:
// every channel gets its terms in k order, so the sums are evalShRaw's, bit for bit, and only one
// quad of coefficients is live at a time.
// The f32 scene's coefficient quads of gaussian i that the evaluated bands use (q[k] = plane k;
// the unused ones zero).
__device__ __forceinline__ void loadShQuads(const SceneRef& s, uint32_t i, int degree, float4 (&q)[12]) {
    const int used = (degree + 1) * (degree + 1);
    const int quads = (used * 3 + 3) / 4;
#pragma unroll
    for (int k = 0; k < 12; ++k) q[k] = k < quads ? s.shPlanes[k * size_t(s.count) + i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// evalShStreamed of an f32 scene over quads loaded by loadShQuads.
__device__ __forceinline__ void evalShQuads(int degree, const float4 (&q)[12], float x, float y, float z,
                                            float out[3]) {
    const int used = (degree + 1) * (degree + 1);
    float b[16];
    shBasis(degree, x, y, z, b);
    float acc[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const float vv[4] = {q[k].x, q[k].y, q[k].z, q[k].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int j = k * 4 + e;
            if (j < used * 3) acc[j % 3] = shTerm(j / 3, acc[j % 3], b[j / 3], vv[e]);
        }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float r = acc[ch] + 0.5f;
        out[ch] = r < 0.0f ? 0.0f : r;
    }
}

template <int FMT>
__device__ __forceinline__ void evalShStreamed(const SceneRef& s, uint32_t i, int degree, float x, float y, float z,
                                               float out[3]) {
    if (FMT == GSR_SCENE_F32_SOA) {
        float4 q[12];
        loadShQuads(s, i, degree, q);
        evalShQuads(degree, q, x, y, z, out);
        return;
    }
    const int used = (degree + 1) * (degree + 1);
    float b[16];
    shBasis(degree, x, y, z, b);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    const size_t plane = s.count;
    {
        const int quads = (used * 3 + 7) / 8;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (q < quads) {
                const uint4 v = s.color[q * plane + i];
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int j = q * 8 + e;
                    const uint32_t h = (e & 1) ? (w[e >> 1] >> 16) : (w[e >> 1] & 0xFFFFu);
                    if (j < used * 3) acc[j % 3] = shTerm(j / 3, acc[j % 3], b[j / 3], halfBitsToFloat(h));
                }
            }
        }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float r = acc[ch] + 0.5f;
        out[ch] = r < 0.0f ? 0.0f : r;
    }
}

// The coefficients of the evaluated bands, channel-major: co[ch][k].
template <int FMT>
__device__ __forceinline__ void loadSh(const SceneRef& s, uint32_t i, int used, float co[3][16]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) co[0][k] = co[1][k] = co[2][k] = 0.0f;
    const size_t plane = s.count;
    if (FMT == GSR_SCENE_F32_SOA) {
        const int quads = (used * 3 + 3) / 4;
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            if (q < quads) {
                const float4 v = s.shPlanes[q * plane + i];
                const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int j = q * 4 + e;
                    if (j < used * 3) co[j % 3][j / 3] = vv[e];
                }
            }
        }
    } else {
        const int quads = (used * 3 + 7) / 8;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (q < quads) {
                const uint4 v = s.color[q * plane + i];
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int j = q * 8 + e;
                    const uint32_t h = (e & 1) ? (w[e >> 1] >> 16) : (w[e >> 1] & 0xFFFFu);
                    if (j < used * 3) co[j % 3][j / 3] = halfBitsToFloat(h);
                }
            }
        }
    }
}

// Whether the conic (a, b, c) is positive definite (det = a c - b^2 from an exact product, so
// that it does not cancel). The preprocess's conics are (the inverse of a positive definite
// covariance); p24 render records round a, b, c to fp16 separately, which can make the conic of a
// thin splat indefinite, and then the reference's `power > 0` test (tile_render.comp:340) drops
// the pixels along its open directions.
__device__ __forceinline__ bool conicPositiveDefinite(float a, float b, float c) {
    const float bb = b * b;
    const float det = fmaf(a, c, -bb) - fmaf(b, b, -bb);
    return det > 0.0f && a > 0.0f && c > 0.0f;
}

// The quadrants (8x8 pixel blocks of the 16x16 tile at (tx0, ty0); bit w: quadrant
// (w & 1, w >> 1)) in which a splat can reach a pixel. With the conic (a, b, c) the render takes a
// pixel iff its q >= log2(1/255), i.e. iff Q = a dx^2 + 2 b dx dy + c dy^2 <= r2,
// r2 = 2 (log2 o - log2(1/255)) / log2 e, up to the FP32 rounding of q. That rounding is relative
// to the size of the terms of the tile-local quadratic the render evaluates, M (centres far from
// the tile and nearly singular conics make those terms large and cancelling), so every test here
// is against lim = r2 + 2e-3 |r2| + 1e-4 + 4e-6 M, several times that rounding: a quadrant is
// dropped only when no pixel centre in it has Q <= lim. The tests:
//   - lim < 0: no pixel anywhere;
//   - the ellipse's bounding box (half-extents sqrt(lim c / det), sqrt(lim a / det), widened by
//     0.1 % and one pixel; det = a c - b^2 from an exact product so that it does not cancel);
//   - the minimum of Q over the quadrant's pixel-centre rectangle, which is on the edge facing
//     the centre (or 0 inside): at x = clamp(cx) with the y minimising Q along that line
//     (dy = -b dx / c), or at y = clamp(cy) likewise.
// A conic that is not positive definite, and NaN anywhere, keep every quadrant.
__device__ __forceinline__ uint32_t quadrantMask(float cx, float cy, float a, float b, float c, float r2, float tx0,
                                                 float ty0) {
    const float bb = b * b;
    const float det = fmaf(a, c, -bb) - fmaf(b, b, -bb);  // a c - b^2: b^2 = bb + the exact rounding error
    if (!(det > 0.0f && a > 0.0f && c > 0.0f)) return 0xFu;
    const float Dx = fabsf(cx - tx0) + 16.0f, Dy = fabsf(cy - ty0) + 16.0f;
    const float M = a * Dx * Dx + 2.0f * fabsf(b) * Dx * Dy + c * Dy * Dy;
    const float lim = r2 + 2e-3f * fabsf(r2) + 1e-4f + 4e-6f * M;
    if (lim < 0.0f) return 0u;
    // The hardware reciprocal and square root (errors near 1e-7, relative) are well inside the
    // 0.1 % widening; 1 / det overflowing gives +inf, caught below.
    const float rdet = __builtin_amdgcn_rcpf(det);
    const float ex = __builtin_amdgcn_sqrtf(lim * c * rdet) * 1.001f + 1.0f;
    const float ey = __builtin_amdgcn_sqrtf(lim * a * rdet) * 1.001f + 1.0f;
    if (!(ex < 1e30f && ey < 1e30f && lim < 1e30f)) return 0xFu;  // also false for NaN
    const float kb = b * __builtin_amdgcn_rcpf(c);
    const float kc = b * __builtin_amdgcn_rcpf(a);
    auto form = [&](float dx, float dy) { return fmaf(dx, fmaf(a, dx, 2.0f * b * dy), c * dy * dy); };
    uint32_t m = 0u;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
        const float x0 = tx0 + 8.0f * float(w & 1u), y0 = ty0 + 8.0f * float(w >> 1);
        const float x1 = x0 + 7.0f, y1 = y0 + 7.0f;
        const bool box = cx + ex >= x0 && cx - ex <= x1 && cy + ey >= y0 && cy - ey <= y1;
        const float dx1 = fminf(fmaxf(cx, x0), x1) - cx;
        const float dy1 = fminf(fmaxf(fmaf(-kb, dx1, cy), y0), y1) - cy;
        const float dy2 = fminf(fmaxf(cy, y0), y1) - cy;
        const float dx2 = fminf(fmaxf(fmaf(-kc, dy2, cx), x0), x1) - cx;
        const bool far = fminf(form(dx1, dy1), form(dx2, dy2)) > lim;  // NaN: kept
        m |= box && !far ? (1u << w) : 0u;
    }
    return m;
}

}  // namespace gsr
