// Differentiable rendering (no counterpart in the reference, which renders forward only): a
// training forward that keeps each pixel's final transmittance and last contributor, the render
// backward (per tile, the list walked back to front, transmittance recovered by division, the
// per-gaussian gradients reduced per wave and in LDS before one global atomic per list entry and
// component), and the preprocess backward (per gaussian, the chain rule through the conic, the
// EWA projection, the 3D covariance, the quaternion, the projection to pixels, the view direction
// and the SH evaluation of preprocess.hip).
//
// The image is out = sum_i c_i a_i T_i + T_N bg (bg = 0), a_i = min(0.99, o_i exp(p_i)),
// p_i = -0.5 (A dx^2 + C dy^2) - B dx dy with (dx, dy) = centre - pixel; a splat is skipped where
// p > 0 or a < 1/255, and the walk ends before the splat whose T (1 - a) falls below 1e-4. The
// gradient is that of this function where it is differentiable (a clamped alpha passes none).
#include "gsr_device.hpp"

namespace gsr {

namespace {

constexpr int kBatch = 256;  // list entries staged in LDS per round (one per thread)
// Per-gaussian screen-space gradient planes (grad2D + v * N): centre x, y; conic A, B, C;
// opacity; colour r, g, b.
constexpr int kGrad2D = 9;

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLog2MinAlpha = -7.994353436858858f;  // log2(1/255)

// One list entry as staged in LDS: the conic scaled into the log2 domain and log2(opacity),
// s0 = (cx, cy, -log2(e) A / 2, -log2(e) B), s1 = (-log2(e) C / 2, log2 o), its colour, and the
// mask of the tile's 8x8 quadrants (bit w: quadrant (w % 2, w / 2)) in which its alpha can reach
// 1/255 (quadrantMask: the ellipse's box, then the exact ellipse-vs-quadrant test, both with a
// margin sized to the rounding of q).
__device__ __forceinline__ uint32_t stageEntry(const float4* splats, uint32_t g, float tx0, float ty0, float4& s0,
                                               float2& s1, float4& col) {
    const float4 a = splats[3 * size_t(g)], b = splats[3 * size_t(g) + 1], c = splats[3 * size_t(g) + 2];
    s0 = make_float4(a.x, a.y, -0.5f * kLog2e * b.x, -kLog2e * b.y);
    s1 = make_float2(-0.5f * kLog2e * b.z, __log2f(b.w));
    col = c;
    const float r2 = 2.0f * (s1.y - kLog2MinAlpha) * (1.0f / kLog2e);
    return quadrantMask(a.x, a.y, b.x, b.y, b.z, r2, tx0, ty0);
}

// The wave's list: the batch slots whose mask has the wave's bit, in slot order; returns the count.
__device__ __forceinline__ uint32_t waveList(const uint32_t* mask, uint32_t* list, uint32_t wave, uint32_t lane) {
    const uint64_t lt = (uint64_t(1) << lane) - 1u;
    uint32_t count = 0;
#pragma unroll
    for (uint32_t j = 0; j < kBatch / 64; ++j) {
        const uint32_t slot = j * 64 + lane;
        const bool hit = (mask[slot] >> wave) & 1u;
        const uint64_t b = __builtin_amdgcn_ballot_w64(hit);
        if (hit) list[count + uint32_t(__popcll(b & lt))] = slot;
        count += uint32_t(__popcll(b));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return count;
}

// q = log2(o exp(power)) at a pixel (dx, dy = centre - pixel), -inf where power > 0 (only for
// a conic that is not positive definite).
__device__ __forceinline__ float qOf(float4 s0, float2 s1, float dx, float dy) {
    const float p = fmaf(fmaf(s0.z, dx, s0.w * dy), dx, s1.x * dy * dy);
    return p <= 0.0f ? p + s1.y : -__builtin_inff();
}

// A splat takes part at a pixel when q >= log2(1/255) (alpha >= 1/255 up to the rounding of q);
// then alpha = min(0.99, 2^q). The training forward and the backward take every decision from
// these two expressions, so the backward retraces the forward's walk exactly. *ex = 2^q (before
// the 0.99 clamp); 0 where the splat does not take part.
__device__ __forceinline__ float alphaOf(float4 s0, float2 s1, float dx, float dy, float* ex) {
    const float q = qOf(s0, s1, dx, dy);
    *ex = 0.0f;
    if (!(q >= kLog2MinAlpha)) return 0.0f;
    const float e = __builtin_amdgcn_exp2f(q);
    *ex = e;
    return kMaxAlpha < e ? kMaxAlpha : e;
}

// Tile order: workgroup b runs on XCD b % 8, and XCD x takes tile rows x, x + 8, x + 16, ... from
// the middle row outwards (the heavy central tiles first, the light ones fill the tail); the grid
// is ceil(tilesY / 8) x tilesX x 8 workgroups, those past the last row return at once.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t centreOut(uint32_t s, uint32_t n) {
    const uint32_t m = (n - 1u) / 2u;
    return (s & 1u) ? m + (s + 1u) / 2u : m - s / 2u;
}
__device__ __forceinline__ uint32_t scheduledTile(uint32_t tilesX, uint32_t tileCount) {
    const uint32_t xcd = blockIdx.x % kXcds, idx = blockIdx.x / kXcds;
    const uint32_t row = xcd + kXcds * centreOut(idx / tilesX, gridDim.x / kXcds / tilesX);
    const uint32_t tile = row * tilesX + idx % tilesX;
    return tile < tileCount ? tile : 0xFFFFFFFFu;
}
uint32_t scheduledGrid(const DevCamera& cam) { return (cam.tilesY + kXcds - 1) / kXcds * cam.tilesX * kXcds; }

// Wave w of a tile's workgroup takes the 8x8 quadrant (w % 2, w / 2): lane l at (l % 8, l / 8).
__device__ __forceinline__ void quadrantPixel(uint32_t tile, uint32_t tilesX, uint32_t* px, uint32_t* py, float* qx,
                                              float* qy) {
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    const uint32_t x0 = (tile % tilesX) * kTileSize + 8 * (w & 1u), y0 = (tile / tilesX) * kTileSize + 8 * (w >> 1);
    *px = x0 + (l & 7u);
    *py = y0 + (l >> 3);
    *qx = float(x0);
    *qy = float(y0);
}

// Training forward: one workgroup per 16x16 tile, one pixel per thread, wave w on the 8x8
// quadrant (w % 2, w / 2), walking the batch entries whose box meets its quadrant four at a time
// (the four alphas first, their LDS reads in flight together, then the blends in list order).
// out = (C, 1 - T); state = (T bits, entries of the tile's range up to and including the last
// contributor).
__global__ __launch_bounds__(256) void kRenderTrain(const float4* __restrict__ splats, const uint32_t* __restrict__ values,
                                                    const uint2* __restrict__ ranges, DevCamera cam,
                                                    float4* __restrict__ out, uint2* __restrict__ state) {
    __shared__ float4 sS0[kBatch];
    __shared__ float2 sS1[kBatch];
    __shared__ float4 sCol[kBatch];
    __shared__ uint32_t sMask[kBatch];
    __shared__ uint32_t sList[4][kBatch];
    const uint32_t tile = scheduledTile(cam.tilesX, cam.tilesX * cam.tilesY);
    if (tile == 0xFFFFFFFFu) return;  // whole workgroup
    uint32_t px, py;
    float qx, qy;
    quadrantPixel(tile, cam.tilesX, &px, &py, &qx, &qy);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const float tx0 = float((tile % cam.tilesX) * kTileSize), ty0 = float((tile / cam.tilesX) * kTileSize);
    const bool inside = px < cam.width && py < cam.height;
    const float pxf = float(px), pyf = float(py);
    const uint2 range = ranges[tile];
    // The lane's state is its transmittance, colour, last contributor and cutoff thr
    // (log2(1/255) while the pixel is open, +inf once done or outside the image): every blend is
    // branch-free, and a splat that no lane of the wave takes is skipped.
    float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
    uint32_t last = 0;
    float thr = inside ? kLog2MinAlpha : __builtin_inff();
    for (uint32_t base = range.x; base < range.y; base += kBatch) {
        if (__syncthreads_count(thr < __builtin_inff()) == 0) break;
        const uint32_t j = base + threadIdx.x;
        sMask[threadIdx.x] = j < range.y ? stageEntry(splats, values[j], tx0, ty0, sS0[threadIdx.x], sS1[threadIdx.x],
                                                      sCol[threadIdx.x])
                                         : 0u;
        __syncthreads();
        const uint32_t count = waveList(sMask, sList[wave], wave, lane);
        for (uint32_t i = 0; i < count; i += 4) {
            if (__builtin_amdgcn_ballot_w64(thr < __builtin_inff()) == 0) break;
            uint32_t slot[4];
            float q[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {  // the four q first, their LDS reads in flight together
                slot[u] = sList[wave][i + u < count ? i + u : i];
                const float4 s0 = sS0[slot[u]];
                q[u] = i + u < count ? qOf(s0, sS1[slot[u]], s0.x - pxf, s0.y - pyf) : -__builtin_inff();
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                if (__builtin_amdgcn_ballot_w64(q[u] >= thr) == 0) continue;
                const bool take = q[u] >= thr;
                const float a = fminf(kMaxAlpha, __builtin_amdgcn_exp2f(q[u]));
                const float testT = T * (1.0f - (take ? a : 0.0f));
                const bool term = take && testT < kMinTransmittance;  // not blended; the pixel ends
                const bool blend = take && !term;
                const float4 col = sCol[slot[u]];
                const float w = blend ? a * T : 0.0f;
                C0 = fmaf(col.x, w, C0);
                C1 = fmaf(col.y, w, C1);
                C2 = fmaf(col.z, w, C2);
                T = blend ? testT : T;
                last = blend ? base + slot[u] - range.x + 1 : last;
                thr = term ? __builtin_inff() : thr;
            }
        }
    }
    if (inside) {
        const size_t p = size_t(py) * cam.width + px;
        out[p] = make_float4(C0, C1, C2, 1.0f - T);
        state[p] = make_uint2(__float_as_uint(T), last);
    }
}

// Lane exchanges for the wave sums, all VALU: v_permlane32_swap / v_permlane16_swap exchange
// half-waves / neighbouring rows of two registers at once, DPP moves permute within a row.
__device__ __forceinline__ float swapSum32(float a, float b) {  // a of lanes 0-31 paired with b of 32-63
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swapSum16(float a, float b) {  // rows 0/1 and 2/3 likewise
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
constexpr int kDppRor8 = 0x128;        // row_ror:8 (lane ^ 8 within a row of 16)
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror (lane -> 7 - lane within 8)
constexpr int kDppXor1 = 0xB1;         // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;         // quad_perm [2, 3, 0, 1]
__device__ __forceinline__ float sum8Lanes(float v) {  // over each group of 8 lanes (lane bits 0-2)
    v += dpp<kDppHalfMirror>(v);
    v += dpp<kDppXor1>(v);
    return v + dpp<kDppXor2>(v);
}

// Render backward: one workgroup per tile (waves on 8x8 quadrants). Each pixel walks its entries from its last contributor
// back to the first, recovering T_i = T_{i+1} / (1 - a_i) and the colour behind,
// B_{i-1} = c_i a_i + (1 - a_i) B_i (B = bg = 0 behind the last); dout/da_i = T_i (c_i - B_i),
// dout/dc_i = a_i T_i. The 9 gradient values of an entry are summed over the wave's 64 pixels
// (shuffles, four entries together), the waves' sums over the tile in LDS, and each entry adds
// its tile's sums to the gaussian's planes with one atomic per component.
__global__ __launch_bounds__(256) void kRenderBackward(const float4* __restrict__ splats,
                                                       const uint32_t* __restrict__ values,
                                                       const uint2* __restrict__ ranges, DevCamera cam,
                                                       const float4* __restrict__ dOut, const uint2* __restrict__ state,
                                                       float* __restrict__ grad2D, uint32_t gaussianCount) {
    __shared__ float4 sS0[kBatch];
    __shared__ float2 sS1[kBatch];
    __shared__ float4 sCol[kBatch];
    __shared__ uint32_t sMask[kBatch];
    __shared__ uint32_t sList[4][kBatch];
    __shared__ float sG[kGrad2D][kBatch];
    __shared__ uint32_t sMax[4];
    const uint32_t tile = scheduledTile(cam.tilesX, cam.tilesX * cam.tilesY);
    if (tile == 0xFFFFFFFFu) return;  // whole workgroup
    uint32_t px, py;
    float qx, qy;
    quadrantPixel(tile, cam.tilesX, &px, &py, &qx, &qy);
    const float tx0 = float((tile % cam.tilesX) * kTileSize), ty0 = float((tile / cam.tilesX) * kTileSize);
    const bool inside = px < cam.width && py < cam.height;
    const float pxf = float(px), pyf = float(py);
    const uint2 range = ranges[tile];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    float T = 1.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    uint32_t last = 0;
    if (inside) {
        const size_t p = size_t(py) * cam.width + px;
        const uint2 st = state[p];
        T = __uint_as_float(st.x);
        last = st.y;
        const float4 d = dOut[p];
        g0 = d.x;
        g1 = d.y;
        g2 = d.z;
    }
    float B0 = 0.0f, B1 = 0.0f, B2 = 0.0f;
    // The tile's largest last-contributor count bounds the walk.
    uint32_t m = last;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, uint32_t(__shfl_xor(int(m), off, 64)));
    if (lane == 0) sMax[wave] = m;
    __syncthreads();
    const uint32_t maxLast = max(max(sMax[0], sMax[1]), max(sMax[2], sMax[3]));
    for (uint32_t end = maxLast; end > 0;) {
        const uint32_t begin = end > uint32_t(kBatch) ? end - uint32_t(kBatch) : 0u;
        const uint32_t n = end - begin;
        __syncthreads();  // the previous round's LDS reads are done
        sMask[threadIdx.x] = threadIdx.x < n ? stageEntry(splats, values[range.x + begin + threadIdx.x], tx0, ty0,
                                                          sS0[threadIdx.x], sS1[threadIdx.x], sCol[threadIdx.x])
                                             : 0u;
#pragma unroll
        for (int v = 0; v < kGrad2D; ++v) sG[v][threadIdx.x] = 0.0f;
        __syncthreads();
        const uint32_t listCount = waveList(sMask, sList[wave], wave, lane);
        // Four entries at a time (list positions i, i - 1, i - 2, i - 3): the pixel's state is
        // updated entry by entry, the 4 x 9 gradient terms are summed over the wave together
        // (32 terms in 32 shuffles, leaving each lane pair one sum; the 4 colour-b terms in 7).
        for (int i = int(listCount) - 1; i >= 0; i -= 4) {
            float v[36];  // v[8 u + q] for the terms q < 8 of entry u, v[32 + u] for term 8
#pragma unroll
            for (int q = 0; q < 36; ++q) v[q] = 0.0f;
            bool act4 = false;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (i - u < 0) break;  // uniform
                const uint32_t slot = sList[wave][i - u];
                const uint32_t j = begin + slot;  // the entry's position in the tile's range
                if (j >= last) continue;
                const float4 s0 = sS0[slot];
                const float2 s1 = sS1[slot];
                const float dx = s0.x - pxf, dy = s0.y - pyf;
                float ex;
                const float alpha = alphaOf(s0, s1, dx, dy, &ex);
                if (alpha == 0.0f) continue;
                act4 = true;
                const float4 col = sCol[slot];
                T = T * __builtin_amdgcn_rcpf(1.0f - alpha);  // T before this splat (1 - alpha >= 0.01)
                const float dAlpha = T * ((col.x - B0) * g0 + (col.y - B1) * g1 + (col.z - B2) * g2);
                v[8 * u + 6] = alpha * T * g0;
                v[8 * u + 7] = alpha * T * g1;
                v[32 + u] = alpha * T * g2;
                B0 = col.x * alpha + (1.0f - alpha) * B0;
                B1 = col.y * alpha + (1.0f - alpha) * B1;
                B2 = col.z * alpha + (1.0f - alpha) * B2;
                if (ex < kMaxAlpha) {  // alpha = o exp(power): d/dpower = alpha, d/do = alpha / o
                    // The pixel's moments of g = dL/dpower; the per-gaussian factors (conic,
                    // 1 / o) are applied once per entry when the tile's sums are flushed.
                    const float gp = dAlpha * alpha;
                    const float gx = gp * dx, gy = gp * dy;
                    v[8 * u + 0] = gx;
                    v[8 * u + 1] = gy;
                    v[8 * u + 2] = gx * dx;
                    v[8 * u + 3] = gx * dy;
                    v[8 * u + 4] = gy * dy;
                    v[8 * u + 5] = gp;
                }
            }
            if (__any(act4)) {  // wave-uniform
                // The 32 terms v[8 u + q]: swap-sums at lane distances 32 (16 pairs) and 16 (8 pairs)
                // leave v[j] = term j + 8 b4 + 16 b5 (b = lane bits); a halving at distance 8
                // leaves v[j] = term j + 4 b3 + 8 b4 + 16 b5 (j < 4), summed over the lane groups
                // of 8: every lane then holds its group's four totals.
#pragma unroll
                for (int j = 0; j < 16; ++j) v[j] = swapSum32(v[j], v[j + 16]);
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = swapSum16(v[j], v[j + 8]);
                const bool h3 = (lane & 8u) != 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float keep = h3 ? v[j + 4] : v[j], give = h3 ? v[j] : v[j + 4];
                    v[j] = sum8Lanes(keep + dpp<kDppRor8>(give));
                }
                // The 4 colour-b terms: swap-sums at 32 and 16 leave term b4 + 2 b5, then the row sum.
                float w0 = swapSum32(v[32], v[34]), w1 = swapSum32(v[33], v[35]);
                w0 = swapSum16(w0, w1);
                w0 = sum8Lanes(w0 + dpp<kDppRor8>(w0));
                const uint32_t jl = lane & 7u;
                if (jl < 4u) {  // lane j of a group adds term j + 4 b3 + 8 b4 + 16 b5
                    const float val = jl == 0u ? v[0] : jl == 1u ? v[1] : jl == 2u ? v[2] : v[3];
                    const uint32_t t = jl + 4u * ((lane >> 3) & 7u), u = t >> 3;
                    if (int(u) <= i) atomicAdd(&sG[t & 7u][sList[wave][i - int(u)]], val);
                }
                if ((lane & 15u) == 0u) {
                    const uint32_t u = (lane >> 4) & 3u;
                    if (int(u) <= i) atomicAdd(&sG[8][sList[wave][i - int(u)]], w0);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < n) {
            // The tile's sums of the moments -> the screen-space gradient: with power =
            // -(A dx^2 + 2 B dx dy + C dy^2) / 2, d/dcentre = -(A m_x + B m_y, B m_x + C m_y),
            // d/d(A, B, C) = -(m_xx / 2, m_xy, m_yy / 2), d/do = (sum of g) / o.
            const uint32_t g = values[range.x + begin + threadIdx.x];
            const float4 s0 = sS0[threadIdx.x];
            const float2 s1 = sS1[threadIdx.x];
            const float A = -2.0f * s0.z * (1.0f / kLog2e), Bc = -s0.w * (1.0f / kLog2e);
            const float Cc = -2.0f * s1.x * (1.0f / kLog2e);
            const float mx = sG[0][threadIdx.x], my = sG[1][threadIdx.x];
            float out[kGrad2D];
            out[0] = -(A * mx + Bc * my);
            out[1] = -(Bc * mx + Cc * my);
            out[2] = -0.5f * sG[2][threadIdx.x];
            out[3] = -sG[3][threadIdx.x];
            out[4] = -0.5f * sG[4][threadIdx.x];
            out[5] = sG[5][threadIdx.x] * __builtin_amdgcn_exp2f(-s1.y);
#pragma unroll
            for (int q = 6; q < kGrad2D; ++q) out[q] = sG[q][threadIdx.x];
#pragma unroll
            for (int q = 0; q < kGrad2D; ++q)
                if (out[q] != 0.0f) atomicAdd(&grad2D[size_t(q) * gaussianCount + g], out[q]);
        }
        end = begin;
    }
}

// The SH basis Y[k] at a direction (degree <= 3; Y[k] = 0 past the degree), and
// sum_k e[k] grad Y_k (the direction gradient of a colour with coefficients e).
__device__ __forceinline__ void shBasis(int degree, float x, float y, float z, float Y[16]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) Y[k] = 0.0f;
    Y[0] = kC0;
    if (degree > 0) {
        Y[1] = -kC1 * y;
        Y[2] = kC1 * z;
        Y[3] = -kC1 * x;
        if (degree > 1) {
            const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            Y[4] = kC2_0 * xy;
            Y[5] = kC2_1 * yz;
            Y[6] = kC2_2 * (2.0f * zz - xx - yy);
            Y[7] = kC2_3 * xz;
            Y[8] = kC2_4 * (xx - yy);
            if (degree > 2) {
                Y[9] = kC3_0 * y * (3.0f * xx - yy);
                Y[10] = kC3_1 * xy * z;
                Y[11] = kC3_2 * y * (4.0f * zz - xx - yy);
                Y[12] = kC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
                Y[13] = kC3_4 * x * (4.0f * zz - xx - yy);
                Y[14] = kC3_5 * z * (xx - yy);
                Y[15] = kC3_6 * x * (xx - 3.0f * yy);
            }
        }
    }
}
__device__ __forceinline__ void shGradDir(int degree, const float* c, float x, float y, float z, float grad[3]) {
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    if (degree > 0) {
        gy += -kC1 * c[1];
        gz += kC1 * c[2];
        gx += -kC1 * c[3];
        if (degree > 1) {
            const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            gx += kC2_0 * y * c[4] - 2.0f * kC2_2 * x * c[6] + kC2_3 * z * c[7] + 2.0f * kC2_4 * x * c[8];
            gy += kC2_0 * x * c[4] + kC2_1 * z * c[5] - 2.0f * kC2_2 * y * c[6] - 2.0f * kC2_4 * y * c[8];
            gz += kC2_1 * y * c[5] + 4.0f * kC2_2 * z * c[6] + kC2_3 * x * c[7];
            if (degree > 2) {
                gx += kC3_0 * 6.0f * xy * c[9] + kC3_1 * yz * c[10] - kC3_2 * 2.0f * xy * c[11] -
                      kC3_3 * 6.0f * xz * c[12] + kC3_4 * (4.0f * zz - 3.0f * xx - yy) * c[13] +
                      kC3_5 * 2.0f * xz * c[14] + kC3_6 * (3.0f * xx - 3.0f * yy) * c[15];
                gy += kC3_0 * (3.0f * xx - 3.0f * yy) * c[9] + kC3_1 * xz * c[10] +
                      kC3_2 * (4.0f * zz - xx - 3.0f * yy) * c[11] - kC3_3 * 6.0f * yz * c[12] -
                      kC3_4 * 2.0f * xy * c[13] - kC3_5 * 2.0f * yz * c[14] - kC3_6 * 6.0f * xy * c[15];
                gz += kC3_1 * xy * c[10] + kC3_2 * 8.0f * yz * c[11] + kC3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy) * c[12] +
                      kC3_4 * 8.0f * xz * c[13] + kC3_5 * (xx - yy) * c[14];
            }
        }
    }
    grad[0] = gx;
    grad[1] = gy;
    grad[2] = gz;
}

// Preprocess backward, one thread per gaussian (f32 scenes). From the screen-space gradient
// planes to the gradients of the mean, log-scales (s = exp), quaternion (as stored, used
// unnormalised), opacity logit (a = sigmoid) and SH coefficients (K x 3, coefficient-major RGB).
// Culled gaussians get zeros.
__global__ __launch_bounds__(256) void kPreprocessBackward(SceneRef scene, DevCamera cam,
                                                           const float4* __restrict__ splats,
                                                           const float* __restrict__ grad2D, float* __restrict__ dMean,
                                                           float* __restrict__ dLogScale, float* __restrict__ dQuat,
                                                           float* __restrict__ dLogit, float* __restrict__ dSh) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= scene.count) return;
    const size_t N = scene.count;
    const bool visible = __float_as_uint(splats[3 * size_t(i)].w) == GSR_CULL_VISIBLE;
    float gM[3] = {0.0f, 0.0f, 0.0f}, gS[3] = {0.0f, 0.0f, 0.0f}, gQ[4] = {0.0f, 0.0f, 0.0f, 0.0f}, gL = 0.0f;
    const int K = int(scene.K);
    if (!visible) {
        if (dMean)
            for (int c = 0; c < 3; ++c) dMean[3 * i + c] = 0.0f;
        if (dLogScale)
            for (int c = 0; c < 3; ++c) dLogScale[3 * i + c] = 0.0f;
        if (dQuat)
            for (int c = 0; c < 4; ++c) dQuat[4 * i + c] = 0.0f;
        if (dLogit) dLogit[i] = 0.0f;
        if (dSh)
            for (int k = 0; k < 3 * K; ++k) dSh[size_t(i) * 3 * K + k] = 0.0f;
        return;
    }
    float g2[kGrad2D];
#pragma unroll
    for (int q = 0; q < kGrad2D; ++q) g2[q] = grad2D[q * N + i];
    const float4 pa = scene.posAlpha[i], sc = scene.scales[i], qv = scene.rotations[i];
    const float mean[3] = {pa.x, pa.y, pa.z};
    const float* v = cam.view;
    float t[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) t[r] = v[r] * mean[0] + v[4 + r] * mean[1] + v[8 + r] * mean[2] + v[12 + r];
    // Opacity.
    gL = g2[5] * pa.w * (1.0f - pa.w);
    // Forward of the covariance (as preprocess.hip).
    const float s[3] = {sc.x, sc.y, sc.z};
    const float qw = qv.x, qx = qv.y, qy = qv.z, qz = qv.w;
    float R[3][3];
    R[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz);
    R[0][1] = 2.0f * (qx * qy + qw * qz);
    R[0][2] = 2.0f * (qx * qz - qw * qy);
    R[1][0] = 2.0f * (qx * qy - qw * qz);
    R[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz);
    R[1][2] = 2.0f * (qy * qz + qw * qx);
    R[2][0] = 2.0f * (qx * qz + qw * qy);
    R[2][1] = 2.0f * (qy * qz - qw * qx);
    R[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
    float M[3][3], S[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) M[c][r] = R[c][r] * s[c];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) S[j][k] = M[0][k] * M[0][j] + M[1][k] * M[1][j] + M[2][k] * M[2][j];
    const float limX = kFrustumClampFactor * cam.tanX, limY = kFrustumClampFactor * cam.tanY;
    const float txtz = t[0] / t[2], tytz = t[1] / t[2];
    const bool clampX = txtz < -limX || txtz > limX, clampY = tytz < -limY || tytz > limY;
    const float tx = (clampX ? (txtz < 0.0f ? -limX : limX) : txtz) * t[2];
    const float ty = (clampY ? (tytz < 0.0f ? -limY : limY) : tytz) * t[2];
    const float tz = t[2];
    const float j00 = cam.focalX / tz, j20 = -(cam.focalX * tx) / (tz * tz);
    const float j11 = cam.focalY / tz, j21 = -(cam.focalY * ty) / (tz * tz);
    float T[3][2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T[j][0] = j00 * v[j * 4 + 0] + j20 * v[j * 4 + 2];
        T[j][1] = j11 * v[j * 4 + 1] + j21 * v[j * 4 + 2];
    }
    float ST[3][2];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) ST[j][k] = S[j][0] * T[0][k] + S[j][1] * T[1][k] + S[j][2] * T[2][k];
    const float c00 = ST[0][0] * T[0][0] + ST[1][0] * T[1][0] + ST[2][0] * T[2][0];
    const float c01 = ST[0][1] * T[0][0] + ST[1][1] * T[1][0] + ST[2][1] * T[2][0];
    const float c11 = ST[0][1] * T[0][1] + ST[1][1] * T[1][1] + ST[2][1] * T[2][1];
    const float ca = c00 + kLowPassFilter, cb = c01, cc = c11 + kLowPassFilter;
    const float detInv = 1.0f / (ca * cc - cb * cb);
    const float ka = cc * detInv, kb = -cb * detInv, kc = ca * detInv;
    // Conic: K = C'^-1, dL/dC' = -K G K with G = [[gA, gB / 2], [gB / 2, gC]].
    const float G00 = g2[2], G01 = 0.5f * g2[3], G11 = g2[4];
    const float KG00 = ka * G00 + kb * G01, KG01 = ka * G01 + kb * G11;
    const float KG10 = kb * G00 + kc * G01, KG11 = kb * G01 + kc * G11;
    const float H00 = -(KG00 * ka + KG01 * kb), H01 = -(KG00 * kb + KG01 * kc), H11 = -(KG10 * kb + KG11 * kc);
    // cov2D = T^T S T: dL/dS = T H T^T, dL/dT = 2 S T H.
    float gSig[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            gSig[a][b] = T[a][0] * (H00 * T[b][0] + H01 * T[b][1]) + T[a][1] * (H01 * T[b][0] + H11 * T[b][1]);
    float gT[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        gT[a][0] = 2.0f * (ST[a][0] * H00 + ST[a][1] * H01);
        gT[a][1] = 2.0f * (ST[a][0] * H01 + ST[a][1] * H11);
    }
    // T = J W.
    float gj00 = 0.0f, gj20 = 0.0f, gj11 = 0.0f, gj21 = 0.0f;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        gj00 += gT[j][0] * v[j * 4 + 0];
        gj20 += gT[j][0] * v[j * 4 + 2];
        gj11 += gT[j][1] * v[j * 4 + 1];
        gj21 += gT[j][1] * v[j * 4 + 2];
    }
    float gt[3] = {0.0f, 0.0f, 0.0f};
    const float tz2 = tz * tz, tz3 = tz2 * tz;
    gt[2] += gj00 * (-cam.focalX / tz2) + gj20 * (2.0f * cam.focalX * tx / tz3) + gj11 * (-cam.focalY / tz2) +
             gj21 * (2.0f * cam.focalY * ty / tz3);
    const float gtx = gj20 * (-cam.focalX / tz2), gty = gj21 * (-cam.focalY / tz2);
    if (clampX) gt[2] += gtx * (txtz < 0.0f ? -limX : limX);
    else gt[0] += gtx;
    if (clampY) gt[2] += gty * (tytz < 0.0f ? -limY : limY);
    else gt[1] += gty;
    // Centre: ctr = ((t0 / (t2 tan) + 1) W - 1) / 2.
    const float hx = 0.5f * float(cam.width) / cam.tanX, hy = 0.5f * float(cam.height) / cam.tanY;
    gt[0] += g2[0] * hx / tz;
    gt[1] += g2[1] * hy / tz;
    gt[2] += -(g2[0] * hx * t[0] + g2[1] * hy * t[1]) / tz2;
    // Sigma = M^T M: dL/dM = 2 M gSig; M = diag(s) R.
    float gR[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float gs = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float gMcr = 2.0f * (M[c][0] * gSig[0][r] + M[c][1] * gSig[1][r] + M[c][2] * gSig[2][r]);
            gs += gMcr * R[c][r];
            gR[c][r] = gMcr * s[c];
        }
        gS[c] = gs * s[c];  // log-scale
    }
    gQ[0] = 2.0f * (gR[0][1] * qz - gR[0][2] * qy - gR[1][0] * qz + gR[1][2] * qx + gR[2][0] * qy - gR[2][1] * qx);
    gQ[1] = 2.0f * (gR[0][1] * qy + gR[0][2] * qz + gR[1][0] * qy - 2.0f * gR[1][1] * qx + gR[1][2] * qw +
                    gR[2][0] * qz - gR[2][1] * qw - 2.0f * gR[2][2] * qx);
    gQ[2] = 2.0f * (-2.0f * gR[0][0] * qy + gR[0][1] * qx - gR[0][2] * qw + gR[1][0] * qx + gR[1][2] * qz +
                    gR[2][0] * qw + gR[2][1] * qz - 2.0f * gR[2][2] * qy);
    gQ[3] = 2.0f * (-2.0f * gR[0][0] * qz + gR[0][1] * qw + gR[0][2] * qx - gR[1][0] * qw - 2.0f * gR[1][1] * qz +
                    gR[1][2] * qy + gR[2][0] * qx + gR[2][1] * qy);
    // Colour: raw = evalSH(dir) + 0.5, colour = max(raw, 0); dir = d / |d|, d = mean - camPos.
    const int degree = scene.degree;
    const int used = (degree + 1) * (degree + 1);
    float co[3][16];
    loadSh<GSR_SCENE_F32_SOA>(scene, i, used, co);
    const float dx = mean[0] - cam.pos[0], dy = mean[1] - cam.pos[1], dz = mean[2] - cam.pos[2];
    const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    const float dir[3] = {dx * inv, dy * inv, dz * inv};
    // The channels' colour gradients (0 where the colour is clamped), folded into one set of
    // coefficients e[k] = sum_ch gc[ch] co[ch][k]: one basis and one direction gradient.
    float gc[3], e[16];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float raw = evalShRaw(degree, co[ch], dir[0], dir[1], dir[2]);
        gc[ch] = raw < 0.0f ? 0.0f : g2[6 + ch];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) e[k] = gc[0] * co[0][k] + gc[1] * co[1][k] + gc[2] * co[2][k];
    float Y[16], gdir[3];
    shBasis(degree, dir[0], dir[1], dir[2], Y);
    shGradDir(degree, e, dir[0], dir[1], dir[2], gdir);
    if (dSh) {
        float* out = dSh + size_t(i) * K * 3;
        if (K == 16 && (reinterpret_cast<uintptr_t>(dSh) & 15u) == 0) {  // 48 floats per gaussian: 12 vector stores
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                float w[4];
#pragma unroll
                for (int l = 0; l < 4; ++l) w[l] = gc[(4 * q + l) % 3] * Y[(4 * q + l) / 3];
                reinterpret_cast<float4*>(out)[q] = make_float4(w[0], w[1], w[2], w[3]);
            }
        } else {
            for (int k = 0; k < K; ++k)
                for (int ch = 0; ch < 3; ++ch) out[k * 3 + ch] = gc[ch] * Y[k];
        }
    }
    const float dd = dir[0] * gdir[0] + dir[1] * gdir[1] + dir[2] * gdir[2];
    // Mean: through t (view rotation) and the direction.
#pragma unroll
    for (int c = 0; c < 3; ++c)
        gM[c] = v[c * 4 + 0] * gt[0] + v[c * 4 + 1] * gt[1] + v[c * 4 + 2] * gt[2] + (gdir[c] - dir[c] * dd) * inv;
    if (dMean)
        for (int c = 0; c < 3; ++c) dMean[3 * i + c] = gM[c];
    if (dLogScale)
        for (int c = 0; c < 3; ++c) dLogScale[3 * i + c] = gS[c];
    if (dQuat)
        for (int c = 0; c < 4; ++c) dQuat[4 * i + c] = gQ[c];
    if (dLogit) dLogit[i] = gL;
}

}  // namespace

int grad2DComponents() { return kGrad2D; }

void launchRenderTrain(const float4* splats, const uint32_t* values, const uint2* ranges, const DevCamera& cam,
                       float4* out, uint2* state, hipStream_t stream) {
    if (cam.tilesX * cam.tilesY == 0) return;
    hipLaunchKernelGGL(kRenderTrain, dim3(scheduledGrid(cam)), dim3(256), 0, stream, splats, values, ranges, cam, out,
                       state);
    markKernel("kRenderTrain", stream);
}

void launchRenderBackward(const float4* splats, const uint32_t* values, const uint2* ranges, const DevCamera& cam,
                          const float4* dOut, const uint2* state, float* grad2D, uint32_t count, hipStream_t stream) {
    if (cam.tilesX * cam.tilesY == 0) return;
    hipLaunchKernelGGL(kRenderBackward, dim3(scheduledGrid(cam)), dim3(256), 0, stream, splats, values, ranges, cam,
                       dOut, state, grad2D, count);
    markKernel("kRenderBackward", stream);
}

void launchPreprocessBackward(const SceneRef& scene, const DevCamera& cam, const float4* splats, const float* grad2D,
                              float* dMean, float* dLogScale, float* dQuat, float* dLogit, float* dSh,
                              hipStream_t stream) {
    const uint32_t blocks = (scene.count + 255) / 256;
    if (blocks == 0) return;
    hipLaunchKernelGGL(kPreprocessBackward, dim3(blocks), dim3(256), 0, stream, scene, cam, splats, grad2D, dMean,
                       dLogScale, dQuat, dLogit, dSh);
    markKernel("kPreprocessBackward", stream);
}

}  // namespace gsr
