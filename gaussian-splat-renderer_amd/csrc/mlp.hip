// Learned colour (no counterpart in the reference): a 16-dim latent per gaussian and one MLP shared
// by all gaussians, [latent(16), view direction(3), 1] -> 64 -> ReLU -> 64 -> ReLU -> 3, stand in
// for the degree-3 SH colour (48 floats per gaussian): colour = max(y + 0.5, 0), the transform the
// SH evaluation ends with. The training targets are the scene's own SH colours (evalSH,
// splat_math.cpp:129-151) for random (gaussian, direction) samples; gaussians whose learned
// colour misses the SH colour by more than a threshold in any of 32 test directions keep SH.
//
// gfx950 design: every product is v_mfma_f32_16x16x16_bf16 (bf16 operands, fp32 accumulation) in
// the "weights x activations" form, with the samples along N: a layer's output tile (lane l holds
// rows 4 (l / 16) + j of column l % 16) is already the next layer's B operand, so the forward pass
// moves no data between layers. The weights are packed once per update into the A-operand lane
// layout (lane l holds W[l % 16][4 (l / 16) + j]: one 8-byte load per lane per tile). The
// backward pass needs the activations and output gradients with the samples along K instead
// (dW = dY X^T): those tiles are transposed through a 1 KB per-wave LDS buffer.
#include <algorithm>

#include "gsr_device.hpp"

namespace gsr {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 parameter block: W1 [64][32] (input 19 = the constant 1: layer 1's bias), W2 [64][64],
// b2 [64], W3 [3][64], b3 [3].
constexpr int kW1 = 0;
constexpr int kW2 = kW1 + 64 * 32;
constexpr int kB2 = kW2 + 64 * 64;
constexpr int kW3 = kB2 + 64;
constexpr int kB3 = kW3 + 3 * 64;
// Packed A tiles (64 lanes x 4 bf16 each): the forward's W1 (4 x 2 tiles), W2 (4 x 4), W3 (rows
// padded to 16: 1 x 4), and the backward's W3^T (4 x 1), W2^T (4 x 4), W1^T rows 0..15 (1 x 4).
constexpr int kTA1 = 0, kTA2 = 8, kTA3 = 24, kTA3T = 28, kTA2T = 32, kTA1T = 48, kTACount = 52;
// The render-time forward's tiles for v_mfma_f32_16x16x32_bf16 (64 lanes x 8 bf16 each, after the
// 16x16x16 tiles): W1 (4 x 1), W2 (4 x 2), W3 (1 x 2). Lane l, slot j holds W[l % 16][k] with
// k = kb + 4 (l / 16) + j for j < 4 and kb + 16 + 4 (l / 16) + j - 4 for j >= 4 (kb = 32 x the
// tile's K index): the two D-layout tiles of rows kb .. kb + 15 and kb + 16 .. kb + 31 of the
// previous layer are then, side by side, the B operand with the same k in every slot. The sum
// over k does not depend on which k the hardware puts in a slot, as long as A and B agree.
constexpr int kTB1 = 0, kTB2 = 4, kTB3 = 12, kTBCount = 14;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kTestDirs = 32;  // fallback test directions per gaussian
constexpr int kParamCount = 64 * 32 + 64 * 64 + 64 + 3 * 64 + 3;

// fp32 -> bf16, round to nearest even (gfx950's v_cvt_pk_bf16_f32, two values per instruction).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t bf16Bits(float f) { return __builtin_bit_cast(uint16_t, __bf16(f)); }
__device__ __forceinline__ uint32_t bf16Pair(float a, float b) {  // one v_cvt_pk_bf16_f32
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}
__device__ __forceinline__ bf16x4 fromPairs(uint32_t lo, uint32_t hi) {
    return __builtin_bit_cast(bf16x4, make_uint2(lo, hi));
}
__device__ __forceinline__ bf16x4 toBf16(f32x4 v) { return fromPairs(bf16Pair(v[0], v[1]), bf16Pair(v[2], v[3])); }
// bf16(max(v, 0)) = max(bf16(v), 0): on the bf16 bit patterns as signed 16-bit integers, every
// negative value (and -0) is below 0 and the non-negative ones keep their order.
__device__ __forceinline__ uint32_t reluPair(uint32_t p) {
    const i16x2 v = __builtin_elementwise_max(__builtin_bit_cast(i16x2, p), i16x2{0, 0});
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ bf16x4 reluBf16(f32x4 v) {
    return fromPairs(reluPair(bf16Pair(v[0], v[1])), reluPair(bf16Pair(v[2], v[3])));
}
__device__ __forceinline__ bf16x4 tileA(const uint2* packs, int tile, uint32_t lane) {
    const uint2 u = packs[tile * 64 + lane];
    bf16x4 r;
    r[0] = short(u.x & 0xFFFFu);
    r[1] = short(u.x >> 16);
    r[2] = short(u.y & 0xFFFFu);
    r[3] = short(u.y >> 16);
    return r;
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 relu(f32x4 v) {
    f32x4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = v[j] > 0.0f ? v[j] : 0.0f;
    return r;
}

// Value j of lane `lane` in packed tile `tile` (the A-operand layout: W[l % 16][4 (l / 16) + j]).
__device__ float packedValue(const float* P, int tile, uint32_t lane, int j) {
    const int r = int(lane % 16u), k = int(4 * (lane / 16u)) + j;
    if (tile < kTA2) {  // W1: mi = tile / 2, ki = tile % 2
        return P[kW1 + (16 * (tile / 2) + r) * 32 + 16 * (tile % 2) + k];
    } else if (tile < kTA3) {  // W2
        const int t = tile - kTA2;
        return P[kW2 + (16 * (t / 4) + r) * 64 + 16 * (t % 4) + k];
    } else if (tile < kTA3T) {  // W3, rows 0..2 of 16
        return r < 3 ? P[kW3 + r * 64 + 16 * (tile - kTA3) + k] : 0.0f;
    } else if (tile < kTA2T) {  // W3^T: rows = hidden 16 mi + r, cols = output k
        return k < 3 ? P[kW3 + k * 64 + 16 * (tile - kTA3T) + r] : 0.0f;
    } else if (tile < kTA1T) {  // W2^T: rows = layer-1 unit 16 mi + r, cols = layer-2 unit 16 ki + k
        const int t = tile - kTA2T;
        return P[kW2 + (16 * (t % 4) + k) * 64 + 16 * (t / 4) + r];
    } else {  // W1^T rows 0..15: rows = input feature r, cols = layer-1 unit 16 ki + k
        return P[kW1 + (16 * (tile - kTA1T) + k) * 32 + r];
    }
}

// Value j of lane `lane` in K=32 tile `tile`.
__device__ float packedValue32(const float* P, int tile, uint32_t lane, int j) {
    const int r = int(lane % 16u), q = int(4 * (lane / 16u));
    const int k = j < 4 ? q + j : 16 + q + j - 4;
    if (tile < kTB2) return P[kW1 + (16 * tile + r) * 32 + k];  // W1: mi = tile
    if (tile < kTB3) {                                          // W2: mi = t / 2, K half t % 2
        const int t = tile - kTB2;
        return P[kW2 + (16 * (t / 2) + r) * 64 + 32 * (t % 2) + k];
    }
    return r < 3 ? P[kW3 + r * 64 + 32 * (tile - kTB3) + k] : 0.0f;  // W3, rows 0..2 of 16
}

// Blocks 0 .. kTACount - 1: the 16x16x16 tiles; then the kTBCount 16x16x32 tiles.
__global__ __launch_bounds__(64) void kPack(const float* __restrict__ params, uint2* __restrict__ packs) {
    const int tile = int(blockIdx.x);
    const uint32_t lane = threadIdx.x;
    if (tile < kTACount) {
        uint16_t h[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) h[j] = bf16Bits(packedValue(params, tile, lane, j));
        packs[tile * 64 + lane] =
            make_uint2(uint32_t(h[0]) | (uint32_t(h[1]) << 16), uint32_t(h[2]) | (uint32_t(h[3]) << 16));
    } else {
        const int t = tile - kTACount;
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            w[j] = uint32_t(bf16Bits(packedValue32(params, t, lane, 2 * j))) |
                   (uint32_t(bf16Bits(packedValue32(params, t, lane, 2 * j + 1))) << 16);
        reinterpret_cast<uint4*>(packs + kTACount * 64)[t * 64 + lane] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

struct Forward {
    f32x4 h1[4], h2[4], y;  // post-ReLU hidden tiles and the output tile (rows = units, cols = samples)
};

// x0: input features 0..15 (the latent), x1: features 16..31 ((dir, 1) in lanes 0..15, else 0),
// both as D-layout tiles of the input matrix (rows = features, cols = the 16 samples).
__device__ __forceinline__ void forward(const uint2* packs, const float* params, f32x4 x0, f32x4 x1, uint32_t lane,
                                        Forward& f) {
    const bf16x4 b0 = toBf16(x0), b1 = toBf16(x1);
    bf16x4 b[4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma(tileA(packs, kTA1 + 2 * mi, lane), b0, acc);
        acc = mfma(tileA(packs, kTA1 + 2 * mi + 1, lane), b1, acc);
        f.h1[mi] = relu(acc);
        b[mi] = toBf16(f.h1[mi]);
    }
    const int rq = int(4 * (lane / 16u));
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        f32x4 acc;
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = params[kB2 + 16 * mi + rq + j];
#pragma unroll
        for (int ki = 0; ki < 4; ++ki) acc = mfma(tileA(packs, kTA2 + 4 * mi + ki, lane), b[ki], acc);
        f.h2[mi] = relu(acc);
    }
    bf16x4 c[4];
#pragma unroll
    for (int ki = 0; ki < 4; ++ki) c[ki] = toBf16(f.h2[ki]);
    f32x4 acc;
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = rq + j < 3 ? params[kB3 + rq + j] : 0.0f;
#pragma unroll
    for (int ki = 0; ki < 4; ++ki) acc = mfma(tileA(packs, kTA3 + ki, lane), c[ki], acc);
    f.y = acc;
}

// The render-time forward (v_mfma_f32_16x16x32_bf16, 14 per 16 samples) with its weights held in
// registers for a wave's lifetime (56 VGPRs of A tiles + 20 of biases): the kernels that use it
// read no weight memory per tile.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
struct Weights {
    bf16x8 w1[4], w2[8], w3[2];
    f32x4 b2[4], b3;
};
__device__ __forceinline__ bf16x8 concat(bf16x4 lo, bf16x4 hi) {
    return __builtin_bit_cast(bf16x8, make_uint4(__builtin_bit_cast(uint2, lo).x, __builtin_bit_cast(uint2, lo).y,
                                                 __builtin_bit_cast(uint2, hi).x, __builtin_bit_cast(uint2, hi).y));
}
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void loadWeights(const uint2* packs, const float* params, uint32_t lane, Weights& w) {
    const uint4* t32 = reinterpret_cast<const uint4*>(packs + kTACount * 64);
#pragma unroll
    for (int t = 0; t < 4; ++t) w.w1[t] = __builtin_bit_cast(bf16x8, t32[(kTB1 + t) * 64 + lane]);
#pragma unroll
    for (int t = 0; t < 8; ++t) w.w2[t] = __builtin_bit_cast(bf16x8, t32[(kTB2 + t) * 64 + lane]);
#pragma unroll
    for (int t = 0; t < 2; ++t) w.w3[t] = __builtin_bit_cast(bf16x8, t32[(kTB3 + t) * 64 + lane]);
    const int rq = int(4 * (lane / 16u));
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int j = 0; j < 4; ++j) w.b2[mi][j] = params[kB2 + 16 * mi + rq + j];
#pragma unroll
    for (int j = 0; j < 4; ++j) w.b3[j] = rq + j < 3 ? params[kB3 + rq + j] : 0.0f;
}
// y (lane l < 16: the 3 outputs of sample l in y[0..2]). The same bf16 operands as forward(); the
// fp32 sums may round differently (32 products per instruction instead of 16).
__device__ __forceinline__ f32x4 forwardY(const Weights& w, f32x4 x0, f32x4 x1) {
    const bf16x8 bx = concat(toBf16(x0), toBf16(x1));
    bf16x4 h[4], c[4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) h[mi] = reluBf16(mfma32(w.w1[mi], bx, f32x4{0.0f, 0.0f, 0.0f, 0.0f}));
    const bf16x8 bh0 = concat(h[0], h[1]), bh1 = concat(h[2], h[3]);
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) c[mi] = reluBf16(mfma32(w.w2[2 * mi + 1], bh1, mfma32(w.w2[2 * mi], bh0, w.b2[mi])));
    const bf16x8 bc0 = concat(c[0], c[1]), bc1 = concat(c[2], c[3]);
    return mfma32(w.w3[1], bc1, mfma32(w.w3[0], bc0, w.b3));
}

// A D-layout tile of a 16 x 16 matrix M (lane l: M[4 (l / 16) + j][l % 16]) re-read in the
// A pattern (lane l: M[l % 16][4 (l / 16) + j]) through the wave's LDS buffer (16 x 17 floats).
__device__ __forceinline__ bf16x4 transposeTile(f32x4 d, float* lds, uint32_t lane) {
    const uint32_t rq = 4 * (lane / 16u), c = lane % 16u;
#pragma unroll
    for (int j = 0; j < 4; ++j) lds[(rq + j) * 17 + c] = d[j];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    f32x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = lds[c * 17 + rq + j];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return toBf16(a);
}

__device__ __forceinline__ float4 bf16Quad(const uint16_t* p) {  // 4 bf16 values, 8-byte aligned
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xFFFF0000u), __uint_as_float(u.y << 16),
                       __uint_as_float(u.y & 0xFFFF0000u));
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {  // a 32-bit integer hash (splitmix-style)
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ float unit01(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// Sample s of a stream: a gaussian and a direction uniform on the sphere. The gaussians follow the
// stream's permutation g = (a s + b) mod count (a coprime to count), so samples 0 .. count - 1 of a
// stream are distinct gaussians: a training step of at most count samples touches each once.
__device__ __forceinline__ void sampleOf(const SampleStream& st, uint32_t s, uint32_t* g, float dir[3]) {
    const uint32_t h0 = mix32(s * 0x9E3779B9u ^ st.seed), h1 = mix32(h0 ^ 0x85EBCA6Bu), h2 = mix32(h1 ^ 0xC2B2AE35u);
    *g = uint32_t((uint64_t(st.mulA) * (s % st.count) + st.addB) % st.count);
    const float z = 2.0f * unit01(h1) - 1.0f, phi = 6.28318530718f * unit01(h2);
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    dir[0] = r * cosf(phi);
    dir[1] = r * sinf(phi);
    dir[2] = z;
}

// Test direction d of kTestDirs (a Fibonacci sphere).
__device__ __forceinline__ void testDir(uint32_t d, float dir[3]) {
    const float z = 1.0f - (2.0f * float(d) + 1.0f) / float(kTestDirs);
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z)), phi = 2.39996322973f * float(d);
    dir[0] = r * cosf(phi);
    dir[1] = r * sinf(phi);
    dir[2] = z;
}

template <int FMT>
__device__ __forceinline__ void shColour(const SceneRef& scene, uint32_t g, const float dir[3], float out[3]) {
    const int used = (scene.degree + 1) * (scene.degree + 1);
    float co[3][16];
    loadSh<FMT>(scene, g, used, co);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = evalShChannel(scene.degree, co[c], dir[0], dir[1], dir[2]);  // clamped
}
// The unclamped SH value (y + 0.5 is trained towards it; both are clamped at 0 when used).
template <int FMT>
__device__ __forceinline__ void shTarget(const SceneRef& scene, uint32_t g, const float dir[3], float out[3]) {
    const int used = (scene.degree + 1) * (scene.degree + 1);
    float co[3][16];
    loadSh<FMT>(scene, g, used, co);
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = evalShRaw(scene.degree, co[c], dir[0], dir[1], dir[2]);
}

// The inputs of one 16-sample tile as D-layout tiles: x0 = latent features 4 (l / 16) .. + 3 of
// sample l % 16 (latQuad), x1 = (dir, 1) in lanes 0..15.
__device__ __forceinline__ void inputTiles(float4 latQuad, const float dir[3], uint32_t lane, f32x4& x0, f32x4& x1) {
    x0 = f32x4{latQuad.x, latQuad.y, latQuad.z, latQuad.w};
    const bool first = lane < 16u;
    x1[0] = first ? dir[0] : 0.0f;
    x1[1] = first ? dir[1] : 0.0f;
    x1[2] = first ? dir[2] : 0.0f;
    x1[3] = first ? 1.0f : 0.0f;
}

// Training pass over samples 0 .. total - 1 of a stream, `perWave` consecutive ones (a multiple of
// 16) per wave, for the loss mean((y + 0.5 - sh)^2) over the total x 3 values; the summed squared
// error into *sse. MODE kTrainLoss: only that. Otherwise the weight gradient: the workgroup's
// waves add theirs in LDS and the workgroup writes its sum to partials[blockIdx.x] (P floats; no
// global atomics; kReduceGrad adds the rows). The latent gradient: kTrainGrad adds it into
// gradLatents (atomics: a gaussian may occur in several samples); kTrainStep (every gaussian of
// the pass distinct: total <= count) applies lazy Adam to each sample's latent in place instead
// (m, v and the latent of the touched gaussians only, bias corrections bc1, bc2 of the step).
template <int FMT, int MODE>
__global__ __launch_bounds__(kThreads) void kTrain(SceneRef scene, const float* __restrict__ params,
                                                   const uint2* __restrict__ packs, float* latents,
                                                   float* __restrict__ partials, float* __restrict__ gradLatents,
                                                   float* __restrict__ mLat, float* __restrict__ vLat, float lrLat,
                                                   float bc1, float bc2, float* __restrict__ sse, SampleStream st,
                                                   uint32_t perWave, uint32_t total) {
    constexpr bool BACKWARD = MODE != kTrainLoss;
    __shared__ float sT[kWaves][16 * 17];
    __shared__ float sGrad[BACKWARD ? kParamCount : 1];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if constexpr (BACKWARD) {
        for (uint32_t i = threadIdx.x; i < uint32_t(kParamCount); i += kThreads) sGrad[i] = 0.0f;
        __syncthreads();
    }
    float* lds = sT[wave];
    const uint32_t rq = 4 * (lane / 16u), col = lane % 16u;
    const float scale = 2.0f / (3.0f * float(total));
    f32x4 dW1[4][2], dW2[4][4], dW3[4];
    float db2[4][4], db3[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        dW3[a] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < 2; ++b) dW1[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < 4; ++b) dW2[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 4; ++j) db2[a][j] = 0.0f;
        db3[a] = 0.0f;
    }
    float err = 0.0f;
    const uint32_t base = (blockIdx.x * kWaves + wave) * perWave;
    for (uint32_t t0 = 0; t0 < perWave && base + t0 < total; t0 += 16) {  // whole tiles: total % 16 == 0
        uint32_t g;
        float dir[3];
        sampleOf(st, base + t0 + col, &g, dir);
        f32x4 x0, x1;
        const float4 lat = *reinterpret_cast<const float4*>(latents + size_t(g) * 16 + rq);
        inputTiles(lat, dir, lane, x0, x1);
        Forward f;
        forward(packs, params, x0, x1, lane, f);
        f32x4 dy = {0.0f, 0.0f, 0.0f, 0.0f};
        if (lane < 16u) {
            float tgt[3];
            shTarget<FMT>(scene, g, dir, tgt);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float e = f.y[c] + 0.5f - tgt[c];
                err += e * e;
                dy[c] = scale * e;
            }
        }
        if (!BACKWARD) continue;
        // dW3 += dY H2^T; dH2 = W3^T dY (masked by ReLU).
        const bf16x4 aDy = transposeTile(dy, lds, lane);
        bf16x4 bH2[4], bH1[4];
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) {
            bH2[ni] = transposeTile(f.h2[ni], lds, lane);
            dW3[ni] = mfma(aDy, bH2[ni], dW3[ni]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) db3[j] += dy[j];
        const bf16x4 bDy = toBf16(dy);
        f32x4 dH2[4];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            f32x4 d = mfma(tileA(packs, kTA3T + mi, lane), bDy, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                d[j] = f.h2[mi][j] > 0.0f ? d[j] : 0.0f;
                db2[mi][j] += d[j];
            }
            dH2[mi] = d;
        }
        // dW2 += dH2 H1^T; dH1 = W2^T dH2 (masked).
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) bH1[ni] = transposeTile(f.h1[ni], lds, lane);
        bf16x4 bDH2[4];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            const bf16x4 a = transposeTile(dH2[mi], lds, lane);
#pragma unroll
            for (int ni = 0; ni < 4; ++ni) dW2[mi][ni] = mfma(a, bH1[ni], dW2[mi][ni]);
            bDH2[mi] = toBf16(dH2[mi]);
        }
        f32x4 dH1[4];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ki = 0; ki < 4; ++ki) d = mfma(tileA(packs, kTA2T + 4 * mi + ki, lane), bDH2[ki], d);
#pragma unroll
            for (int j = 0; j < 4; ++j) d[j] = f.h1[mi][j] > 0.0f ? d[j] : 0.0f;
            dH1[mi] = d;
        }
        // dW1 += dH1 X^T; dLatent = W1^T[0..15] dH1.
        const bf16x4 bX0 = transposeTile(x0, lds, lane), bX1 = transposeTile(x1, lds, lane);
        bf16x4 bDH1[4];
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            const bf16x4 a = transposeTile(dH1[mi], lds, lane);
            dW1[mi][0] = mfma(a, bX0, dW1[mi][0]);
            dW1[mi][1] = mfma(a, bX1, dW1[mi][1]);
            bDH1[mi] = toBf16(dH1[mi]);
        }
        f32x4 dl = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ki = 0; ki < 4; ++ki) dl = mfma(tileA(packs, kTA1T + ki, lane), bDH1[ki], dl);
        if constexpr (MODE == kTrainGrad) {
#pragma unroll
            for (int j = 0; j < 4; ++j) atomicAdd(&gradLatents[size_t(g) * 16 + rq + j], dl[j]);
        } else {  // lazy Adam on this sample's latent entries (the only sample of gaussian g in the pass)
            constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
            const size_t li = size_t(g) * 16 + rq;
            float4 m4 = *reinterpret_cast<const float4*>(mLat + li), v4 = *reinterpret_cast<const float4*>(vLat + li);
            float mv[4] = {m4.x, m4.y, m4.z, m4.w}, vv[4] = {v4.x, v4.y, v4.z, v4.w};
            float pv[4] = {lat.x, lat.y, lat.z, lat.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                mv[j] = b1 * mv[j] + (1.0f - b1) * dl[j];
                vv[j] = b2 * vv[j] + (1.0f - b2) * dl[j] * dl[j];
                pv[j] -= lrLat * (mv[j] / bc1) / (sqrtf(vv[j] / bc2) + eps);
            }
            *reinterpret_cast<float4*>(mLat + li) = make_float4(mv[0], mv[1], mv[2], mv[3]);
            *reinterpret_cast<float4*>(vLat + li) = make_float4(vv[0], vv[1], vv[2], vv[3]);
            *reinterpret_cast<float4*>(latents + li) = make_float4(pv[0], pv[1], pv[2], pv[3]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) err += __shfl_xor(err, off, 64);
    if (lane == 0) atomicAdd(sse, err);
    if constexpr (!BACKWARD) return;
    // The workgroup's weight gradient: every wave adds its entries into LDS.
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = 16 * mi + int(rq) + j;
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) atomicAdd(&sGrad[kW1 + row * 32 + 16 * ni + int(col)], dW1[mi][ni][j]);
#pragma unroll
            for (int ni = 0; ni < 4; ++ni) atomicAdd(&sGrad[kW2 + row * 64 + 16 * ni + int(col)], dW2[mi][ni][j]);
        }
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (int(rq) + j < 3) atomicAdd(&sGrad[kW3 + (int(rq) + j) * 64 + 16 * ni + int(col)], dW3[ni][j]);
    // Bias sums over the 16 sample columns (lanes with the same l / 16).
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float v = db2[mi][j];
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (col == 0u) atomicAdd(&sGrad[kB2 + 16 * mi + int(rq) + j], v);
        }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        float v = db3[j];
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0u) atomicAdd(&sGrad[kB3 + j], v);
    }
    __syncthreads();
    float* row = partials + size_t(blockIdx.x) * kParamCount;
    for (uint32_t i = threadIdx.x; i < uint32_t(kParamCount); i += kThreads) row[i] = sGrad[i];
}

// grad[i] = the sum of the workgroups' rows partials[b][i], b < rows: a workgroup per 64 entries,
// its 4 waves summing every 4th row (8 loads in flight per lane), then added in LDS.
__global__ __launch_bounds__(256) void kReduceGrad(const float* __restrict__ partials, uint32_t rows,
                                                   float* __restrict__ grad) {
    __shared__ float sSum[4][64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 64u + lane;
    float s = 0.0f;
    if (i < uint32_t(kParamCount)) {
        uint32_t b = wave;
        for (; b + 28u < rows; b += 32u) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partials[size_t(b + 4u * u) * kParamCount + i];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; b < rows; b += 4u) s += partials[size_t(b) * kParamCount + i];
    }
    sSum[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && i < uint32_t(kParamCount)) grad[i] = (sSum[0][lane] + sSum[1][lane]) + (sSum[2][lane] + sSum[3][lane]);
}

// Adam on n values; the gradients are cleared for the next step.
__global__ __launch_bounds__(256) void kAdam(float* __restrict__ p, float* __restrict__ grad, float* __restrict__ m,
                                             float* __restrict__ v, uint32_t n, float lr, float bc1, float bc2) {
    constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float gi = grad[i];
        grad[i] = 0.0f;
        const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

// The latents the renderer reads: bf16 (rounded as the MFMA operands are), 32 bytes per gaussian.
__global__ __launch_bounds__(256) void kToBf16(const float* __restrict__ in, uint16_t* __restrict__ out, uint32_t n) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) out[i] = bf16Bits(in[i]);
}

// Latent initialisation: uniform in [-scale, scale) from a hash of (seed, index).
__global__ __launch_bounds__(256) void kInitLatents(float* __restrict__ out, uint32_t n, uint32_t seed, float scale) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        out[i] = scale * (2.0f * unit01(mix32(i * 0x9E3779B9u ^ mix32(seed))) - 1.0f);
}

// Latent initialisation from the scene's SH (principal components). kShMoments: per workgroup,
// the sums of the D = 3 (degree + 1)^2 SH values x (feature f = 3 k + c, the stored order) and of
// the products x_i x_j (i <= j, row-major upper triangle) over its gaussians, staged 64 gaussians
// at a time in LDS; row b of `partials` = D sums, then D (D + 1) / 2 product sums.
constexpr int kMomentRows = 64;
constexpr int kMaxShValues = 48;
template <int FMT>
__global__ __launch_bounds__(256) void kShMoments(SceneRef scene, float* __restrict__ partials) {
    __shared__ float sX[kMomentRows][kMaxShValues + 1];
    const int used = (scene.degree + 1) * (scene.degree + 1), D = used * 3, pairs = D * (D + 1) / 2;
    const uint32_t t = threadIdx.x;
    constexpr int kPer = (kMaxShValues * (kMaxShValues + 1) / 2 + 255) / 256;  // 5 pairs per thread
    int pi[kPer], pj[kPer];
    float acc[kPer], sum = 0.0f;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        int p = int(t) + q * 256, i = 0;
        const int pp = p;
        while (p >= D - i && i < D) {  // row i of the upper triangle holds D - i pairs
            p -= D - i;
            ++i;
        }
        pi[q] = pp < pairs ? i : 0;
        pj[q] = pp < pairs ? i + p : 0;
        acc[q] = 0.0f;
    }
    for (uint32_t g0 = blockIdx.x * kMomentRows; g0 < scene.count; g0 += gridDim.x * kMomentRows) {
        if (t < uint32_t(kMomentRows)) {
            const uint32_t g = g0 + t;
            float co[3][16];
            if (g < scene.count) {
                loadSh<FMT>(scene, g, used, co);
            } else {
#pragma unroll
                for (int k = 0; k < 16; ++k) co[0][k] = co[1][k] = co[2][k] = 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 16; ++k)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (3 * k + c < kMaxShValues) sX[t][3 * k + c] = co[c][k];
        }
        __syncthreads();
        for (int r = 0; r < kMomentRows; ++r) {
#pragma unroll
            for (int q = 0; q < kPer; ++q) acc[q] = fmaf(sX[r][pi[q]], sX[r][pj[q]], acc[q]);
            if (int(t) < D) sum += sX[r][t];
        }
        __syncthreads();
    }
    float* row = partials + size_t(blockIdx.x) * size_t(D + pairs);
    if (int(t) < D) row[t] = sum;
#pragma unroll
    for (int q = 0; q < kPer; ++q)
        if (int(t) + q * 256 < pairs) row[D + int(t) + q * 256] = acc[q];
}

// latents[g][r] = proj[r] . (x_g - mu) for the 16 rows of proj (D values each, stored 48 apart).
template <int FMT>
__global__ __launch_bounds__(256) void kInitLatentsPca(SceneRef scene, const float* __restrict__ projMu,
                                                       float* __restrict__ latents) {
    __shared__ float sP[16 * kMaxShValues + kMaxShValues];
    for (uint32_t i = threadIdx.x; i < 16 * kMaxShValues + kMaxShValues; i += 256) sP[i] = projMu[i];
    __syncthreads();
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= scene.count) return;
    const int used = (scene.degree + 1) * (scene.degree + 1);
    float co[3][16];
    loadSh<FMT>(scene, g, used, co);
    float x[kMaxShValues];
#pragma unroll
    for (int k = 0; k < 16; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) x[3 * k + c] = co[c][k] - sP[16 * kMaxShValues + 3 * k + c];
    float out[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        float a = 0.0f;
#pragma unroll
        for (int f = 0; f < kMaxShValues; ++f) a = fmaf(sP[r * kMaxShValues + f], x[f], a);
        out[r] = a;
    }
    float4* dst = reinterpret_cast<float4*>(latents + size_t(g) * 16);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = make_float4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
}

// The (gaussian, direction) pairs of samples first .. first + n - 1 of stream `seed`.
__global__ __launch_bounds__(256) void kSamples(SampleStream st, uint32_t first, uint32_t n,
                                                uint32_t* __restrict__ gaussians, float* __restrict__ dirs) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t g;
    float dir[3];
    sampleOf(st, first + i, &g, dir);
    gaussians[i] = g;
    dirs[3 * size_t(i)] = dir[0];
    dirs[3 * size_t(i) + 1] = dir[1];
    dirs[3 * size_t(i) + 2] = dir[2];
}

// Fallback: per gaussian, the largest channel difference between max(y + 0.5, 0) (with the bf16
// latent the renderer uses) and the clamped SH colour over the kTestDirs test directions;
// useMlp[g] = 1 when it is at most `threshold`, and err[g] holds the difference. A wave takes
// two gaussians per round (2 x 32 directions = 4 tiles).
template <int FMT>
__global__ __launch_bounds__(kThreads) void kFallback(SceneRef scene, const float* __restrict__ params,
                                                      const uint2* __restrict__ packs,
                                                      const uint16_t* __restrict__ latentsBf16, float threshold,
                                                      uint8_t* __restrict__ useMlp, float* __restrict__ errOut) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t col = lane % 16u, rq = 4 * (lane / 16u);
    const uint32_t waves = gridDim.x * kWaves;
    Weights w;
    loadWeights(packs, params, lane, w);
    for (uint32_t g0 = (blockIdx.x * kWaves + wave) * 2; g0 < scene.count; g0 += waves * 2) {
        float worst[2] = {0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t g = g0 + uint32_t(t / 2);
            const uint32_t gs = g < scene.count ? g : scene.count - 1;
            float dir[3];
            testDir(uint32_t(t % 2) * 16 + col, dir);
            f32x4 x0, x1;
            inputTiles(bf16Quad(latentsBf16 + size_t(gs) * 16 + rq), dir, lane, x0, x1);
            const f32x4 y = forwardY(w, x0, x1);
            float e = 0.0f;
            if (lane < 16u) {
                float sh[3];
                shColour<FMT>(scene, gs, dir, sh);
#pragma unroll
                for (int c = 0; c < 3; ++c) e = fmaxf(e, fabsf(fmaxf(y[c] + 0.5f, 0.0f) - sh[c]));
            }
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) e = fmaxf(e, __shfl_xor(e, off, 64));
            worst[t / 2] = fmaxf(worst[t / 2], e);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const uint32_t g = g0 + uint32_t(k);
                if (g < scene.count) {
                    useMlp[g] = worst[k] <= threshold ? 1u : 0u;
                    errOut[g] = worst[k];
                }
            }
        }
    }
}

// The render's colour for the gaussians that use the MLP and cover a tile (after the preprocess,
// which left their colour at 0): dir = normalize(mean - camPos) as for SH; writes the splat's
// o2.xyz and, in the packed render-splat modes, the fp16 colour of the packed record. Persistent
// waves with the weights in registers; a wave takes 4 tiles of 16 consecutive gaussians per
// round, issues all their loads first (the round is latency-bound otherwise) and skips the MFMAs
// of a tile none of whose gaussians needs the MLP.
constexpr int kColourTiles = 4;
template <int FMT>
__global__ __launch_bounds__(kThreads) void kColourMlp(SceneRef scene, DevCamera cam, const float* __restrict__ params,
                                                       const uint2* __restrict__ packs,
                                                       const uint16_t* __restrict__ latentsBf16,
                                                       const uint8_t* __restrict__ useMlp,
                                                       const uint32_t* __restrict__ tileCounts,
                                                       float4* __restrict__ splats, uint32_t renderSplat,
                                                       uint32_t* __restrict__ packed) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t col = lane % 16u, rq = 4 * (lane / 16u);
    const uint32_t waves = gridDim.x * kWaves, rounds = (scene.count + 16 * kColourTiles - 1) / (16 * kColourTiles);
    Weights w;
    loadWeights(packs, params, lane, w);
    for (uint32_t round = blockIdx.x * kWaves + (threadIdx.x >> 6); round < rounds; round += waves) {
        // Every load of the round first, unconditionally (no load waits on another's value), and
        // all of them in registers before the one skip test: one memory round trip per round.
        uint32_t mark[kColourTiles], cover[kColourTiles];
        float3 pos[kColourTiles];
        float4 lat[kColourTiles];
#pragma unroll
        for (int u = 0; u < kColourTiles; ++u) {
            const uint32_t g = (round * kColourTiles + uint32_t(u)) * 16 + col;
            const uint32_t gs = g < scene.count ? g : scene.count - 1;
            mark[u] = useMlp[gs];
            cover[u] = tileCounts[gs];
            if (FMT == GSR_SCENE_F32_SOA) {
                const float4 pa = scene.posAlpha[gs];
                pos[u] = make_float3(pa.x, pa.y, pa.z);
            } else {
                const uint4 gv = scene.geometry[gs];
                pos[u] = make_float3(__uint_as_float(gv.x), __uint_as_float(gv.y), __uint_as_float(gv.z));
            }
            lat[u] = bf16Quad(latentsBf16 + size_t(gs) * 16 + rq);
        }
        bool need[kColourTiles], any = false;
#pragma unroll
        for (int u = 0; u < kColourTiles; ++u) {
            const uint32_t g = (round * kColourTiles + uint32_t(u)) * 16 + col;
            need[u] = (g < scene.count) & (mark[u] != 0u) & (cover[u] != 0u);
            any |= need[u];
            asm volatile("" ::"v"(pos[u].x), "v"(pos[u].y), "v"(pos[u].z), "v"(lat[u].x), "v"(lat[u].y),
                         "v"(lat[u].z), "v"(lat[u].w));
        }
        if (!__any(any)) continue;  // wave-uniform: no gaussian of the round needs the MLP
#pragma unroll
        for (int u = 0; u < kColourTiles; ++u) {
            const uint32_t g = (round * kColourTiles + uint32_t(u)) * 16 + col;
            const float dx = pos[u].x - cam.pos[0], dy = pos[u].y - cam.pos[1], dz = pos[u].z - cam.pos[2];
            const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
            const float dir[3] = {dx * inv, dy * inv, dz * inv};
            f32x4 x0, x1;
            inputTiles(lat[u], dir, lane, x0, x1);
            const f32x4 y = forwardY(w, x0, x1);  // every lane: the MFMAs need the whole wave
            if (lane < 16u && need[u]) {
                const float r = fmaxf(y[0] + 0.5f, 0.0f), gr = fmaxf(y[1] + 0.5f, 0.0f),
                            b = fmaxf(y[2] + 0.5f, 0.0f);
                *reinterpret_cast<float3*>(splats + 3 * size_t(g) + 2) = make_float3(r, gr, b);
                const uint32_t rg = halfBits(r) | (halfBits(gr) << 16);
                if (renderSplat == GSR_RENDER_SPLAT_P32) {
                    uint32_t* pw = packed + 8 * size_t(g);
                    pw[6] = rg;
                    pw[7] = (pw[7] & 0xFFFF0000u) | halfBits(b);  // the radius stays in the high half
                } else if (renderSplat == GSR_RENDER_SPLAT_P24) {
                    uint32_t* pw = packed + 6 * size_t(g);
                    pw[4] = rg;
                    pw[5] = (pw[5] & 0xFFFF0000u) | halfBits(b);
                }
            }
        }
    }
}

// Evaluation at given (gaussian, direction) pairs: y + 0.5 unclamped (tests).
__global__ __launch_bounds__(kThreads) void kEvaluate(const float* __restrict__ params, const uint2* __restrict__ packs,
                                                      const float* __restrict__ latents,
                                                      const uint32_t* __restrict__ gaussians,
                                                      const float* __restrict__ dirs, uint32_t n,
                                                      float* __restrict__ rgb) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t col = lane % 16u;
    const uint32_t t0 = (blockIdx.x * kWaves + wave) * 16;
    if (t0 >= n) return;  // whole waves
    const uint32_t s = t0 + col < n ? t0 + col : n - 1;
    const float dir[3] = {dirs[3 * s], dirs[3 * s + 1], dirs[3 * s + 2]};
    f32x4 x0, x1;
    inputTiles(*reinterpret_cast<const float4*>(latents + size_t(gaussians[s]) * 16 + 4 * (lane / 16u)), dir, lane,
               x0, x1);
    Weights w;
    loadWeights(packs, params, lane, w);
    const f32x4 y = forwardY(w, x0, x1);
    if (lane < 16u && t0 + col < n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) rgb[3 * size_t(s) + c] = y[c] + 0.5f;
    }
}

uint32_t gridFor(uint32_t n, uint32_t perBlock, uint32_t cap) {
    const uint32_t b = (n + perBlock - 1) / perBlock;
    return b > cap ? cap : b;
}

}  // namespace

static_assert(kParamCount == kB3 + 3, "the parameter block");
int mlpParamCount() { return kParamCount; }

SampleStream sampleStream(uint32_t seed, uint32_t count) {
    // a: a hash of the seed reduced to [1, count), moved up to the next value coprime to count
    // (a permutation of 0 .. count - 1); b: another hash.
    auto h = [](uint64_t x) {
        x ^= x >> 33;
        x *= 0xFF51AFD7ED558CCDull;
        x ^= x >> 33;
        x *= 0xC4CEB9FE1A85EC53ull;
        x ^= x >> 33;
        return x;
    };
    SampleStream st{seed, 1u, 0u, count ? count : 1u};
    if (count > 1) {
        uint32_t a = uint32_t(h(uint64_t(seed) * 2 + 1) % (count - 1)) + 1u;
        auto gcd = [](uint32_t x, uint32_t y) {
            while (y) {
                const uint32_t t = x % y;
                x = y;
                y = t;
            }
            return x;
        };
        while (gcd(a, count) != 1u) a = a + 1u < count ? a + 1u : 1u;
        st.mulA = a;
        st.addB = uint32_t(h(uint64_t(seed) * 2 + 2) % count);
    }
    return st;
}
int mlpPackedTiles() { return kTACount + 2 * kTBCount; }  // units of 64 x uint2

void launchMlpPack(const float* params, uint2* packs, hipStream_t stream) {
    hipLaunchKernelGGL(kPack, dim3(kTACount + kTBCount), dim3(64), 0, stream, params, packs);
    markKernel("kPack", stream);
}

uint32_t mlpTrainBlocks(uint32_t samples) {
    const uint32_t tiles = samples / 16;
    const uint32_t tilesPerWave = std::max(2u, (tiles + 2047) / 2048);
    const uint32_t waves = (tiles + tilesPerWave - 1) / tilesPerWave;
    return (waves + kWaves - 1) / kWaves;
}

void launchMlpTrain(const SceneRef& scene, const float* params, const uint2* packs, float* latents, float* gradParams,
                    float* partials, float* gradLatents, float* mLat, float* vLat, float lrLat, uint32_t step,
                    float* sse, const SampleStream& st, uint32_t samples, int mode, hipStream_t stream) {
    // Each wave takes perWave consecutive samples (a multiple of 16, at least 32); at most 2048
    // waves (512 workgroups, the rows of `partials`). samples is a multiple of 16.
    const uint32_t tiles = samples / 16;
    if (tiles == 0 || scene.count == 0) return;
    const uint32_t tilesPerWave = std::max(2u, (tiles + 2047) / 2048);
    const uint32_t blocks = mlpTrainBlocks(samples);
    const uint32_t perWave = 16 * tilesPerWave;
    const float bc1 = 1.0f - powf(0.9f, float(step)), bc2 = 1.0f - powf(0.999f, float(step));
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kThreads), 0, stream, scene, params, packs, latents, partials,
                           gradLatents, mLat, vLat, lrLat, bc1, bc2, sse, st, perWave, samples);
        markKernel("kTrain", stream);
    };
    auto pick = [&](auto fmt) {
        constexpr int FMT = decltype(fmt)::value;
        if (mode == kTrainLoss) go(kTrain<FMT, kTrainLoss>);
        else if (mode == kTrainGrad) go(kTrain<FMT, kTrainGrad>);
        else go(kTrain<FMT, kTrainStep>);
    };
    if (scene.format == GSR_SCENE_F32_SOA) pick(std::integral_constant<int, GSR_SCENE_F32_SOA>{});
    else pick(std::integral_constant<int, GSR_SCENE_F16>{});
    if (mode != kTrainLoss) {
        hipLaunchKernelGGL(kReduceGrad, dim3((kParamCount + 63) / 64), dim3(256), 0, stream, partials, blocks,
                           gradParams);
        markKernel("kReduceGrad", stream);
    }
}

void launchAdam(float* p, float* grad, float* m, float* v, uint32_t n, float lr, uint32_t step, hipStream_t stream) {
    const float bc1 = 1.0f - powf(0.9f, float(step)), bc2 = 1.0f - powf(0.999f, float(step));
    const uint32_t blocks = gridFor(n, 256, 4096);
    if (blocks == 0) return;
    hipLaunchKernelGGL(kAdam, dim3(blocks), dim3(256), 0, stream, p, grad, m, v, n, lr, bc1, bc2);
    markKernel("kAdam", stream);
}

void launchToBf16(const float* in, uint16_t* out, uint32_t n, hipStream_t stream) {
    const uint32_t blocks = gridFor(n, 256, 4096);
    if (blocks == 0) return;
    hipLaunchKernelGGL(kToBf16, dim3(blocks), dim3(256), 0, stream, in, out, n);
    markKernel("kToBf16", stream);
}

uint32_t shMomentBlocks(uint32_t count) { return std::min<uint32_t>(512u, (count + kMomentRows - 1) / kMomentRows); }

void launchShMoments(const SceneRef& scene, float* partials, hipStream_t stream) {
    const uint32_t blocks = shMomentBlocks(scene.count);
    if (blocks == 0) return;
    if (scene.format == GSR_SCENE_F32_SOA)
        hipLaunchKernelGGL(kShMoments<GSR_SCENE_F32_SOA>, dim3(blocks), dim3(256), 0, stream, scene, partials);
    else
        hipLaunchKernelGGL(kShMoments<GSR_SCENE_F16>, dim3(blocks), dim3(256), 0, stream, scene, partials);
    markKernel("kShMoments", stream);
}

void launchInitLatentsPca(const SceneRef& scene, const float* projMu, float* latents, hipStream_t stream) {
    const uint32_t blocks = (scene.count + 255) / 256;
    if (blocks == 0) return;
    if (scene.format == GSR_SCENE_F32_SOA)
        hipLaunchKernelGGL(kInitLatentsPca<GSR_SCENE_F32_SOA>, dim3(blocks), dim3(256), 0, stream, scene, projMu,
                           latents);
    else
        hipLaunchKernelGGL(kInitLatentsPca<GSR_SCENE_F16>, dim3(blocks), dim3(256), 0, stream, scene, projMu, latents);
    markKernel("kInitLatentsPca", stream);
}

void launchInitLatents(float* out, uint32_t n, uint32_t seed, float scale, hipStream_t stream) {
    const uint32_t blocks = gridFor(n, 256, 4096);
    if (blocks == 0) return;
    hipLaunchKernelGGL(kInitLatents, dim3(blocks), dim3(256), 0, stream, out, n, seed, scale);
    markKernel("kInitLatents", stream);
}

void launchMlpSamples(const SampleStream& st, uint32_t first, uint32_t n, uint32_t* gaussians, float* dirs,
                      hipStream_t stream) {
    const uint32_t blocks = (n + 255) / 256;
    if (blocks == 0) return;
    hipLaunchKernelGGL(kSamples, dim3(blocks), dim3(256), 0, stream, st, first, n, gaussians, dirs);
    markKernel("kSamples", stream);
}

void launchMlpFallback(const SceneRef& scene, const float* params, const uint2* packs, const uint16_t* latentsBf16,
                       float threshold, uint8_t* useMlp, float* err, hipStream_t stream) {
    const uint32_t blocks = gridFor(scene.count, 2 * kWaves, 2048);
    if (blocks == 0) return;
    if (scene.format == GSR_SCENE_F32_SOA)
        hipLaunchKernelGGL(kFallback<GSR_SCENE_F32_SOA>, dim3(blocks), dim3(kThreads), 0, stream, scene, params, packs,
                           latentsBf16, threshold, useMlp, err);
    else
        hipLaunchKernelGGL(kFallback<GSR_SCENE_F16>, dim3(blocks), dim3(kThreads), 0, stream, scene, params, packs,
                           latentsBf16, threshold, useMlp, err);
    markKernel("kFallback", stream);
}

void launchColourMlp(const SceneRef& scene, const DevCamera& cam, const float* params, const uint2* packs,
                     const uint16_t* latentsBf16, const uint8_t* useMlp, const uint32_t* tileCounts, float4* splats,
                     uint32_t renderSplat, uint32_t* packed, hipStream_t stream) {
    // Persistent: at most 4 waves per SIMD (each loads the weights once), 64 gaussians per round.
    const uint32_t blocks = gridFor(scene.count, 16 * kColourTiles * kWaves, 1024);
    if (blocks == 0) return;
    if (scene.format == GSR_SCENE_F32_SOA)
        hipLaunchKernelGGL(kColourMlp<GSR_SCENE_F32_SOA>, dim3(blocks), dim3(kThreads), 0, stream, scene, cam, params,
                           packs, latentsBf16, useMlp, tileCounts, splats, renderSplat, packed);
    else
        hipLaunchKernelGGL(kColourMlp<GSR_SCENE_F16>, dim3(blocks), dim3(kThreads), 0, stream, scene, cam, params,
                           packs, latentsBf16, useMlp, tileCounts, splats, renderSplat, packed);
    markKernel("kColourMlp", stream);
}

void launchMlpEvaluate(const float* params, const uint2* packs, const float* latents, const uint32_t* gaussians,
                       const float* dirs, uint32_t n, float* rgb, hipStream_t stream) {
    const uint32_t blocks = (n + 16 * kWaves - 1) / (16 * kWaves);
    if (blocks == 0) return;
    hipLaunchKernelGGL(kEvaluate, dim3(blocks), dim3(kThreads), 0, stream, params, packs, latents, gaussians, dirs, n,
                       rgb);
    markKernel("kEvaluate", stream);
}

}  // namespace gsr
