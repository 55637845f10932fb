// Per-tile sorts of list ranges by (depth bits, gaussian index), shared by the tile depth sort
// kernel (tilesort.hip) and the culled render, which sorts its own tile when the frame fuses the
// two: wave forms (a range of up to 64 E words in one wave's registers, a bitonic network of
// register, DPP and lane-shuffle stages), workgroup forms (4 waves, cross-wave stages through LDS)
// and the long form (LDS chunks of 2048, then merge rounds through a scratch list). Every form
// reads whole rows of 64 list values (the value buffers are padded past the list) and writes the
// sorted values in place.
#pragma once

#include "gsr_internal.hpp"

namespace gsr {

namespace {  // each translation unit compiles its own copy for its kernels' register budget


constexpr int kTsThreads = 256;
constexpr int kTsWaves = kTsThreads / 64;
constexpr uint32_t kWaveChunk = 2048;  // entries one wave sorts in registers (E = 32)
constexpr uint64_t kPad = ~uint64_t(0);

// (depth bits, gaussian index) of list value g.
__device__ __forceinline__ uint64_t orderWord(const uint32_t* __restrict__ depthBits, uint32_t g) {
    return (uint64_t(depthBits[g]) << 32) | g;
}

__device__ __forceinline__ uint64_t shflXor64(uint64_t v, uint32_t mask) {
    const uint32_t lo = __shfl_xor(uint32_t(v), int(mask), 64);
    const uint32_t hi = __shfl_xor(uint32_t(v >> 32), int(mask), 64);
    return (uint64_t(hi) << 32) | lo;
}

// The word of lane ^ M for a constant M < 16: DPP moves within a row of 16 lanes, VALU only
// (quad_perm for xor 1, 2, 3; row_half_mirror = xor 7; row_ror:8 = xor 8; row_mirror = xor 15;
// xor 4 = the quad mirror of the half-row mirror).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, 0xF, 0xF, false));
}
template <uint32_t M>
__device__ __forceinline__ uint32_t rowXor(uint32_t v) {
    if constexpr (M == 1) return dpp<0xB1>(v);
    else if constexpr (M == 2) return dpp<0x4E>(v);
    else if constexpr (M == 3) return dpp<0x1B>(v);
    else if constexpr (M == 4) return dpp<0x1B>(dpp<0x141>(v));
    else if constexpr (M == 7) return dpp<0x141>(v);
    else if constexpr (M == 8) return dpp<0x128>(v);
    else {
        static_assert(M == 15, "xor masks with a DPP form");
        return dpp<0x140>(v);
    }
}
template <uint32_t M>
__device__ __forceinline__ uint64_t rowXor64(uint64_t v) {
    return (uint64_t(rowXor<M>(uint32_t(v >> 32))) << 32) | rowXor<M>(uint32_t(v));
}

// Compare-exchange of two registers: minimum to the first.
__device__ __forceinline__ void exchange(uint64_t& a, uint64_t& b) {
    const uint64_t x = a, y = b;
    a = __builtin_elementwise_min(x, y);
    b = __builtin_elementwise_max(x, y);
}

// The lower position keeps the minimum of the pair, the upper one the maximum.
__device__ __forceinline__ uint64_t keep(bool lower, uint64_t mine, uint64_t partner) {
    return (lower == (mine < partner)) ? mine : partner;
}

// In-lane stages (partners r and r ^ M of the same lane; M < E).
template <int E, int M>
__device__ __forceinline__ void registerStage(uint64_t (&w)[E]) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int p = r ^ M;
        if (p > r) exchange(w[r], w[p]);
    }
}
template <int E, int J>
__device__ __forceinline__ void registerHalves(uint64_t (&w)[E]) {  // stages J, J / 2, ..., 1
    if constexpr (J >= 1) {
        registerStage<E, J>(w);
        registerHalves<E, J / 2>(w);
    }
}
// Merge levels k = 2 .. E: every lane sorts its own E registers.
template <int E, int K = 2>
__device__ __forceinline__ void registerLevels(uint64_t (&w)[E]) {
    if constexpr (K <= E) {
        registerStage<E, K - 1>(w);  // flip
        registerHalves<E, K / 4>(w);
        registerLevels<E, K * 2>(w);
    }
}

// The 32-word sort keeps lane shuffles for every lane stage: in the DPP form it spills within
// the 96 VGPRs of 5 workgroups per CU, and at 4 workgroups per CU (128 VGPRs) the kernel took
// 71.9 us per headline frame against 58.5 us with this split (64.0 us with shuffles only).
template <int E>
constexpr bool kDppStages = E <= 16;

// Lane stages (partners m lanes apart, m = j / E): same register. M > 0: the constant lane
// distance, exchanged by DPP; M == 0: the run-time distance m, by lane shuffles.
template <int E, uint32_t M>
__device__ __forceinline__ void laneHalfBy(uint64_t (&w)[E], bool lower, uint32_t m) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
        uint64_t y;
        if constexpr (M == 0) y = shflXor64(w[r], m);
        else y = rowXor64<M>(w[r]);
        w[r] = keep(lower, w[r], y);
    }
}
template <int E>
__device__ __forceinline__ void laneHalf(uint64_t (&w)[E], uint32_t lane, uint32_t m) {
    const bool lower = (lane & m) == 0u;
    if constexpr (!kDppStages<E>) {
        laneHalfBy<E, 0>(w, lower, m);
        return;
    }
    switch (m) {  // uniform per wave
    case 1: laneHalfBy<E, 1>(w, lower, m); break;
    case 2: laneHalfBy<E, 2>(w, lower, m); break;
    case 4: laneHalfBy<E, 4>(w, lower, m); break;
    case 8: laneHalfBy<E, 8>(w, lower, m); break;
    default: laneHalfBy<E, 0>(w, lower, m); break;  // 16, 32
    }
}
// Flip of a level k > E: position lane * E + r pairs with (lane ^ m) * E + (E - 1 - r),
// m = k / E - 1; the lower position is the one whose lane has bit k / (2 E) clear.
template <int E, uint32_t M>
__device__ __forceinline__ void laneFlipBy(uint64_t (&w)[E], bool lower, uint32_t m) {
    auto partner = [&](uint64_t v) {
        if constexpr (M == 0) return shflXor64(v, m);
        else return rowXor64<M>(v);
    };
    if constexpr (E == 1) {
        w[0] = keep(lower, w[0], partner(w[0]));
    } else {
#pragma unroll
        for (int r = 0; r < E / 2; ++r) {
            const uint64_t ya = partner(w[E - 1 - r]);
            const uint64_t yb = partner(w[r]);
            w[r] = keep(lower, w[r], ya);
            w[E - 1 - r] = keep(lower, w[E - 1 - r], yb);
        }
    }
}
template <int E>
__device__ __forceinline__ void laneFlip(uint64_t (&w)[E], uint32_t lane, uint32_t m, uint32_t half) {
    const bool lower = (lane & half) == 0u;
    if constexpr (!kDppStages<E>) {
        laneFlipBy<E, 0>(w, lower, m);
        return;
    }
    switch (m) {  // uniform per wave
    case 1: laneFlipBy<E, 1>(w, lower, m); break;
    case 3: laneFlipBy<E, 3>(w, lower, m); break;
    case 7: laneFlipBy<E, 7>(w, lower, m); break;
    case 15: laneFlipBy<E, 15>(w, lower, m); break;
    default: laneFlipBy<E, 0>(w, lower, m); break;  // 31, 63
    }
}

// Ascending bitonic sort of the wave's 64 E words (positions lane * E + r).
template <int E>
__device__ __forceinline__ void waveSort(uint64_t (&w)[E], uint32_t lane) {
    constexpr uint32_t N = 64u * E;
    registerLevels<E>(w);
#pragma unroll 1
    for (uint32_t k = 2u * E; k <= N; k <<= 1) {
        laneFlip<E>(w, lane, k / E - 1u, k / (2u * E));
#pragma unroll 1
        for (uint32_t j = k >> 2; j >= uint32_t(E); j >>= 1) laneHalf<E>(w, lane, j / E);
        registerHalves<E, E / 2>(w);
    }
}

// Words r of a lane: list entries first + r * 64 + lane (rows of 64, read whole: the value buffers
// are padded), their depth bits looked up, kPad past the n entries.
template <int E>
__device__ __forceinline__ void loadWords(uint64_t (&w)[E], const uint32_t* values,
                                          const uint32_t* __restrict__ depthBits, uint32_t first, uint32_t n,
                                          uint32_t lane) {
    const uint32_t* row = values + first + lane;
#pragma unroll
    for (int r = 0; r < E; ++r) w[r] = row[r * 64];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const bool valid = uint32_t(r) * 64u + lane < n;
        const uint32_t g = valid ? uint32_t(w[r]) : 0u;
        const uint64_t word = orderWord(depthBits, g);
        w[r] = valid ? word : kPad;
    }
}

// The wave sorts entries [at, at + n) in place (n <= 64 E).
template <int E>
__device__ __forceinline__ void waveSortRange(uint32_t* values, const uint32_t* __restrict__ depthBits, uint32_t at,
                                              uint32_t n, uint32_t lane) {
    uint64_t w[E];
    loadWords<E>(w, values, depthBits, at, n, lane);
    waveSort<E>(w, lane);
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const uint32_t e = lane * E + uint32_t(r);
        if (e < n) values[at + e] = uint32_t(w[r]);
    }
}

__device__ __forceinline__ void waveSortAny(uint32_t* values, const uint32_t* __restrict__ depthBits, uint32_t at,
                                            uint32_t n, uint32_t lane) {
    if (n <= 64u) waveSortRange<1>(values, depthBits, at, n, lane);  // uniform per wave
    else if (n <= 128u) waveSortRange<2>(values, depthBits, at, n, lane);
    else if (n <= 256u) waveSortRange<4>(values, depthBits, at, n, lane);
    else if (n <= 512u) waveSortRange<8>(values, depthBits, at, n, lane);
    else if (n <= 1024u) waveSortRange<16>(values, depthBits, at, n, lane);
    else waveSortRange<32>(values, depthBits, at, n, lane);
}

// ---- long ranges: the whole workgroup ----

constexpr uint32_t kGroupChunk = 4u * kWaveChunk;  // ranges up to 8192: the register form below
constexpr uint32_t kWaveAlone = kWaveChunk;  // ranges a single wave sorts (longer: the whole workgroup)

// Storage slot in sW of chunk position q = lane * E + r (slot r * 64 + lane: a wave's accesses
// to one register index hit 64 consecutive words).
__device__ __forceinline__ uint32_t chunkSlot(uint32_t lane, uint32_t r) { return r * 64u + lane; }

constexpr uint32_t kLdsChunk = 2048;  // entries sorted in LDS at once (16 KB)

// Cross-wave compare-exchange of two chunks of 64 E words held in registers by a lower and an
// upper wave: position q of the lower pairs with position q of the upper, or with 64 E - 1 - q
// when `flip`. The lower keeps the minimum, the upper the maximum, exchanged in place through s:
// the upper writes its words, the lower reads each partner, keeps the minimum and writes the
// maximum into the same slot, the upper reads it back. `active` (the same for both waves of the
// pair) false: the pair does nothing. Every thread of the workgroup calls it (barriers).
template <int E>
__device__ __forceinline__ void exchangeStep(uint64_t (&w)[E], uint32_t lane, bool upper, bool active, bool flip,
                                             uint64_t* s) {
    if (active && upper) {
#pragma unroll
        for (uint32_t r = 0; r < uint32_t(E); ++r) s[chunkSlot(lane, r)] = w[r];
    }
    __syncthreads();
    if (active && !upper) {
#pragma unroll
        for (uint32_t r = 0; r < uint32_t(E); ++r) {
            const uint32_t slot = flip ? chunkSlot(63u - lane, uint32_t(E) - 1u - r) : chunkSlot(lane, r);
            const uint64_t y = s[slot];
            s[slot] = __builtin_elementwise_max(w[r], y);
            w[r] = __builtin_elementwise_min(w[r], y);
        }
    }
    __syncthreads();
    if (active && upper) {
#pragma unroll
        for (uint32_t r = 0; r < uint32_t(E); ++r) w[r] = s[chunkSlot(lane, r)];
    }
    __syncthreads();
}

// Two disjoint pairs of waves exchanged at once (pair 0 in sW[0, 64 E), pair 1 after it), or one
// after the other where two chunks do not fit in the kLdsChunk words of sW.
template <int E>
__device__ __forceinline__ void exchangePairs(uint64_t (&w)[E], uint32_t lane, uint32_t pair, bool upper,
                                              bool active, bool flip, uint64_t* sW) {
    constexpr uint32_t C = 64u * E;
    if constexpr (2u * C <= kLdsChunk) {
        exchangeStep<E>(w, lane, upper, active, flip, sW + pair * C);
    } else {
        exchangeStep<E>(w, lane, upper, active && pair == 0u, flip, sW);
        exchangeStep<E>(w, lane, upper, active && pair == 1u, flip, sW);
    }
}

// The stages of a merge level below the chunk size (j = 32 E .. 1), within each wave's chunk.
template <int E>
__device__ __forceinline__ void mergeChunk(uint64_t (&w)[E], uint32_t lane) {
#pragma unroll 1
    for (uint32_t j = 32u * E; j >= uint32_t(E); j >>= 1) laneHalf<E>(w, lane, j / E);
    registerHalves<E, E / 2>(w);
}

// A range of n <= 256 E entries sorted by the 4 waves of the workgroup: positions
// v * 64 E + lane * E + r (wave v), each wave's chunk of 64 E sorted in registers, then the
// bitonic merge levels 128 E and 256 E: the stages whose partners are in another wave through
// sW (the two pairs of a stage at once), the rest in registers (the 16-word form: DPP for lane
// distances below 16). Padding words sort last. The waves holding entries are a prefix; a wave
// past the range (all padding) is only ever the upper wave of a pair, where max(x, padding) leaves
// it as it is, so it skips its sorts and its pairs skip their exchanges.
template <int E>
__device__ __attribute__((noinline)) void groupSortRegisters(uint32_t* values, const uint32_t* __restrict__ depthBits,
                                                             uint2 range, uint32_t wave, uint32_t lane, uint64_t* sW) {
    constexpr uint32_t C = 64u * E;  // a wave's chunk
    const uint32_t n = range.y - range.x;
    const bool mine = n > wave * C;  // this wave's chunk holds entries
    uint64_t w[E];
    if (mine) {
        loadWords<E>(w, values, depthBits, range.x + wave * C, n - wave * C, lane);
        waveSort<E>(w, lane);
    } else {
#pragma unroll
        for (int r = 0; r < E; ++r) w[r] = kPad;
    }
    // Level 2C: chunks (0, 1) and (2, 3), flip.
    const uint32_t pairA = wave >> 1;
    const bool upperA = (wave & 1u) != 0u, activeA = n > (2u * pairA + 1u) * C;
    exchangePairs<E>(w, lane, pairA, upperA, activeA, true, sW);
    if (mine) mergeChunk<E>(w, lane);
    // Level 4C: the flip pairs chunk v with chunk 3 - v (position q with C - 1 - q), then the
    // stage C apart pairs chunk v with chunk v ^ 1 (same position).
    const uint32_t pairB = (wave == 1u || wave == 2u) ? 1u : 0u;
    exchangePairs<E>(w, lane, pairB, wave >= 2u, n > (pairB == 0u ? 3u : 2u) * C, true, sW);
    exchangePairs<E>(w, lane, pairA, upperA, activeA, false, sW);
    if (mine) mergeChunk<E>(w, lane);
#pragma unroll
    for (uint32_t r = 0; r < uint32_t(E); ++r) {
        const uint32_t e = wave * C + lane * E + r;
        if (e < n) values[range.x + e] = uint32_t(w[r]);
    }
}




// Ascending bitonic sort of s[0, n2) (n2 a power of two) by the workgroup.
__device__ __forceinline__ void bitonicSortLds(uint64_t* s, uint32_t n2) {
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2 / 2; i += kTsThreads) {
                const uint32_t a = ((i & ~(j - 1u)) << 1) | (i & (j - 1u));
                const uint32_t b = a + j;
                const bool up = (a & k) == 0u;
                const uint64_t x = s[a], y = s[b];
                if ((x > y) == up) {
                    s[a] = y;
                    s[b] = x;
                }
            }
            __syncthreads();
        }
    }
}

// Number of elements of run A (length la) among the first d outputs of merge(A, B).
__device__ __forceinline__ uint32_t coRank(const uint32_t* __restrict__ depthBits, const uint32_t* va, uint32_t la,
                                           const uint32_t* vb, uint32_t lb, uint32_t d) {
    uint32_t lo = d > lb ? d - lb : 0u, hi = d < la ? d : la;
    while (lo < hi) {
        const uint32_t i = (lo + hi) >> 1;
        if (orderWord(depthBits, va[i]) < orderWord(depthBits, vb[d - i - 1])) lo = i + 1;
        else hi = i;
    }
    return lo;
}

__device__ void groupSortLong(uint32_t* values, const uint32_t* __restrict__ depthBits, uint32_t* valuesTmp,
                              uint2 range, uint64_t* sW) {
    const uint32_t n = range.y - range.x;

    // 1. Chunks of up to 2048 entries, sorted in LDS, in place.
    for (uint32_t c0 = 0; c0 < n; c0 += kLdsChunk) {
        const uint32_t m = n - c0 < kLdsChunk ? n - c0 : kLdsChunk;
        uint32_t n2 = 2;
        while (n2 < m) n2 <<= 1;
        const uint32_t base = range.x + c0;
        for (uint32_t i = threadIdx.x; i < n2; i += kTsThreads)
            sW[i] = i < m ? orderWord(depthBits, values[base + i]) : kPad;
        __syncthreads();
        bitonicSortLds(sW, n2);
        for (uint32_t i = threadIdx.x; i < m; i += kTsThreads) values[base + i] = uint32_t(sW[i]);
        __syncthreads();
    }

    // 2. Merge rounds: runs of width w -> 2w, alternating between the list and the scratch.
    uint32_t* vs = values + range.x;
    uint32_t* vd = valuesTmp + range.x;
    for (uint32_t w = kLdsChunk; w < n; w <<= 1) {
        for (uint32_t lo = 0; lo < n; lo += 2 * w) {
            const uint32_t mid = lo + w < n ? lo + w : n;
            const uint32_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            const uint32_t la = mid - lo, lb = hi - mid, total = hi - lo;
            const uint32_t per = (total + kTsThreads - 1) / kTsThreads;
            const uint32_t d0 = threadIdx.x * per < total ? threadIdx.x * per : total;
            const uint32_t d1 = d0 + per < total ? d0 + per : total;
            uint32_t ia = coRank(depthBits, vs + lo, la, vs + mid, lb, d0);
            uint32_t ib = d0 - ia;
            for (uint32_t d = d0; d < d1; ++d) {
                bool takeA;
                if (ia >= la) takeA = false;
                else if (ib >= lb) takeA = true;
                else takeA = orderWord(depthBits, vs[lo + ia]) < orderWord(depthBits, vs[mid + ib]);
                if (takeA) {
                    vd[lo + d] = vs[lo + ia];
                    ++ia;
                } else {
                    vd[lo + d] = vs[mid + ib];
                    ++ib;
                }
            }
        }
        __syncthreads();  // the round's output is the next round's input (same workgroup)
        uint32_t* tv = vs;
        vs = vd;
        vd = tv;
    }
    if (vs != values + range.x) {  // the result is in the scratch: copy it back
        for (uint32_t i = threadIdx.x; i < n; i += kTsThreads) values[range.x + i] = vs[i];
    }
    __syncthreads();  // before the next long range reuses the LDS
}

}  // namespace

}  // namespace gsr
