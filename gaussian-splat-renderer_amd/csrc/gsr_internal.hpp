// Internal types shared by the host orchestration (capi.cpp) and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/gsr.h"

namespace gsr {

// INRIA constants (reference splat/splat_math.h:25-40).
constexpr float kNearCullDepth = 0.2f;
constexpr float kLowPassFilter = 0.3f;
constexpr float kFrustumClampFactor = 1.3f;
constexpr float kMaxAlpha = 0.99f;
constexpr float kMinAlpha = 1.0f / 255.0f;
constexpr float kMinTransmittance = 1e-4f;
constexpr int kMaxSortWindow = 16;      // StopThePop: largest window held in registers
constexpr int kMaxSortWindowHeap = 64;  // StopThePop: largest window (LDS heaps beyond kMaxSortWindow)
constexpr int kTileSize = 16;
constexpr uint32_t kNoGaussian = 0xFFFFFFFFu;

// Preprocess: one thread per gaussian, 256 per block; the block also reduces its tile
// counts and cull counts into one BlockStats record (so no global atomics are needed).
constexpr int kPreprocessThreads = 256;
struct BlockStats {
    uint32_t tileSum;       // sum of the block's tile counts
    uint32_t nonZero;       // gaussians with a non-zero tile count
    uint32_t cull[5];       // per cull reason
    uint32_t pad;
};
static_assert(sizeof(BlockStats) == 32, "BlockStats is 32 bytes");

// Kernel-side camera (SplatCamera, splat_math.h:43-50). view[c*4+r], column-major.
// tilesX/Y: INRIA's 16-px tiles (getRect); binSize, binsX/Y: the grid the lists are built on
// (TileSettings::tileSize, 16 or 8 px; splat_tile_renderer.cpp:537-541).
struct DevCamera {
    float view[16];
    float pos[3];
    float tanX, tanY, focalX, focalY;
    uint32_t width, height, tilesX, tilesY;
    uint32_t binSize, binsX, binsY;
};

// Device view of a scene. Records wider than 16 B are stored as planes of 16-B quads:
// quad q of gaussian i is at plane[q * count + i].
struct SceneRef {
    uint32_t count;
    uint32_t K;              // coefficients per channel stored
    int32_t format;          // gsr_scene_format
    int32_t degree;          // bands evaluated
    const float4* posAlpha;  // F32
    const float4* scales;
    const float4* rotations;
    const float4* shPlanes;  // F32: shQuads planes; quad q = floats [4q, 4q + 4) of the K*3 record
    uint32_t shQuads;
    const uint4* geometry;   // F16: 2 planes (the two uint4 of the 32-B geometry record)
    const uint4* color;      // F16: colorQuads planes
    uint32_t colorQuads;     // uint4s per colour record
};

// Scene layout (scene.hip): record-per-gaussian arrays -> planes of 16-B quads.
void launchFloatsToPlanes(const float* src, uint32_t count, uint32_t floatsPer, uint32_t quads, float4* planes,
                          hipStream_t stream);
void launchQuadsToPlanes(const uint4* src, uint32_t count, uint32_t quads, uint4* planes, hipStream_t stream);

// Sorting: stable LSD radix sort, 8-bit digits, 64-bit keys (tile << 32 | depth bits),
// one sweep per digit over tiles of 2048 keys.
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortKeysPerLane = 8;                               // rounds of 64 keys per wave
constexpr int kSortTile = kSortThreads * kSortKeysPerLane;        // 2048 keys per tile
constexpr uint32_t kMaxSortPasses = 8;                            // 64-bit keys

// Device work areas of one sort. Reduce-then-scan (3 kernels per digit): per-block digit
// histograms and digit totals. One-sweep (1 kernel per digit after one histogram pass):
// global digit histograms (passes x 256), per-pass tile tickets, look-back status words
// (tiles x 256, never cleared: epoch-tagged) and a timeout counter (0 unless a look-back
// wait gave up).
constexpr uint32_t kMaxSortBlocks = 4096;                         // reduce-then-scan grid: one tile per block up to 8.4M keys
enum SortAlgorithm : int32_t { kSortAlgReduceThenScan = 0, kSortAlgOnesweep = 1 };
struct SortWork {
    int32_t algorithm;
    uint32_t* globalHist;  // one-sweep: passes x 256 digit counts; reduce-then-scan: 256 digit totals
    uint32_t* tickets;     // one-sweep
    uint64_t* status;      // one-sweep
    uint32_t* timeouts;    // one-sweep
    uint32_t* blockHist;   // reduce-then-scan: 2^bits x blocks counts, blocks <= kMaxSortBlocks
};
size_t sortStatusBytes(uint32_t maxCount);

// Per-kernel timing (profile_stages = 2). While a frame is recorded with marks set for the calling
// thread, every launch wrapper records an event on its stream after its kernel, with the
// kernel's name; kernel i took elapsed(events[i], events[i + 1]) (events[0] is recorded before
// the frame's first kernel).
struct KernelMarks {
    static constexpr int kMax = 96;
    hipEvent_t events[kMax + 1] = {};
    const char* names[kMax] = {};
    int count = 0;
};
extern thread_local KernelMarks* t_kernelMarks;
inline void markKernel(const char* name, hipStream_t stream) {
    KernelMarks* m = t_kernelMarks;
    if (m && m->count < KernelMarks::kMax) {
        (void)hipEventRecord(m->events[m->count + 1], stream);
        m->names[m->count++] = name;
    }
}

// Launch wrappers (each in its .hip file).
// Global mode: the preprocess also adds each block's BlockStats into the sums of its superblock of
// kSuperBlocks blocks, superStats[block / kSuperBlocks] (atomics; one 32-B record per superblock,
// so a workgroup's atomics meet one cache line), and zeroes the control block's accumulated
// counters; the duplicate pass then finds its block's first entry from those sums and the earlier
// blocks of its superblock (no frame-setup scan), and its workgroup 0 writes the control block;
// the ranges pass zeroes the sums for the next frame. superStats = nullptr: binning mode
// (kFrameSetup scans the block sums).
constexpr uint32_t kSuperBlocks = 64;
inline uint32_t superBlockCount(uint32_t count) {
    const uint32_t blocks = (count + kPreprocessThreads - 1) / kPreprocessThreads;
    return (blocks + kSuperBlocks - 1) / kSuperBlocks;
}
// spanInfo / spanPlanes (2 planes of uint4, count each): per gaussian with tiles, its coverage as
// the duplicate pass walks it (SpanRecord, gsr_device.hpp), written with the tile count; so is
// depthBits[i] (the splat's depth as bits, the list's depth key; nullptr: not written).
void launchPreprocess(const SceneRef& scene, const DevCamera& cam, uint32_t bounds, float4* splats,
                      uint32_t* tileCounts, BlockStats* blockStats, BlockStats* superStats, gsr_control* control,
                      uint32_t* spanInfo, uint4* spanPlanes, uint32_t* depthBits, uint32_t* rangesWords,
                      uint32_t rangeWordCount, uint32_t renderSplat, uint32_t* packed, float4* rayDepth,
                      const uint8_t* useMlp, hipStream_t stream);
// scanVisible = false: blockOffsets = exclusive scan of the blocks' tile sums (global mode);
// true: of their visible (non-zero count) gaussians (binning mode's compaction).
void launchFrameSetup(const BlockStats* blockStats, uint32_t numBlocks, uint32_t gaussianCount, uint32_t capacity,
                      bool scanVisible, uint32_t* blockOffsets, gsr_control* control, hipStream_t stream);
// The lists' grid: getRect's 16-px tiles and the bins of binSize (16 or 8) px the entries name.
struct BinGrid {
    uint32_t tilesX, tilesY, binSize, binsX, binsY;
};
// Global mode's duplicate: block b's first entry from the preprocess's superblock and block sums;
// the list is clamped to `capacity`; the last workgroup writes the control block.
struct DupFrame {
    const BlockStats* blockStats;
    const BlockStats* superStats;  // numSuper sums
    uint32_t numSuper, capacity;
};
void launchDuplicate(const float4* splats, const uint32_t* tileCounts, const DupFrame& frame, const uint32_t* spanInfo,
                     const uint4* spanPlanes, const uint32_t* depthBits, uint32_t count, uint32_t bounds,
                     const BinGrid& grid, uint64_t* keys, uint32_t* values, gsr_control* control, hipStream_t stream);
// The same list with 32-bit keys = tile id (the depth keys stay in the preprocess's depthBits).
void launchDuplicateTiles(const float4* splats, const uint32_t* tileCounts, const DupFrame& frame,
                          const uint32_t* spanInfo, const uint4* spanPlanes, uint32_t count, uint32_t bounds,
                          const BinGrid& grid, uint32_t* keys, uint32_t* values, gsr_control* control,
                          hipStream_t stream);
// Sorts by key bits [firstBit, keyBits) (stable). Returns which buffer (0 = keysA/valuesA,
// 1 = B) holds the sorted result. *epoch is the caller's status-epoch counter (one per status
// array), advanced once per pass.
int launchRadixSort(const SortWork& work, uint64_t* keysA, uint32_t* valuesA, uint64_t* keysB, uint32_t* valuesB,
                    uint32_t maxCount, const uint32_t* countDevice, uint32_t firstBit, uint32_t keyBits,
                    uint32_t* epoch, hipStream_t stream);
// Per-tile depth order (tilesort.hip): after a stable sort of the list by its tile ids, sorts
// each tile's range of gaussian indices by (depth bits, gaussian index) in place, the depth bits
// of gaussian g being depthBits[g]; the lists then equal those of the full 64-bit (tile, depth)
// sort. values must hold kListPad readable entries past the list (it reads whole rows of 64).
// valuesTmp: scratch of the list's size (ranges longer than 8192 entries). order: a permutation
// of the tile ids, the order in which the tiles are taken (scheduling only; the lists do not
// depend on it).
constexpr uint32_t kListPad = 8192;
void launchTileDepthSort(uint32_t* values, const uint32_t* depthBits, uint32_t* valuesTmp, const uint2* ranges,
                         const uint32_t* order, uint32_t tileCount, hipStream_t stream);
int launchRadixSort32(const SortWork& work, uint32_t* keysA, uint32_t* valuesA, uint32_t* keysB, uint32_t* valuesB,
                      uint32_t maxCount, const uint32_t* countDevice, uint32_t keyBits, hipStream_t stream);
void launchCopyPairs(const uint64_t* ks, const uint32_t* vs, uint64_t* kd, uint32_t* vd, const uint32_t* countDevice,
                     uint32_t maxCount, hipStream_t stream);
// zeroWords[0, zeroCount): words the ranges pass zeroes (global mode's superblock sums; none: 0).
void launchRanges(const uint64_t* keys, const gsr_control* control, uint32_t capacity, uint32_t tileCount,
                  uint2* ranges, uint32_t* zeroWords, uint32_t zeroCount, hipStream_t stream);
void launchRanges32(const uint32_t* keys, const gsr_control* control, uint32_t capacity, uint32_t tileCount,
                    uint2* ranges, uint32_t* zeroWords, uint32_t zeroCount, hipStream_t stream);

// Binning mode (the reference's default tile sort, splat_tile_renderer.h:8-26): compact the
// visible splats (depth bits, index) in index order, sort them by depth (32-bit keys), scan
// their tile counts in that order, duplicate with 32-bit tile keys, stable sort by tile.
void launchCompactVisible(const float4* splats, const uint32_t* tileCounts, const uint32_t* blockOffsets,
                          uint32_t count, uint32_t* depthKeys, uint32_t* depthValues, hipStream_t stream);
void launchGatherCounts(const uint32_t* tileCounts, const uint32_t* order, const gsr_control* control,
                        uint32_t maxVisible, uint32_t* sortedCounts, uint32_t* blockSums, hipStream_t stream);
// Exclusive scan of the ceil(visible / 256) block sums in place (one workgroup).
void launchScanVisibleBlocks(uint32_t* blockSums, const gsr_control* control, hipStream_t stream);
void launchDuplicateBinning(const float4* splats, const uint32_t* sortedCounts, const uint32_t* order,
                            const uint32_t* blockOffsets, const gsr_control* control, const uint32_t* spanInfo,
                            const uint4* spanPlanes, uint32_t maxVisible, uint32_t bounds, const BinGrid& grid,
                            uint32_t* keys, uint32_t* values, gsr_control* controlRW, hipStream_t stream);
// What the culled render of one frame's tiles reads and writes (16-px tiles, tilesX per row).
// sortValues != nullptr: the list is grouped by tile but each tile's range is not yet in depth
// order; the render's workgroup sorts its tile's range in place first (sortValues = values; the
// depth bits of gaussian g in depthBits[g]; sortScratch: a list-sized scratch for ranges over 4096
// entries), as the tile depth sort would (tilesort.hpp).
struct TileJob {
    const float4* splats;
    const uint32_t* packed;
    const uint32_t* values;
    const uint2* ranges;
    uint32_t width, height, tilesX, gaussianCount;
    void* out;
    uint32_t* tileStats;
    uint32_t* sortValues = nullptr;
    const uint32_t* depthBits = nullptr;
    uint32_t* sortScratch = nullptr;
};
// The per-tile depth sort fused into the culled render (TileJob::sortValues).
struct FusedSort {
    uint32_t* values;
    const uint32_t* depthBits;
    uint32_t* scratch;
};
// One persistent culled render over the tiles of `views` frames (jobs and tileStart on the
// device: tileStart[v] = the tiles of views 0..v-1, tileStart[views] = all); `counter`: one device
// word, zeroed by the launch itself.
void launchRenderViews(const TileJob* jobs, const uint32_t* tileStart, uint32_t views, uint32_t workgroups,
                       uint32_t* counter, gsr_target_format format, int renderSplat, hipStream_t stream);
// sortWindow > 0: StopThePop with that window (rayDepth: 3 float4 per gaussian), reference alpha.
// tileStats (culled loop only; null = none): 8 counters per tile, zeroed by the caller (the
// reference's TileStat, splat_tile_renderer.h:130-141).
constexpr int kTileStatCount = 8;
// fused != nullptr: the culled 16-px render (loop CULLED, one pixel per thread, fast alpha, no
// window) sorts each tile's range by depth itself (TileJob::sortValues).
void launchRender(const float4* splats, const uint32_t* packed, int renderSplat, const uint32_t* values,
                  const uint2* ranges, const DevCamera& cam, uint32_t gaussianCount, gsr_target_format format,
                  int pixelsPerThread, bool exactAlpha, int loop, int sortWindow, const float4* rayDepth, void* out,
                  uint32_t* tileStats, int batch, hipStream_t stream, const FusedSort* fused = nullptr);

// Learned colour (mlp.hip): a 16-value latent per gaussian and one shared MLP. Parameters:
// mlpParamCount() floats; packs: mlpPackedTiles() x 64 uint2 (the bf16 A-operand tiles, rebuilt
// by launchMlpPack after every parameter change).
int mlpParamCount();
int mlpPackedTiles();
void launchMlpPack(const float* params, uint2* packs, hipStream_t stream);
// A training sample stream: sample s is gaussian (mulA s + addB) mod count (mulA coprime to count:
// samples 0 .. count - 1 are distinct gaussians) and a direction hashed from (seed, s).
struct SampleStream {
    uint32_t seed, mulA, addB, count;
};
SampleStream sampleStream(uint32_t seed, uint32_t count);
enum MlpTrainMode : int { kTrainLoss = 0, kTrainGrad = 1, kTrainStep = 2 };
// The workgroups of a training pass over `samples` samples (the rows of its `partials`, at most 512).
uint32_t mlpTrainBlocks(uint32_t samples);
// Samples 0 .. samples - 1 (samples % 16 == 0) of stream st: the squared-error sum into *sse and,
// unless kTrainLoss, the weight gradient of the mean squared error into gradParams (written, via
// per-workgroup rows in partials: mlpTrainBlocks(samples) x mlpParamCount() floats). The latent
// gradient: kTrainGrad adds it into gradLatents; kTrainStep (samples <= count: distinct gaussians)
// applies Adam step `step` (lrLat, moments mLat / vLat) to the touched latents in place.
void launchMlpTrain(const SceneRef& scene, const float* params, const uint2* packs, float* latents, float* gradParams,
                    float* partials, float* gradLatents, float* mLat, float* vLat, float lrLat, uint32_t step,
                    float* sse, const SampleStream& st, uint32_t samples, int mode, hipStream_t stream);
// One Adam step (step >= 1) on n values; clears the gradients.
void launchAdam(float* p, float* grad, float* m, float* v, uint32_t n, float lr, uint32_t step, hipStream_t stream);
void launchToBf16(const float* in, uint16_t* out, uint32_t n, hipStream_t stream);
void launchInitLatents(float* out, uint32_t n, uint32_t seed, float scale, hipStream_t stream);
// Principal-component latents: launchShMoments writes shMomentBlocks(count) rows of D SH sums and
// D (D + 1) / 2 product sums (D = 3 (degree + 1)^2 <= 48); launchInitLatentsPca sets latent r of
// every gaussian to projMu[r * 48 ..] . (x - mu), mu = projMu[16 * 48 ..] (48 floats each).
uint32_t shMomentBlocks(uint32_t count);
void launchShMoments(const SceneRef& scene, float* partials, hipStream_t stream);
void launchInitLatentsPca(const SceneRef& scene, const float* projMu, float* latents, hipStream_t stream);
void launchMlpSamples(const SampleStream& st, uint32_t first, uint32_t n, uint32_t* gaussians, float* dirs,
                      hipStream_t stream);
void launchMlpFallback(const SceneRef& scene, const float* params, const uint2* packs, const uint16_t* latentsBf16,
                       float threshold, uint8_t* useMlp, float* err, hipStream_t stream);
// After the preprocess (run with the same useMlp): the MLP colour of the gaussians with useMlp set
// and a non-zero tile count, into the splats (and the packed records in the p32 / p24 modes).
void launchColourMlp(const SceneRef& scene, const DevCamera& cam, const float* params, const uint2* packs,
                     const uint16_t* latentsBf16, const uint8_t* useMlp, const uint32_t* tileCounts, float4* splats,
                     uint32_t renderSplat, uint32_t* packed, hipStream_t stream);
void launchMlpEvaluate(const float* params, const uint2* packs, const float* latents, const uint32_t* gaussians,
                       const float* dirs, uint32_t n, float* rgb, hipStream_t stream);

// Differentiable rendering (backward.hip). The training forward (full records, RGBA32F out, exact
// alpha) keeps per pixel (T bits, entries up to the last contributor); the render backward adds
// each gaussian's screen-space gradient into grad2D (grad2DComponents() planes of N floats:
// centre x, y, conic A, B, C, opacity, colour r, g, b; cleared by the caller); the preprocess
// backward turns them into the gradients of the mean, log-scales, quaternion, opacity logit and
// SH coefficients (f32 scenes; each output may be null).
int grad2DComponents();
void launchRenderTrain(const float4* splats, const uint32_t* values, const uint2* ranges, const DevCamera& cam,
                       float4* out, uint2* state, hipStream_t stream);
void launchRenderBackward(const float4* splats, const uint32_t* values, const uint2* ranges, const DevCamera& cam,
                          const float4* dOut, const uint2* state, float* grad2D, uint32_t count, hipStream_t stream);
void launchPreprocessBackward(const SceneRef& scene, const DevCamera& cam, const float4* splats, const float* grad2D,
                              float* dMean, float* dLogScale, float* dQuat, float* dLogit, float* dSh,
                              hipStream_t stream);

// Stand-alone scan (GpuScan): reduce, block-sum scan, downsweep.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanBlock = kScanThreads * kScanItems;  // 2048
void launchExclusiveScan(const uint32_t* in, uint32_t* out, uint32_t count, uint32_t* blockSums, uint32_t* total,
                         hipStream_t stream);

}  // namespace gsr
