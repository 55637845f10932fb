// Host side of the splat-tiles path: the C ABI of include/gsr.h. Replaces the Vulkan
// SplatTileRenderer host code (reference render/splat_tile_renderer.cpp:126-954): it owns
// the per-gaussian and list buffers, records the stages of a frame on one HIP stream and
// provides the blocking test readbacks. No count is read back during a frame: V, M and
// entries stay in the device control block and every kernel reads them there.
#include "gsr_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

using namespace gsr;

namespace {

thread_local std::string g_lastError;

gsr_status fail(gsr_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_lastError = buf;
    return s;
}

}  // namespace

namespace gsr {
thread_local KernelMarks* t_kernelMarks = nullptr;

gsr_status setError(gsr_status s, const char* message) {
    g_lastError = message;
    return s;
}
}  // namespace gsr

namespace {

gsr_status hipFail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? GSR_ERROR_OUT_OF_MEMORY : GSR_ERROR_HIP, "%s: %s", what,
                hipGetErrorString(e));
}

#define GSR_HIP(call)                                          \
    do {                                                       \
        const hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hipFail(e_, #call);       \
    } while (0)

// A device allocation that frees itself.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t allocate(size_t n) {
        release();
        const size_t b = std::max<size_t>(n, 256);
        hipError_t e = hipMalloc(&ptr, b);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        bytes = b;
        return hipSuccess;
    }
    // Grows (contents not kept).
    hipError_t ensure(size_t n) { return bytes >= n && ptr ? hipSuccess : allocate(n); }
    // Stream-ordered allocation and release (the adaptive list's buffers): usable by work
    // enqueued after the allocation on `stream`, freed after the work enqueued before the release.
    hipError_t allocateAsync(size_t n, hipStream_t stream) {
        release();
        const size_t b = std::max<size_t>(n, 256);
        hipError_t e = hipMallocAsync(&ptr, b, stream);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        bytes = b;
        return hipSuccess;
    }
    hipError_t releaseAsync(hipStream_t stream) {
        hipError_t e = ptr ? hipFreeAsync(ptr, stream) : hipSuccess;
        ptr = nullptr;
        bytes = 0;
        return e;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

uint32_t bitsFor(uint32_t v) {  // bits needed to hold values 0..v
    uint32_t b = 0;
    while (b < 32 && (uint64_t(1) << b) <= v) ++b;
    return b;
}

}  // namespace

// --- Fixed scratch of the allocation-free sort (gsr_sort_pairs_noalloc) ---------------------------
// For kNoAllocSortMaxKeys keys: the ping-pong copy (8-B keys, then 4-B values) and the
// reduce-then-scan work area (digit totals and per-block digit histograms), in one file-scope
// device array. Every GPU the library runs on holds its own copy until the process ends. In host
// code the array's name is a host-side stand-in, not a device address: host code uses only the
// address hipGetSymbolAddress gives for the current device (resolved once per device at
// gsr_context_create). Nothing in it must start at zero: the histogram pass writes the digit
// counts of every block that holds keys and the scan every digit total before they are read.
namespace {
constexpr uint32_t kNoAllocSortMaxKeys = 1u << 20;
constexpr size_t kNoAllocValuesOffset = size_t(kNoAllocSortMaxKeys) * 8;
constexpr size_t kNoAllocWorkOffset = size_t(kNoAllocSortMaxKeys) * 12;
constexpr size_t kNoAllocSortBlocks = (kNoAllocSortMaxKeys + kSortTile - 1) / kSortTile;  // the grid for that count
constexpr size_t kNoAllocWorkWords =
    size_t(kMaxSortPasses) * kRadix + kMaxSortPasses + 1 + size_t(kRadix) * kNoAllocSortBlocks;
constexpr size_t kNoAllocScratchBytes = kNoAllocWorkOffset + kNoAllocWorkWords * 4;
}  // namespace
static __device__ __align__(256) uint64_t scratch[(kNoAllocScratchBytes + 7) / 8];

namespace {
std::mutex g_scratchMutex;
std::vector<void*> g_scratchByDevice;  // device ordinal -> that device's address of `scratch`

// With `device` current: its address of `scratch`, looked up once per device.
hipError_t resolveScratch(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratchByDevice.size() <= size_t(device)) g_scratchByDevice.resize(size_t(device) + 1, nullptr);
    if (!g_scratchByDevice[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchByDevice[device] = p;
    }
    *out = g_scratchByDevice[device];
    return hipSuccess;
}
}  // namespace

struct gsr_context {
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffer work;        // stand-alone sort/scan work area
    void* noAllocScratch = nullptr;  // this device's copy of `scratch` (gsr_sort_pairs_noalloc)
    DeviceBuffer sortStatus;  // stand-alone sort: look-back status words (epoch-tagged, zeroed once)
    uint32_t sortEpoch = 0;
};

struct gsr_buffer {
    gsr_context* ctx = nullptr;
    DeviceBuffer mem;
};

struct gsr_scene {
    gsr_context* ctx = nullptr;
    uint32_t count = 0;
    int32_t shDegree = 0;
    uint32_t K = 1;
    gsr_scene_format format = GSR_SCENE_F32_SOA;
    DeviceBuffer posAlpha, scales, rotations, sh, geometry, color;  // owned copies
    SceneRef ref{};
};

// Learned colour: the MLP's parameters and the per-gaussian latents (fp32 for training, bf16 for
// rendering), each with Adam moments and gradients, the packed bf16 weight tiles and the fallback
// marks.
struct gsr_colour_model {
    gsr_context* ctx = nullptr;
    const gsr_scene* scene = nullptr;
    uint32_t count = 0;
    uint32_t seed = 0;
    DeviceBuffer params, gradParams, mParams, vParams, packs;
    DeviceBuffer latents, gradLatents, mLatents, vLatents, latentsBf16;
    DeviceBuffer useMlp, errors, sse;
    DeviceBuffer partials;    // the training pass's per-workgroup weight-gradient rows (sized on use)
    uint32_t step = 0;        // Adam steps since creation or the last set
    uint32_t streamIndex = 0; // training sample streams used
    uint32_t mlpCount = 0;    // gaussians marked by the last fallback pass
};

struct gsr_renderer {
    gsr_context* ctx = nullptr;
    const gsr_scene* scene = nullptr;
    gsr_settings settings{};
    uint32_t capacity = 0;
    int32_t degree = 0;
    // Per gaussian (+ the packed render splats in the p32 / p24 modes).
    DeviceBuffer splats, tileCounts, blockOffsets, blockStats, control, packed;
    // Per gaussian with tiles: its coverage as the duplicate pass walks it (4-B info + 2 planes of
    // 16 B: the first 8 row spans), written by the preprocess.
    DeviceBuffer spanInfo, spanPlanes;
    DeviceBuffer rayDepth;  // StopThePop: 3 float4 per gaussian
    // Differentiable renderers: per pixel (T bits, last contributor) of the last frame, sized for
    // the largest camera seen; per gaussian the screen-space gradient planes of gsr_backward.
    DeviceBuffer pixelState, grad2D;
    size_t pixelStatePixels = 0;
    // The list (ping-pong; keys are 32-bit tile ids in binning mode and in global mode with the
    // per-tile depth sort, 64-bit (tile, depth) for the full-key and one-sweep sorts) and the
    // sort's work areas. Global mode with the per-tile depth sort: per gaussian the depth bits the
    // duplicate pass writes for the tile sort.
    DeviceBuffer keysA, keysB, valuesA, valuesB, sortSmall, sortStatus, depthBits;
    // Global mode: the preprocess blocks' sums per superblock (superBlockCount BlockStats, zero
    // between frames: the ranges pass clears them).
    DeviceBuffer superStats;
    // Binning mode, per gaussian: compacted depth keys / indices (ping-pong), depth-ordered
    // tile counts and their per-256 sums.
    DeviceBuffer depthKeysA, depthKeysB, depthValuesA, depthValuesB, sortedCounts, visibleSums;
    bool binning = false;
    // Global mode sorted by tile id (reduce-then-scan, 2 passes of the 32-bit keys), then each
    // tile's range by depth (the tile depth sort); otherwise global mode sorts whole 64-bit keys.
    bool tileSort = false;
    bool tileKeys32() const { return binning || tileSort; }
    uint32_t sortEpoch = 0;
    // Adaptive list (splat_tile_renderer.cpp:444-486) and the per-frame statistics read back
    // without stalling: each frame's control block is copied to pinned host memory and an event
    // marks it; the next frames collect the completed copies.
    uint32_t maxCapacity = 0, listResizes = 0;
    uint64_t frameIndex = 0;
    static constexpr int kStatSlots = 4;
    gsr_control* statHost = nullptr;  // kStatSlots pinned control blocks
    hipEvent_t statEvent[kStatSlots] = {};
    uint64_t statFrame[kStatSlots] = {};
    bool statPending[kStatSlots] = {};
    gsr_control lastStats{};
    uint64_t lastStatsFrame = 0, lastAdaptedFrame = ~uint64_t(0);
    bool lastStatsValid = false;
    uint32_t shrinkFrames = 0, shrinkWindowMax = 0;
    // Per tile (sized for the largest camera seen).
    DeviceBuffer ranges;
    uint32_t rangesTiles = 0;
    // The tile depth sort's tile order, one table per tile grid rendered so far: centre-out (row
    // distance from the middle row, then column distance from the middle column).
    struct TileOrder {
        uint32_t tilesX = 0, tilesY = 0;
        DeviceBuffer table;
    };
    std::vector<std::unique_ptr<TileOrder>> tileOrders;
    const TileOrder* findOrder(uint32_t tx, uint32_t ty) const {
        for (const auto& o : tileOrders)
            if (o->tilesX == tx && o->tilesY == ty) return o.get();
        return nullptr;
    }
    // Last frame.
    gsr_camera lastCam{};
    bool haveFrame = false;
    int sortedIn = 0;  // 0 = A, 1 = B
    hipStream_t lastStream = nullptr;
    hipEvent_t order = nullptr;  // marks the last frame on lastStream (frames on another stream wait on it)
    hipEvent_t events[GSR_STAGE_COUNT] = {};
    bool eventsRecorded = false;
    std::unique_ptr<KernelMarks> marks;  // profile_stages = 2: an event after every kernel
    bool marksRecorded = false;
    // The lists' bins: TileSettings::tileSize (16 or 8 px).
    uint32_t binSize = 16;
    uint32_t binsX(const gsr_camera& c) const { return (c.width + binSize - 1) / binSize; }
    uint32_t binsY(const gsr_camera& c) const { return (c.height + binSize - 1) / binSize; }
    // tile_stats: kTileStatCount counters per tile of the last frame (sized for the largest grid).
    DeviceBuffer tileStats;
    uint32_t tileStatsTiles = 0;
    gsr_colour_model* colour = nullptr;  // learned colour, or null (SH)
    // A captured batch holds raw pointers to this renderer's buffers: `id` names the renderer for
    // its lifetime (never reused) and `generation` goes up whenever a buffer a frame writes or
    // reads is reallocated or the colour binding changes, so a replay can refuse stale graphs.
    uint64_t id = 0;
    uint64_t generation = 0;
};

namespace {
// Live renderers (id by address), so that a replay can tell a destroyed renderer from a live one,
// even when a new renderer was allocated at the same address.
std::mutex g_liveMutex;
std::unordered_map<const gsr_renderer*, uint64_t> g_liveRenderers;
std::atomic<uint64_t> g_nextRendererId{1};

bool rendererLive(const gsr_renderer* r, uint64_t id) {
    std::lock_guard<std::mutex> lock(g_liveMutex);
    const auto it = g_liveRenderers.find(r);
    return it != g_liveRenderers.end() && it->second == id;
}
}  // namespace

extern "C" {

const char* gsr_last_error(void) { return g_lastError.c_str(); }
const char* gsr_version(void) { return "gsr-mi355x 0.1 (gfx950)"; }

gsr_status gsr_context_create(int32_t device, gsr_context** out) {
    if (!out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_context_create: out is null");
    *out = nullptr;
    int n = 0;
    GSR_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_context_create: device %d of %d", device, n);
    GSR_HIP(hipSetDevice(device));
    gsr_context* c = new (std::nothrow) gsr_context();
    if (!c) return fail(GSR_ERROR_OUT_OF_MEMORY, "gsr_context_create: host allocation");
    c->device = device;
    hipError_t e = resolveScratch(device, &c->noAllocScratch);
    if (e != hipSuccess) {
        delete c;
        return hipFail(e, "gsr_context_create: hipGetSymbolAddress(scratch)");
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hipFail(e, "hipStreamCreateWithFlags");
    }
    *out = c;
    return GSR_OK;
}

gsr_status gsr_context_synchronize(gsr_context* ctx) {
    if (!ctx) return fail(GSR_ERROR_INVALID_ARGUMENT, "null context");
    GSR_HIP(hipSetDevice(ctx->device));
    GSR_HIP(hipDeviceSynchronize());
    return GSR_OK;
}

void gsr_context_destroy(gsr_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->work.release();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

gsr_status gsr_buffer_create(gsr_context* ctx, size_t bytes, gsr_buffer** out) {
    if (!ctx || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_buffer_create: null argument");
    *out = nullptr;
    GSR_HIP(hipSetDevice(ctx->device));
    gsr_buffer* b = new (std::nothrow) gsr_buffer();
    if (!b) return fail(GSR_ERROR_OUT_OF_MEMORY, "gsr_buffer_create: host allocation");
    b->ctx = ctx;
    const hipError_t e = b->mem.allocate(bytes);
    if (e != hipSuccess) {
        delete b;
        return hipFail(e, "gsr_buffer_create: hipMalloc");
    }
    *out = b;
    return GSR_OK;
}

void* gsr_buffer_device_ptr(gsr_buffer* buf) { return buf ? buf->mem.ptr : nullptr; }

gsr_status gsr_buffer_read(gsr_buffer* buf, void* dst, size_t bytes) {
    if (!buf || !dst || bytes > buf->mem.bytes) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_buffer_read: bad arguments");
    GSR_HIP(hipSetDevice(buf->ctx->device));
    GSR_HIP(hipDeviceSynchronize());
    GSR_HIP(hipMemcpy(dst, buf->mem.ptr, bytes, hipMemcpyDeviceToHost));
    return GSR_OK;
}

gsr_status gsr_buffer_write(gsr_buffer* buf, const void* src, size_t bytes) {
    if (!buf || !src || bytes > buf->mem.bytes) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_buffer_write: bad arguments");
    GSR_HIP(hipSetDevice(buf->ctx->device));
    GSR_HIP(hipDeviceSynchronize());
    GSR_HIP(hipMemcpy(buf->mem.ptr, src, bytes, hipMemcpyHostToDevice));
    return GSR_OK;
}

void gsr_buffer_destroy(gsr_buffer* buf) {
    if (!buf) return;
    (void)hipSetDevice(buf->ctx->device);
    (void)hipDeviceSynchronize();
    delete buf;
}

// --- Scene ---------------------------------------------------------------------------------

gsr_status gsr_scene_create(gsr_context* ctx, const gsr_scene_desc* d, gsr_scene** out) {
    if (!ctx || !d || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_scene_create: null argument");
    *out = nullptr;
    if (d->sh_degree < 0 || d->sh_degree > 3)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_scene_create: sh_degree %d outside 0..3", d->sh_degree);
    if (d->format != GSR_SCENE_F32_SOA && d->format != GSR_SCENE_F16)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_scene_create: unknown format %d", int(d->format));
    const uint32_t K = uint32_t((d->sh_degree + 1) * (d->sh_degree + 1));
    if (d->count > 0) {
        if (d->format == GSR_SCENE_F32_SOA && (!d->pos_alpha || !d->scales || !d->rotations || !d->sh))
            return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_scene_create: f32 SoA needs pos_alpha, scales, rotations, sh");
        if (d->format == GSR_SCENE_F16 && (!d->geometry || !d->color))
            return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_scene_create: f16 needs geometry and color records");
    }
    GSR_HIP(hipSetDevice(ctx->device));
    gsr_scene* s = new (std::nothrow) gsr_scene();
    if (!s) return fail(GSR_ERROR_OUT_OF_MEMORY, "gsr_scene_create: host allocation");
    s->ctx = ctx;
    s->count = d->count;
    s->shDegree = d->sh_degree;
    s->K = K;
    s->format = d->format;
    const size_t n = d->count;
    const uint32_t colorHalves = (K * 3 + 7) / 8 * 8;
    auto bind = [&](DeviceBuffer& buf, const void* src, size_t bytes, const void** dst) -> hipError_t {
        if (d->device_pointers || n == 0) {
            *dst = src;
            return hipSuccess;
        }
        hipError_t e = buf.allocate(bytes);
        if (e != hipSuccess) return e;
        e = hipMemcpy(buf.ptr, src, bytes, hipMemcpyHostToDevice);
        *dst = buf.ptr;
        return e;
    };
    // Records wider than 16 B are re-laid out as planes of 16-B quads in library-owned memory
    // (also for device pointers): the source goes to the device as it is (or is read in
    // place), and a layout kernel writes the planes.
    hipStream_t stream = ctx->stream;
    DeviceBuffer staging;
    auto planes = [&](DeviceBuffer& buf, const void* src, size_t srcBytes, size_t planeBytes,
                      const void** devSrc) -> hipError_t {
        hipError_t e = buf.allocate(planeBytes);
        if (e != hipSuccess) return e;
        if (d->device_pointers) {
            *devSrc = src;
            return hipSuccess;
        }
        if ((e = staging.allocate(srcBytes)) != hipSuccess) return e;
        *devSrc = staging.ptr;
        return hipMemcpy(staging.ptr, src, srcBytes, hipMemcpyHostToDevice);
    };
    SceneRef& r = s->ref;
    r.count = d->count;
    r.K = K;
    r.format = d->format;
    r.degree = d->sh_degree;
    r.colorQuads = colorHalves / 8;
    r.shQuads = (K * 3 + 3) / 4;
    hipError_t e = hipSuccess;
    const void* p = nullptr;
    if (d->format == GSR_SCENE_F32_SOA) {
        if ((e = bind(s->posAlpha, d->pos_alpha, n * 16, &p)) == hipSuccess) r.posAlpha = static_cast<const float4*>(p);
        if (e == hipSuccess && (e = bind(s->scales, d->scales, n * 16, &p)) == hipSuccess)
            r.scales = static_cast<const float4*>(p);
        if (e == hipSuccess && (e = bind(s->rotations, d->rotations, n * 16, &p)) == hipSuccess)
            r.rotations = static_cast<const float4*>(p);
        if (e == hipSuccess && n > 0 && (e = planes(s->sh, d->sh, n * K * 3 * 4, n * r.shQuads * 16, &p)) == hipSuccess) {
            launchFloatsToPlanes(static_cast<const float*>(p), uint32_t(n), K * 3, r.shQuads, s->sh.as<float4>(), stream);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            r.shPlanes = s->sh.as<float4>();
        }
    } else if (n > 0) {
        if ((e = planes(s->geometry, d->geometry, n * 32, n * 32, &p)) == hipSuccess) {
            launchQuadsToPlanes(static_cast<const uint4*>(p), uint32_t(n), 2, s->geometry.as<uint4>(), stream);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            r.geometry = s->geometry.as<uint4>();
        }
        if (e == hipSuccess && (e = planes(s->color, d->color, n * colorHalves * 2, n * colorHalves * 2, &p)) == hipSuccess) {
            launchQuadsToPlanes(static_cast<const uint4*>(p), uint32_t(n), r.colorQuads, s->color.as<uint4>(), stream);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            r.color = s->color.as<uint4>();
        }
    }
    if (e != hipSuccess) {
        delete s;  // frees the partial upload (gpu_scene.cpp:108-114)
        return hipFail(e, "gsr_scene_create: upload");
    }
    *out = s;
    return GSR_OK;
}

void gsr_scene_destroy(gsr_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipDeviceSynchronize();
    delete s;
}

// --- fp16 encoder (host) ---------------------------------------------------------------------

namespace {
uint16_t floatToHalfRne(float value) {
    uint32_t f;
    std::memcpy(&f, &value, 4);
    const uint32_t sign = (f >> 16) & 0x8000u;
    const uint32_t a = f & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return uint16_t(sign | 0x7C00u | (a > 0x7F800000u ? 0x200u : 0u));  // Inf / NaN
    if (a >= 0x477FF000u) return uint16_t(sign | 0x7C00u);  // rounds past 65504
    if (a < 0x38800000u) {                                  // half subnormal or zero
        if (a < 0x33000000u) return uint16_t(sign);
        const uint32_t e = a >> 23;
        const uint32_t m = (a & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - e;
        uint32_t h = m >> shift;
        const uint32_t rest = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1u);
        if (rest > half || (rest == half && (h & 1u))) ++h;
        return uint16_t(sign | h);
    }
    uint32_t h = (a - 0x38000000u) >> 13;  // rebias 127 -> 15
    const uint32_t rest = a & 0x1FFFu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) ++h;
    return uint16_t(sign | h);
}
uint32_t asBits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
}  // namespace

gsr_status gsr_encode_f16(uint32_t count, int32_t shDegree, const float* posAlpha, const float* scales,
                          const float* rotations, const float* sh, uint32_t* geometry, uint16_t* color) {
    if (shDegree < 0 || shDegree > 3) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_encode_f16: sh_degree %d", shDegree);
    if (count && (!posAlpha || !scales || !rotations || !sh || !geometry || !color))
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_encode_f16: null array");
    const uint32_t K = uint32_t((shDegree + 1) * (shDegree + 1));
    const uint32_t colorHalves = (K * 3 + 7) / 8 * 8;
    for (uint32_t i = 0; i < count; ++i) {
        const float* pa = posAlpha + size_t(i) * 4;
        const float* sc = scales + size_t(i) * 4;
        const float* q = rotations + size_t(i) * 4;
        uint32_t* w = geometry + size_t(i) * 8;
        w[0] = asBits(pa[0]);
        w[1] = asBits(pa[1]);
        w[2] = asBits(pa[2]);
        w[3] = uint32_t(floatToHalfRne(pa[3])) | (uint32_t(floatToHalfRne(sc[0])) << 16);
        w[4] = uint32_t(floatToHalfRne(sc[1])) | (uint32_t(floatToHalfRne(sc[2])) << 16);
        w[5] = uint32_t(floatToHalfRne(q[0])) | (uint32_t(floatToHalfRne(q[1])) << 16);
        w[6] = uint32_t(floatToHalfRne(q[2])) | (uint32_t(floatToHalfRne(q[3])) << 16);
        w[7] = 0u;
        uint16_t* c = color + size_t(i) * colorHalves;
        const float* src = sh + size_t(i) * K * 3;
        for (uint32_t j = 0; j < colorHalves; ++j) c[j] = j < K * 3 ? floatToHalfRne(src[j]) : uint16_t(0);
    }
    return GSR_OK;
}

// --- Camera ------------------------------------------------------------------------------------

gsr_status gsr_camera_make(const float viewInria[16], uint32_t width, uint32_t height, float tanX, float tanY,
                           gsr_camera* out) {
    if (!viewInria || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_camera_make: null argument");
    if (!(tanX > 0.0f) || !(tanY > 0.0f)) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_camera_make: tan(fov/2) <= 0");
    gsr_camera c{};
    std::memcpy(c.view_inria, viewInria, sizeof(c.view_inria));
    // Camera centre c = -R^T t (viewInria * c = 0 for a rigid transform).
    const float* v = viewInria;
    for (int i = 0; i < 3; ++i) {
        const float s = v[i * 4 + 0] * v[12] + v[i * 4 + 1] * v[13] + v[i * 4 + 2] * v[14];
        c.position[i] = -s;
    }
    c.width = width;
    c.height = height;
    c.tan_half_fov_x = tanX;
    c.tan_half_fov_y = tanY;
    c.focal_x = float(width) / (2.0f * tanX);
    c.focal_y = float(height) / (2.0f * tanY);
    c.tiles_x = (width + kTileSize - 1) / kTileSize;
    c.tiles_y = (height + kTileSize - 1) / kTileSize;
    *out = c;
    return GSR_OK;
}

gsr_status gsr_camera_from_viewer(const float view[16], const float proj[16], const float model[16], uint32_t width,
                                  uint32_t height, gsr_camera* out) {
    if (!view || !proj || !model || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_camera_from_viewer: null argument");
    // viewInria = diag(1, -1, -1, 1) * view * model, column-major products summed in order.
    auto mul = [](const float* A, const float* B, float* R) {
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i)
                R[j * 4 + i] =
                    A[i] * B[j * 4 + 0] + A[4 + i] * B[j * 4 + 1] + A[8 + i] * B[j * 4 + 2] + A[12 + i] * B[j * 4 + 3];
    };
    const float flip[16] = {1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1};
    float fv[16], vi[16];
    mul(flip, view, fv);
    mul(fv, model, vi);
    const float tanX = 1.0f / proj[0];
    const float tanY = 1.0f / std::fabs(proj[5]);
    return gsr_camera_make(vi, width, height, tanX, tanY, out);
}

// --- Renderer ----------------------------------------------------------------------------------

static uint32_t capacityForBytes(double bytes, bool binning) {
    // Per entry: key + value (64-bit keys in global mode), both ping-ponged, + the sort's
    // block histograms (0.5 B): 16.5 or 24.5 B (splat_tile_renderer.cpp:117-124), in blocks
    // of 2048 entries.
    const double maxEntries = double(1u << 30);
    double c = bytes / (binning ? 16.5 : 24.5);
    c = std::min(std::max(c, 2048.0), maxEntries);
    const uint32_t e = uint32_t(c);
    return e - e % 2048u;
}

static uint32_t roundUpToBlock(uint64_t n) {
    const uint64_t b = kSortTile;
    return uint32_t(std::min<uint64_t>((n + b - 1) / b * b, 0xFFFFF800ull));
}

// The list buffers for `capacity` entries: keys and values ping-pong (+ the one-sweep status
// words), zeroed so that the lists start as valid keys and gaussian indices. The value buffers
// hold kListPad more entries (zero: gaussian 0), which the tile depth sort may read past the
// list's end. Adaptive lists are stream-ordered allocations, so a resize never waits for the
// frames in flight.
static hipError_t allocateList(gsr_renderer* r, uint32_t capacity, hipStream_t stream) {
    const size_t keyBytes = r->tileKeys32() ? 4 : 8;
    const bool async = r->settings.adaptive_list != 0;
    hipError_t e = hipSuccess;
    for (DeviceBuffer* b : {&r->keysA, &r->keysB, &r->valuesA, &r->valuesB, &r->sortStatus}) {
        const size_t bytes = b == &r->sortStatus ? sortStatusBytes(capacity)
                             : (b == &r->keysA || b == &r->keysB) ? size_t(capacity) * keyBytes
                                                                  : (size_t(capacity) + kListPad) * 4;
        if (e == hipSuccess) e = async ? b->allocateAsync(bytes, stream) : b->allocate(bytes);
        if (e == hipSuccess) e = async ? hipMemsetAsync(b->ptr, 0, b->bytes, stream) : hipMemset(b->ptr, 0, b->bytes);
    }
    if (e == hipSuccess) r->capacity = capacity;
    return e;
}

// Makes `stream` wait for the renderer's last frame when that frame ran on another stream. The
// frame's end is an event recorded on its own stream (never a handle to that stream, which may be
// gone by now). Frames recorded into a capture record no event, so a replayed graph is ordered by
// the stream it is launched on.
static hipError_t orderAfterLastFrame(gsr_renderer* r, hipStream_t stream, bool capturing) {
    if (capturing || !r->lastStream || r->lastStream == stream) return hipSuccess;
    return hipStreamWaitEvent(stream, r->order, 0);
}

// Collects the frame statistics whose copies have completed (lastStats = the newest).
static void collectStats(gsr_renderer* r) {
    for (int s = 0; s < gsr_renderer::kStatSlots; ++s) {
        if (!r->statPending[s] || hipEventQuery(r->statEvent[s]) != hipSuccess) continue;
        r->statPending[s] = false;
        if (r->lastStatsValid && r->statFrame[s] < r->lastStatsFrame) continue;
        r->lastStats = r->statHost[s];
        r->lastStatsFrame = r->statFrame[s];
        r->lastStatsValid = true;
    }
}

// adaptCapacity (splat_tile_renderer.cpp:444-486): grow at once to 1.5 M when M > 0.9 capacity
// (M is the unclamped entry count, the true need even for a frame that overflowed); shrink to
// 1.5 x the window's largest M after 120 consecutive frames below 0.4 capacity; never below
// min(256k, max), never above the budget, and not for a shrink of less than a quarter.
static hipError_t adaptCapacity(gsr_renderer* r, hipStream_t stream) {
    if (!r->settings.adaptive_list || !r->lastStatsValid || r->lastStatsFrame == r->lastAdaptedFrame) return hipSuccess;
    r->lastAdaptedFrame = r->lastStatsFrame;
    const uint32_t m = r->lastStats.duplicates;
    const uint32_t cap = r->capacity;
    uint64_t target = 0;
    if (double(m) > 0.9 * cap) {
        target = uint64_t(1.5 * m);
        r->shrinkFrames = 0;
        r->shrinkWindowMax = 0;
    } else if (double(m) < 0.4 * cap) {
        r->shrinkWindowMax = std::max(r->shrinkWindowMax, m);
        if (++r->shrinkFrames >= 120) {
            target = uint64_t(1.5 * r->shrinkWindowMax);
            r->shrinkFrames = 0;
            r->shrinkWindowMax = 0;
        }
    } else {
        r->shrinkFrames = 0;
        r->shrinkWindowMax = 0;
    }
    if (target == 0) return hipSuccess;
    const uint32_t minCap = std::min<uint32_t>(256u * 1024u, r->maxCapacity);
    const uint32_t newCap = std::clamp(roundUpToBlock(target), minCap, r->maxCapacity);
    if (newCap == cap || (newCap < cap && double(newCap) > 0.75 * cap)) return hipSuccess;
    // The frames in flight may still read the old list: its release is ordered after them
    // (gsr_render has made `stream` wait for the renderer's last frame).
    for (DeviceBuffer* b : {&r->keysA, &r->keysB, &r->valuesA, &r->valuesB, &r->sortStatus}) {
        const hipError_t e = b->releaseAsync(stream);
        if (e != hipSuccess) return e;
    }
    ++r->generation;
    const hipError_t e = allocateList(r, newCap, stream);
    if (e == hipSuccess) {
        ++r->listResizes;
        return e;
    }
    // The new size did not fit: back to the old one (the frame reports the failure); if even
    // that fails, the renderer has no lists and every later frame reports OUT_OF_MEMORY.
    (void)hipGetLastError();
    for (DeviceBuffer* b : {&r->keysA, &r->keysB, &r->valuesA, &r->valuesB, &r->sortStatus}) (void)b->releaseAsync(stream);
    r->capacity = 0;
    (void)allocateList(r, cap, stream);
    return e;
}

gsr_status gsr_renderer_create(gsr_context* ctx, const gsr_scene* scene, const gsr_settings* settings,
                               gsr_renderer** out) {
    if (!ctx || !scene || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: null argument");
    *out = nullptr;
    gsr_settings s{};
    if (settings) s = *settings;
    else {
        s.bounds = GSR_BOUNDS_TIGHT;
        s.max_sh_degree = 3;
    }
    if (s.bounds != GSR_BOUNDS_RECT && s.bounds != GSR_BOUNDS_TIGHT && s.bounds != GSR_BOUNDS_CUTOFF_ONLY)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: unknown bounds %d", int(s.bounds));
    if (s.format != GSR_TARGET_RGBA8 && s.format != GSR_TARGET_RGBA32F)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: unknown target format %d", int(s.format));
    if (s.budget_mb == 0) s.budget_mb = 256;
    if (s.render_pixels_per_thread == 0) s.render_pixels_per_thread = 1;
    if (s.sort_algorithm != GSR_SORT_REDUCE_THEN_SCAN && s.sort_algorithm != GSR_SORT_ONESWEEP)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: unknown sort_algorithm %d", int(s.sort_algorithm));
    if (s.render_splat != GSR_RENDER_SPLAT_FULL && s.render_splat != GSR_RENDER_SPLAT_P32 &&
        s.render_splat != GSR_RENDER_SPLAT_P24)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: unknown render_splat %d", int(s.render_splat));
    if (s.sort_window > uint32_t(kMaxSortWindowHeap))
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: sort_window %u must be <= %d", s.sort_window,
                    kMaxSortWindowHeap);
    if (s.sort_mode != GSR_SORT_MODE_BINNING && s.sort_mode != GSR_SORT_MODE_GLOBAL)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: unknown sort_mode %d", int(s.sort_mode));
    if (s.sort_mode == GSR_SORT_MODE_BINNING && s.sort_algorithm == GSR_SORT_ONESWEEP)
        return fail(GSR_ERROR_INVALID_ARGUMENT,
                    "gsr_renderer_create: the one-sweep sort takes 64-bit keys (sort_mode GLOBAL)");
    if (s.render_loop != GSR_RENDER_LOOP_CULLED && s.render_loop != GSR_RENDER_LOOP_DIVERGENT &&
        s.render_loop != GSR_RENDER_LOOP_PREDICATED)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: unknown render_loop %d", int(s.render_loop));
    if (s.render_pixels_per_thread != 1 && s.render_pixels_per_thread != 2 && s.render_pixels_per_thread != 4)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: render_pixels_per_thread must be 1, 2 or 4");
    if (s.differentiable &&
        (s.format != GSR_TARGET_RGBA32F || s.render_splat != GSR_RENDER_SPLAT_FULL || s.sort_window != 0 ||
         scene->format != GSR_SCENE_F32_SOA))
        return fail(GSR_ERROR_UNSUPPORTED,
                    "gsr_renderer_create: differentiable needs format RGBA32F, render_splat FULL, sort_window 0 "
                    "and an f32 scene");
    if (s.profile_stages < 0 || s.profile_stages > 2)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_create: profile_stages must be 0, 1 or 2");
    if (s.tile_size == 0) s.tile_size = 16;
    if (s.tile_size != 8 && s.tile_size != 16)
        return fail(GSR_ERROR_INVALID_ARGUMENT,
                    "gsr_renderer_create: tile_size must be 8 or 16 (a divisor of INRIA's 16-px tiles), got %d",
                    s.tile_size);
    const bool bins8 = s.tile_size == 8;
    if (s.render_batch == 0) s.render_batch = bins8 ? 64u : 256u;
    if (bins8 ? (s.render_batch != 32 && s.render_batch != 64 && s.render_batch != 128)
              : (s.render_batch != 128 && s.render_batch != 256 && s.render_batch != 512))
        return fail(GSR_ERROR_INVALID_ARGUMENT,
                    "gsr_renderer_create: render_batch %u must be %s", s.render_batch,
                    bins8 ? "32, 64 or 128 (splats per 8x8 bin and batch)" : "128, 256 or 512 (splats per tile and batch)");
    if (s.render_batch != (bins8 ? 64u : 256u) &&
        (s.render_loop != GSR_RENDER_LOOP_CULLED || s.render_pixels_per_thread != 1 || s.render_exact_alpha ||
         s.sort_window != 0 || s.differentiable || s.tile_stats))
        return fail(GSR_ERROR_UNSUPPORTED,
                    "gsr_renderer_create: render_batch other than the default needs the culled loop (one pixel per "
                    "thread, fast alpha, sort_window 0, not differentiable, no tile_stats)");
    if (s.tile_size == 8 &&
        (s.render_loop != GSR_RENDER_LOOP_CULLED || s.render_pixels_per_thread != 1 || s.render_exact_alpha ||
         s.sort_window != 0 || s.differentiable || s.tile_stats))
        return fail(GSR_ERROR_UNSUPPORTED,
                    "gsr_renderer_create: tile_size 8 renders with the culled loop only (render_loop CULLED, one "
                    "pixel per thread, fast alpha, sort_window 0, not differentiable, no tile_stats)");
    if (s.tile_stats &&
        (s.render_loop != GSR_RENDER_LOOP_CULLED || s.render_pixels_per_thread != 1 || s.render_exact_alpha ||
         s.sort_window != 0 || s.differentiable))
        return fail(GSR_ERROR_UNSUPPORTED,
                    "gsr_renderer_create: tile_stats counts the culled render loop (render_loop CULLED, one pixel per "
                    "thread, fast alpha, sort_window 0, not differentiable)");
    GSR_HIP(hipSetDevice(ctx->device));
    gsr_renderer* r = new (std::nothrow) gsr_renderer();
    if (!r) return fail(GSR_ERROR_OUT_OF_MEMORY, "gsr_renderer_create: host allocation");
    r->ctx = ctx;
    r->scene = scene;
    r->settings = s;
    r->binSize = uint32_t(s.tile_size);
    r->degree = std::min(scene->shDegree, std::max(0, std::min(3, s.max_sh_degree)));
    r->binning = s.sort_mode == GSR_SORT_MODE_BINNING;
    r->tileSort = !r->binning && !s.full_key_sort && s.sort_algorithm == GSR_SORT_REDUCE_THEN_SCAN;
    r->maxCapacity = capacityForBytes(double(s.budget_mb) * 1024.0 * 1024.0, r->binning);
    uint32_t initialCapacity = r->maxCapacity;
    if (s.adaptive_list) {
        // A first guess (4 entries per gaussian); the measured count corrects it a few frames later.
        const uint64_t guess = s.adaptive_initial_entries ? uint64_t(s.adaptive_initial_entries)
                                                          : std::max<uint64_t>(4ull * scene->count, 1u << 20);
        initialCapacity = std::min(r->maxCapacity, roundUpToBlock(guess));
    }
    const size_t n = std::max<size_t>(scene->count, 1);
    const size_t gBlocks = (n + kPreprocessThreads - 1) / kPreprocessThreads;
    hipError_t e = hipSuccess;
    auto alloc = [&](DeviceBuffer& b, size_t bytes) {
        if (e == hipSuccess) e = b.allocate(bytes);
    };
    alloc(r->splats, n * 48);
    if (s.render_splat != GSR_RENDER_SPLAT_FULL) alloc(r->packed, n * (s.render_splat == GSR_RENDER_SPLAT_P32 ? 32 : 24));
    if (s.sort_window > 0) alloc(r->rayDepth, n * 48);
    if (s.differentiable) alloc(r->grad2D, n * 4 * size_t(grad2DComponents()));
    alloc(r->tileCounts, n * 4);
    alloc(r->spanInfo, n * 4);
    alloc(r->spanPlanes, n * 32);
    alloc(r->blockOffsets, gBlocks * 4);
    alloc(r->control, sizeof(gsr_control));
    if (r->binning) {
        for (DeviceBuffer* b : {&r->depthKeysA, &r->depthKeysB, &r->depthValuesA, &r->depthValuesB, &r->sortedCounts})
            alloc(*b, n * 4);
        alloc(r->visibleSums, gBlocks * 4);
    }
    if (!r->binning) alloc(r->depthBits, n * 4);
    if (!r->binning) alloc(r->superStats, size_t(superBlockCount(uint32_t(n))) * sizeof(BlockStats));
    alloc(r->sortSmall, (kMaxSortPasses * kRadix + kMaxSortPasses + size_t(kRadix) * kMaxSortBlocks) * 4);
    if (e == hipSuccess) e = allocateList(r, initialCapacity, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemset(r->control.ptr, 0, sizeof(gsr_control));
    if (e == hipSuccess && r->superStats.ptr) e = hipMemset(r->superStats.ptr, 0, r->superStats.bytes);
    for (int i = 0; i < GSR_STAGE_COUNT && e == hipSuccess; ++i) e = hipEventCreate(&r->events[i]);
    if (s.profile_stages == 2 && e == hipSuccess) {
        r->marks.reset(new (std::nothrow) KernelMarks());
        if (!r->marks) e = hipErrorOutOfMemory;
        for (int i = 0; i <= KernelMarks::kMax && e == hipSuccess; ++i) e = hipEventCreate(&r->marks->events[i]);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->order, hipEventDisableTiming);
    if (s.adaptive_list || s.frame_stats) {
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&r->statHost), sizeof(gsr_control) * gsr_renderer::kStatSlots,
                              hipHostMallocDefault);
        for (int i = 0; i < gsr_renderer::kStatSlots && e == hipSuccess; ++i)
            e = hipEventCreateWithFlags(&r->statEvent[i], hipEventDisableTiming);
    }
    r->id = g_nextRendererId.fetch_add(1);
    {
        std::lock_guard<std::mutex> lock(g_liveMutex);
        g_liveRenderers[r] = r->id;
    }
    if (e != hipSuccess) {
        gsr_renderer_destroy(r);
        return hipFail(e, "gsr_renderer_create: allocation");
    }
    *out = r;
    return GSR_OK;
}

void gsr_renderer_destroy(gsr_renderer* r) {
    if (!r) return;
    {
        std::lock_guard<std::mutex> lock(g_liveMutex);
        g_liveRenderers.erase(r);
    }
    (void)hipSetDevice(r->ctx->device);
    (void)hipDeviceSynchronize();
    if (r->order) (void)hipEventDestroy(r->order);
    for (hipEvent_t& ev : r->events)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t& ev : r->statEvent)
        if (ev) (void)hipEventDestroy(ev);
    if (r->marks)
        for (hipEvent_t& ev : r->marks->events)
            if (ev) (void)hipEventDestroy(ev);
    if (r->statHost) (void)hipHostFree(r->statHost);
    delete r;
}

uint32_t gsr_renderer_capacity(const gsr_renderer* r) { return r ? r->capacity : 0; }
uint32_t gsr_renderer_max_capacity(const gsr_renderer* r) { return r ? r->maxCapacity : 0; }
uint32_t gsr_renderer_list_resizes(const gsr_renderer* r) { return r ? r->listResizes : 0; }

gsr_status gsr_renderer_last_stats(gsr_renderer* r, gsr_control* out, uint64_t* frameIndex) {
    if (!r || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_last_stats: null argument");
    if (!r->statHost)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_last_stats: needs frame_stats or adaptive_list");
    GSR_HIP(hipSetDevice(r->ctx->device));
    collectStats(r);
    if (!r->lastStatsValid)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_last_stats: no frame has completed yet");
    *out = r->lastStats;
    if (frameIndex) *frameIndex = r->lastStatsFrame;
    return GSR_OK;
}

uint64_t gsr_renderer_total_bytes(const gsr_renderer* r) {
    if (!r) return 0;
    const DeviceBuffer* bufs[] = {&r->splats, &r->tileCounts, &r->blockOffsets, &r->blockStats, &r->control,
                                  &r->keysA,  &r->keysB,      &r->valuesA,      &r->valuesB,    &r->sortSmall,
                                  &r->sortStatus, &r->ranges, &r->depthKeysA, &r->depthKeysB, &r->depthValuesA,
                                  &r->depthValuesB, &r->sortedCounts, &r->visibleSums, &r->packed, &r->rayDepth,
                                  &r->pixelState, &r->grad2D, &r->tileStats, &r->depthBits, &r->superStats, &r->spanInfo, &r->spanPlanes};
    uint64_t s = 0;
    for (const DeviceBuffer* b : bufs) s += b->bytes;
    for (const auto& o : r->tileOrders) s += o->table.bytes;
    return s;
}

// One frame on `stream`. deferred != null: everything but the render pass, whose tiles are
// described in *deferred instead (a batch's persistent render draws them later).
static gsr_status renderFrame(gsr_renderer* r, const gsr_camera* cam, void* outImage, hipStream_t streamArg,
                              TileJob* deferred) {
    if (!r || !cam || !outImage) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_render: null argument");
    if (cam->width == 0 || cam->height == 0 || cam->width > 65536 || cam->height > 65536)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_render: image size %ux%u", cam->width, cam->height);
    if (cam->tiles_x != (cam->width + 15) / 16 || cam->tiles_y != (cam->height + 15) / 16)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_render: camera tiles do not match its size (use gsr_camera_make)");
    GSR_HIP(hipSetDevice(r->ctx->device));
    hipStream_t stream = streamArg ? streamArg : r->ctx->stream;
    hipStreamCaptureStatus captureStatus = hipStreamCaptureStatusNone;
    GSR_HIP(hipStreamIsCapturing(stream, &captureStatus));
    const bool capturing = captureStatus != hipStreamCaptureStatusNone;
    // A frame reuses the previous frame's buffers: on another stream it waits for that frame
    // (a captured batch is ordered by its own fork event instead).
    GSR_HIP(orderAfterLastFrame(r, stream, capturing));
    const gsr_scene* sc = r->scene;
    // The lists' grid: bins of binSize px (the camera's tiles_x/y are INRIA's 16-px tiles).
    const uint32_t binsX = r->binsX(*cam), binsY = r->binsY(*cam);
    const uint32_t tiles = binsX * binsY;
    // Per-size buffers (onRenderTargetsChanged, splat_tile_renderer.cpp:537-570).
    if (tiles > r->rangesTiles) {
        GSR_HIP(hipStreamSynchronize(stream));
        ++r->generation;
        r->rangesTiles = 0;  // no buffer until the allocation succeeds (a failed one is retried)
        GSR_HIP(r->ranges.allocate(size_t(tiles) * 8));
        r->rangesTiles = tiles;
    }
    if (r->settings.differentiable && size_t(cam->width) * cam->height > r->pixelStatePixels) {
        GSR_HIP(hipStreamSynchronize(stream));
        ++r->generation;
        r->pixelStatePixels = 0;
        GSR_HIP(r->pixelState.allocate(size_t(cam->width) * cam->height * 8));
        r->pixelStatePixels = size_t(cam->width) * cam->height;
    }
    if (r->settings.tile_stats && tiles > r->tileStatsTiles) {
        GSR_HIP(hipStreamSynchronize(stream));
        ++r->generation;
        r->tileStatsTiles = 0;
        GSR_HIP(r->tileStats.allocate(size_t(tiles) * kTileStatCount * 4));
        r->tileStatsTiles = tiles;
    }
    const uint32_t* tileOrder = nullptr;
    if (r->tileSort) {
        const gsr_renderer::TileOrder* o = r->findOrder(binsX, binsY);
        if (!o) {  // a new grid: its table, kept for later frames of the size
            std::vector<uint32_t> order(tiles);
            for (uint32_t t = 0; t < tiles; ++t) order[t] = t;
            const int64_t tx = binsX, ty = binsY;
            auto centreKey = [&](uint32_t t) {
                const int64_t row = t / tx, col = t % tx;
                return std::llabs(2 * row - (ty - 1)) * (2 * tx + 1) + std::llabs(2 * col - (tx - 1));
            };
            std::stable_sort(order.begin(), order.end(),
                             [&](uint32_t a, uint32_t b) { return centreKey(a) < centreKey(b); });
            auto entry = std::make_unique<gsr_renderer::TileOrder>();
            GSR_HIP(entry->table.allocate(size_t(tiles) * 4));
            GSR_HIP(hipMemcpy(entry->table.ptr, order.data(), size_t(tiles) * 4, hipMemcpyHostToDevice));
            entry->tilesX = binsX;
            entry->tilesY = binsY;
            r->tileOrders.push_back(std::move(entry));
            o = r->tileOrders.back().get();
        }
        tileOrder = o->table.as<uint32_t>();
    }
    const uint32_t rangeWords = 2 * tiles;
    const uint32_t preBlocks = (std::max(sc->count, rangeWords) + kPreprocessThreads - 1) / kPreprocessThreads;
    if (r->blockStats.bytes < size_t(preBlocks) * sizeof(BlockStats)) {
        GSR_HIP(hipStreamSynchronize(stream));
        ++r->generation;
        GSR_HIP(r->blockStats.allocate(size_t(preBlocks) * sizeof(BlockStats)));
    }
    DevCamera dc{};
    std::memcpy(dc.view, cam->view_inria, sizeof(dc.view));
    std::memcpy(dc.pos, cam->position, sizeof(dc.pos));
    dc.tanX = cam->tan_half_fov_x;
    dc.tanY = cam->tan_half_fov_y;
    dc.focalX = cam->focal_x;
    dc.focalY = cam->focal_y;
    dc.width = cam->width;
    dc.height = cam->height;
    dc.tilesX = cam->tiles_x;
    dc.tilesY = cam->tiles_y;
    dc.binSize = r->binSize;
    dc.binsX = binsX;
    dc.binsY = binsY;
    const BinGrid grid{cam->tiles_x, cam->tiles_y, r->binSize, binsX, binsY};
    SceneRef ref = sc->ref;
    ref.degree = r->degree;

    if (r->statHost) {
        collectStats(r);
        GSR_HIP(adaptCapacity(r, stream));
    }
    if (r->capacity == 0)
        return fail(GSR_ERROR_OUT_OF_MEMORY, "gsr_render: the entry lists could not be reallocated after a failed resize");
    const bool prof = r->settings.profile_stages != 0;
    auto mark = [&](int i) -> hipError_t { return prof ? hipEventRecord(r->events[i], stream) : hipSuccess; };
    gsr_control* control = r->control.as<gsr_control>();
    const uint32_t gBlocks = (sc->count + kPreprocessThreads - 1) / kPreprocessThreads;
    const uint32_t tileBits = bitsFor(tiles);  // tile ids 0..tiles (tiles = the sentinel)
    uint32_t* small = r->sortSmall.as<uint32_t>();
    SortWork work{r->settings.sort_algorithm, small, small + kMaxSortPasses * kRadix, r->sortStatus.as<uint64_t>(),
                  &control->sort_timeouts, small + kMaxSortPasses * kRadix + kMaxSortPasses};
    const uint32_t bounds = uint32_t(r->settings.bounds);

    GSR_HIP(mark(0));
    // Per-kernel marks: the launch wrappers record them while t_kernelMarks is set (cleared on
    // every return path by the guard).
    struct MarksGuard {
        ~MarksGuard() { t_kernelMarks = nullptr; }
    } marksGuard;
    if (r->marks && !capturing) {
        r->marks->count = 0;
        GSR_HIP(hipEventRecord(r->marks->events[0], stream));
        t_kernelMarks = r->marks.get();
    }
    BlockStats* superStats = r->binning ? nullptr : r->superStats.as<BlockStats>();
    const uint32_t numSuper = superBlockCount(sc->count);
    launchPreprocess(ref, dc, bounds, r->splats.as<float4>(), r->tileCounts.as<uint32_t>(),
                     r->blockStats.as<BlockStats>(), superStats, control, r->spanInfo.as<uint32_t>(),
                     r->spanPlanes.as<uint4>(), r->binning ? nullptr : r->depthBits.as<uint32_t>(),
                     r->ranges.as<uint32_t>(), rangeWords,
                     uint32_t(r->settings.render_splat), r->packed.as<uint32_t>(), r->rayDepth.as<float4>(),
                     r->colour ? r->colour->useMlp.as<uint8_t>() : nullptr, stream);
    GSR_HIP(hipGetLastError());
    if (r->colour && r->colour->mlpCount > 0) {
        const gsr_colour_model* cm = r->colour;
        launchColourMlp(ref, dc, cm->params.as<float>(), cm->packs.as<uint2>(), cm->latentsBf16.as<uint16_t>(),
                        cm->useMlp.as<uint8_t>(), r->tileCounts.as<uint32_t>(), r->splats.as<float4>(),
                        uint32_t(r->settings.render_splat),
                        r->packed.as<uint32_t>(), stream);
        GSR_HIP(hipGetLastError());
    }
    GSR_HIP(mark(1));
    // Binning mode scans the block sums here; global mode's duplicate pass takes its block offsets
    // from the preprocess's superblock sums and writes the control block itself.
    if (r->binning) {
        launchFrameSetup(r->blockStats.as<BlockStats>(), gBlocks, sc->count, r->capacity, true,
                         r->blockOffsets.as<uint32_t>(), control, stream);
        GSR_HIP(hipGetLastError());
    }
    const DupFrame dupFrame{r->blockStats.as<BlockStats>(), superStats, numSuper, r->capacity};
    const uint32_t superWords = superStats ? numSuper * uint32_t(sizeof(BlockStats) / 4) : 0u;
    int which = 0;
    const uint32_t* sortedValues = nullptr;
    // The per-tile depth sort runs inside the culled 16-px render (its workgroup sorts its tile
    // first) when that render draws the frame now; otherwise as its own pass.
    const bool fuseSort = r->tileSort && !deferred && !r->settings.differentiable && r->settings.sort_window == 0 &&
                          r->settings.render_loop == GSR_RENDER_LOOP_CULLED &&
                          r->settings.render_pixels_per_thread == 1 && !r->settings.render_exact_alpha &&
                          r->binSize == uint32_t(kTileSize);
    FusedSort fused{};
    if (r->binning) {
        // Binning (splat_tile_renderer.cpp:680-760): compact -> depth sort -> gather + scan
        // of the depth-ordered counts -> duplicate (tile keys) -> stable tile sort.
        launchCompactVisible(r->splats.as<float4>(), r->tileCounts.as<uint32_t>(), r->blockOffsets.as<uint32_t>(),
                             sc->count, r->depthKeysA.as<uint32_t>(), r->depthValuesA.as<uint32_t>(), stream);
        GSR_HIP(hipGetLastError());
        GSR_HIP(mark(2));
        const int dw = launchRadixSort32(work, r->depthKeysA.as<uint32_t>(), r->depthValuesA.as<uint32_t>(),
                                         r->depthKeysB.as<uint32_t>(), r->depthValuesB.as<uint32_t>(), sc->count,
                                         &control->visible, 32, stream);
        GSR_HIP(hipGetLastError());
        GSR_HIP(mark(3));
        const uint32_t* order = dw ? r->depthValuesB.as<uint32_t>() : r->depthValuesA.as<uint32_t>();
        launchGatherCounts(r->tileCounts.as<uint32_t>(), order, control, sc->count, r->sortedCounts.as<uint32_t>(),
                           r->visibleSums.as<uint32_t>(), stream);
        GSR_HIP(hipGetLastError());
        launchScanVisibleBlocks(r->visibleSums.as<uint32_t>(), control, stream);
        GSR_HIP(hipGetLastError());
        launchDuplicateBinning(r->splats.as<float4>(), r->sortedCounts.as<uint32_t>(), order,
                               r->visibleSums.as<uint32_t>(), control, r->spanInfo.as<uint32_t>(),
                               r->spanPlanes.as<uint4>(), sc->count, bounds, grid,
                               r->keysA.as<uint32_t>(), r->valuesA.as<uint32_t>(), control, stream);
        GSR_HIP(hipGetLastError());
        GSR_HIP(mark(4));
        which = launchRadixSort32(work, r->keysA.as<uint32_t>(), r->valuesA.as<uint32_t>(), r->keysB.as<uint32_t>(),
                                  r->valuesB.as<uint32_t>(), r->capacity, &control->entries, tileBits, stream);
        GSR_HIP(hipGetLastError());
        GSR_HIP(mark(5));
        sortedValues = which ? r->valuesB.as<uint32_t>() : r->valuesA.as<uint32_t>();
        launchRanges32(which ? r->keysB.as<uint32_t>() : r->keysA.as<uint32_t>(), control, r->capacity, tiles,
                       r->ranges.as<uint2>(), nullptr, 0, stream);
    } else if (r->tileSort) {
        // The tile ids by the radix sort (2 passes of 32-bit keys), then each tile's range by
        // (depth, index) in the tile depth sort: the list of the full 64-bit sort.
        GSR_HIP(mark(2));
        GSR_HIP(mark(3));
        launchDuplicateTiles(r->splats.as<float4>(), r->tileCounts.as<uint32_t>(), dupFrame, r->spanInfo.as<uint32_t>(),
                             r->spanPlanes.as<uint4>(), sc->count, bounds, grid, r->keysA.as<uint32_t>(),
                             r->valuesA.as<uint32_t>(), control, stream);
        GSR_HIP(hipGetLastError());
        GSR_HIP(mark(4));
        which = launchRadixSort32(work, r->keysA.as<uint32_t>(), r->valuesA.as<uint32_t>(), r->keysB.as<uint32_t>(),
                                  r->valuesB.as<uint32_t>(), r->capacity, &control->entries, tileBits, stream);
        GSR_HIP(hipGetLastError());
        uint32_t* vals = which ? r->valuesB.as<uint32_t>() : r->valuesA.as<uint32_t>();
        launchRanges32(which ? r->keysB.as<uint32_t>() : r->keysA.as<uint32_t>(), control, r->capacity, tiles,
                       r->ranges.as<uint2>(), reinterpret_cast<uint32_t*>(superStats), superWords, stream);
        GSR_HIP(hipGetLastError());
        uint32_t* scratch = which ? r->valuesA.as<uint32_t>() : r->valuesB.as<uint32_t>();
        if (fuseSort) {
            fused = FusedSort{vals, r->depthBits.as<uint32_t>(), scratch};
        } else {
            launchTileDepthSort(vals, r->depthBits.as<uint32_t>(), scratch, r->ranges.as<uint2>(), tileOrder, tiles,
                                stream);
            GSR_HIP(hipGetLastError());
        }
        GSR_HIP(mark(5));
        sortedValues = vals;
    } else {
        // The whole 64-bit (tile, depth) keys by the radix sort (full_key_sort, or the one-sweep
        // sort).
        GSR_HIP(mark(2));
        GSR_HIP(mark(3));
        launchDuplicate(r->splats.as<float4>(), r->tileCounts.as<uint32_t>(), dupFrame, r->spanInfo.as<uint32_t>(),
                        r->spanPlanes.as<uint4>(), r->depthBits.as<uint32_t>(), sc->count, bounds, grid,
                        r->keysA.as<uint64_t>(), r->valuesA.as<uint32_t>(), control, stream);
        GSR_HIP(hipGetLastError());
        GSR_HIP(mark(4));
        which = launchRadixSort(work, r->keysA.as<uint64_t>(), r->valuesA.as<uint32_t>(), r->keysB.as<uint64_t>(),
                                r->valuesB.as<uint32_t>(), r->capacity, &control->entries, 0u, 32 + tileBits,
                                &r->sortEpoch, stream);
        GSR_HIP(hipGetLastError());
        GSR_HIP(mark(5));
        launchRanges(which ? r->keysB.as<uint64_t>() : r->keysA.as<uint64_t>(), control, r->capacity, tiles,
                     r->ranges.as<uint2>(), reinterpret_cast<uint32_t*>(superStats), superWords, stream);
        sortedValues = which ? r->valuesB.as<uint32_t>() : r->valuesA.as<uint32_t>();
    }
    GSR_HIP(hipGetLastError());
    GSR_HIP(mark(6));
    if (deferred)
        *deferred = TileJob{r->splats.as<float4>(), r->packed.as<uint32_t>(), sortedValues, r->ranges.as<uint2>(),
                            cam->width, cam->height, binsX, sc->count, outImage, nullptr};
    else if (r->settings.differentiable)
        launchRenderTrain(r->splats.as<float4>(), sortedValues, r->ranges.as<uint2>(), dc,
                          static_cast<float4*>(outImage), r->pixelState.as<uint2>(), stream);
    else {
        uint32_t* tileStats = nullptr;
        if (r->settings.tile_stats) {
            tileStats = r->tileStats.as<uint32_t>();
            GSR_HIP(hipMemsetAsync(tileStats, 0, size_t(tiles) * kTileStatCount * 4, stream));
        }
        launchRender(r->splats.as<float4>(), r->packed.as<uint32_t>(), r->settings.render_splat, sortedValues,
                     r->ranges.as<uint2>(), dc, sc->count, r->settings.format,
                     r->settings.render_pixels_per_thread, r->settings.render_exact_alpha != 0,
                     r->settings.render_loop, int(r->settings.sort_window), r->rayDepth.as<float4>(), outImage,
                     tileStats, int(r->settings.render_batch), stream, fuseSort ? &fused : nullptr);
    }
    GSR_HIP(hipGetLastError());
    GSR_HIP(mark(7));
    r->marksRecorded = t_kernelMarks != nullptr;
    t_kernelMarks = nullptr;
    if (r->statHost) {  // this frame's control block, read back without a stall
        int slot = -1;
        for (int s = 0; s < gsr_renderer::kStatSlots && slot < 0; ++s)
            if (!r->statPending[s]) slot = s;
        if (slot < 0) {  // every slot in flight: wait for the oldest
            uint64_t oldest = ~uint64_t(0);
            for (int s = 0; s < gsr_renderer::kStatSlots; ++s)
                if (r->statFrame[s] < oldest) oldest = r->statFrame[s], slot = s;
            GSR_HIP(hipEventSynchronize(r->statEvent[slot]));
            collectStats(r);
        }
        GSR_HIP(hipMemcpyAsync(&r->statHost[slot], control, sizeof(gsr_control), hipMemcpyDeviceToHost, stream));
        GSR_HIP(hipEventRecord(r->statEvent[slot], stream));
        r->statFrame[slot] = r->frameIndex;
        r->statPending[slot] = true;
    }
    if (!capturing) GSR_HIP(hipEventRecord(r->order, stream));
    ++r->frameIndex;
    r->lastCam = *cam;
    r->haveFrame = true;
    r->sortedIn = which;
    r->lastStream = stream;
    r->eventsRecorded = prof;
    return GSR_OK;
}

gsr_status gsr_render(gsr_renderer* r, const gsr_camera* cam, void* outImage, void* streamArg) {
    return renderFrame(r, cam, outImage, static_cast<hipStream_t>(streamArg), nullptr);
}

gsr_status gsr_read_control(gsr_renderer* r, gsr_control* out) {
    if (!r || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_control: null argument");
    GSR_HIP(hipSetDevice(r->ctx->device));
    GSR_HIP(hipDeviceSynchronize());
    GSR_HIP(hipMemcpy(out, r->control.ptr, sizeof(gsr_control), hipMemcpyDeviceToHost));
    return GSR_OK;
}

gsr_status gsr_read_lists(gsr_renderer* r, uint32_t* tileKeys, uint32_t* depthKeys, uint32_t* values,
                          uint32_t* ranges, uint32_t maxEntries) {
    if (!r) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_lists: null renderer");
    if (!r->haveFrame) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_lists: no frame rendered yet");
    gsr_control c;
    gsr_status st = gsr_read_control(r, &c);
    if (st != GSR_OK) return st;
    const uint32_t n = std::min(c.entries, maxEntries);
    const void* keySrc = r->sortedIn ? r->keysB.ptr : r->keysA.ptr;
    const void* valueSrc = r->sortedIn ? r->valuesB.ptr : r->valuesA.ptr;
    std::vector<uint32_t> vals;
    if (n > 0 && (values || (depthKeys && r->tileKeys32()))) {
        vals.resize(n);
        GSR_HIP(hipMemcpy(vals.data(), valueSrc, size_t(n) * 4, hipMemcpyDeviceToHost));
        if (values) std::memcpy(values, vals.data(), size_t(n) * 4);
    }
    if (n > 0 && r->tileKeys32()) {
        if (tileKeys) GSR_HIP(hipMemcpy(tileKeys, keySrc, size_t(n) * 4, hipMemcpyDeviceToHost));
        if (depthKeys) {  // 32-bit tile keys carry no depth: it is the listed splat's depth bits
            std::vector<float4> sp(size_t(r->scene->count) * 3);
            GSR_HIP(hipMemcpy(sp.data(), r->splats.ptr, sp.size() * sizeof(float4), hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < n; ++i) {
                const float d = vals[i] < r->scene->count ? sp[size_t(vals[i]) * 3].z : 0.0f;
                std::memcpy(&depthKeys[i], &d, 4);
            }
        }
    } else if (n > 0 && (tileKeys || depthKeys)) {
        std::vector<uint64_t> k(n);
        GSR_HIP(hipMemcpy(k.data(), keySrc, size_t(n) * 8, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            if (tileKeys) tileKeys[i] = uint32_t(k[i] >> 32);
            if (depthKeys) depthKeys[i] = uint32_t(k[i]);
        }
    }
    if (ranges)
        GSR_HIP(hipMemcpy(ranges, r->ranges.ptr, size_t(r->binsX(r->lastCam)) * r->binsY(r->lastCam) * 8,
                          hipMemcpyDeviceToHost));
    return GSR_OK;
}

gsr_status gsr_read_splats(gsr_renderer* r, gsr_projected_splat* splats, uint32_t* tileCounts) {
    if (!r) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_splats: null renderer");
    if (!r->haveFrame) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_splats: no frame rendered yet");
    GSR_HIP(hipSetDevice(r->ctx->device));
    GSR_HIP(hipDeviceSynchronize());
    const size_t n = r->scene->count;
    if (n && splats) GSR_HIP(hipMemcpy(splats, r->splats.ptr, n * 48, hipMemcpyDeviceToHost));
    if (n && tileCounts) GSR_HIP(hipMemcpy(tileCounts, r->tileCounts.ptr, n * 4, hipMemcpyDeviceToHost));
    return GSR_OK;
}

gsr_status gsr_read_packed_splats(gsr_renderer* r, uint32_t* words) {
    if (!r || !words) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_packed_splats: null argument");
    if (r->settings.render_splat == GSR_RENDER_SPLAT_FULL)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_packed_splats: render_splat is FULL (no packed records)");
    if (!r->haveFrame) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_packed_splats: no frame rendered yet");
    GSR_HIP(hipSetDevice(r->ctx->device));
    GSR_HIP(hipDeviceSynchronize());
    const size_t bytes = size_t(r->scene->count) * (r->settings.render_splat == GSR_RENDER_SPLAT_P32 ? 32 : 24);
    if (bytes) GSR_HIP(hipMemcpy(words, r->packed.ptr, bytes, hipMemcpyDeviceToHost));
    return GSR_OK;
}

gsr_status gsr_read_stage_times(gsr_renderer* r, float* ms, uint32_t n) {
    if (!r || !ms) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_stage_times: null argument");
    if (!r->eventsRecorded) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_stage_times: no profiled frame (profile_stages = 0?)");
    GSR_HIP(hipSetDevice(r->ctx->device));
    GSR_HIP(hipEventSynchronize(r->events[GSR_STAGE_COUNT - 1]));
    float t[GSR_STAGE_COUNT];
    for (int i = 0; i < GSR_STAGE_COUNT - 1; ++i) GSR_HIP(hipEventElapsedTime(&t[i], r->events[i], r->events[i + 1]));
    GSR_HIP(hipEventElapsedTime(&t[GSR_STAGE_TOTAL], r->events[0], r->events[GSR_STAGE_COUNT - 1]));
    for (uint32_t i = 0; i < n && i < uint32_t(GSR_STAGE_COUNT); ++i) ms[i] = t[i];
    return GSR_OK;
}

gsr_status gsr_read_kernel_times(gsr_renderer* r, const char** names, float* ms, uint32_t max, uint32_t* count) {
    if (!r) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_kernel_times: null renderer");
    if (!r->marks || !r->marksRecorded)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_kernel_times: no frame with kernel marks (profile_stages = 2?)");
    GSR_HIP(hipSetDevice(r->ctx->device));
    const KernelMarks& m = *r->marks;
    if (m.count > 0) GSR_HIP(hipEventSynchronize(m.events[m.count]));
    for (int i = 0; i < m.count && uint32_t(i) < max; ++i) {
        if (names) names[i] = m.names[i];
        if (ms) GSR_HIP(hipEventElapsedTime(&ms[i], m.events[i], m.events[i + 1]));
    }
    if (count) *count = uint32_t(m.count);
    return GSR_OK;
}

gsr_status gsr_read_tile_stats(gsr_renderer* r, uint32_t* stats, uint32_t maxTiles) {
    if (!r || !stats) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_tile_stats: null argument");
    if (!r->settings.tile_stats)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_tile_stats: the renderer was created without tile_stats");
    if (!r->haveFrame) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_read_tile_stats: no frame rendered yet");
    GSR_HIP(hipSetDevice(r->ctx->device));
    GSR_HIP(hipDeviceSynchronize());
    const uint32_t tiles = std::min(r->binsX(r->lastCam) * r->binsY(r->lastCam), maxTiles);
    if (tiles) GSR_HIP(hipMemcpy(stats, r->tileStats.ptr, size_t(tiles) * kTileStatCount * 4, hipMemcpyDeviceToHost));
    return GSR_OK;
}

// --- Stand-alone sort and scan -------------------------------------------------------------------

gsr_status gsr_sort_pairs(gsr_context* ctx, uint64_t* keys, uint32_t* values, uint32_t maxCount,
                          const uint32_t* countDevice, uint32_t keyBits, gsr_sort_algorithm algorithm,
                          void* streamArg) {
    if (!ctx || (maxCount && (!keys || !values))) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_sort_pairs: null argument");
    if (keyBits == 0 || keyBits > 64) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_sort_pairs: key_bits %u", keyBits);
    if (algorithm != GSR_SORT_REDUCE_THEN_SCAN && algorithm != GSR_SORT_ONESWEEP)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_sort_pairs: unknown algorithm %d", int(algorithm));
    if (maxCount == 0) return GSR_OK;
    GSR_HIP(hipSetDevice(ctx->device));
    hipStream_t stream = streamArg ? static_cast<hipStream_t>(streamArg) : ctx->stream;
    const size_t n = maxCount;
    const size_t smallBytes = (kMaxSortPasses * kRadix + kMaxSortPasses + 1 + size_t(kRadix) * kMaxSortBlocks) * 4;
    const size_t need = n * 8 + n * 4 + smallBytes + 256;
    if (ctx->work.bytes < need) {
        GSR_HIP(hipDeviceSynchronize());
        GSR_HIP(ctx->work.allocate(need));
    }
    if (ctx->sortStatus.bytes < sortStatusBytes(maxCount)) {
        GSR_HIP(hipDeviceSynchronize());
        GSR_HIP(ctx->sortStatus.allocate(sortStatusBytes(maxCount)));
        GSR_HIP(hipMemset(ctx->sortStatus.ptr, 0, ctx->sortStatus.bytes));
        ctx->sortEpoch = 0;
    }
    char* base = ctx->work.as<char>();
    uint64_t* keysB = reinterpret_cast<uint64_t*>(base);
    uint32_t* valuesB = reinterpret_cast<uint32_t*>(base + n * 8);
    uint32_t* small = reinterpret_cast<uint32_t*>(base + n * 12);
    SortWork work{int32_t(algorithm), small, small + kMaxSortPasses * kRadix, ctx->sortStatus.as<uint64_t>(),
                  small + kMaxSortPasses * kRadix + kMaxSortPasses,
                  small + kMaxSortPasses * kRadix + kMaxSortPasses + 1};
    const int which = launchRadixSort(work, keys, values, keysB, valuesB, maxCount, countDevice, 0u, keyBits,
                                      &ctx->sortEpoch, stream);
    GSR_HIP(hipGetLastError());
    if (which) {  // odd pass count: the first `count` pairs are in the scratch copy
        launchCopyPairs(keysB, valuesB, keys, values, countDevice, maxCount, stream);
        GSR_HIP(hipGetLastError());
    }
    return GSR_OK;
}

gsr_status gsr_sort_pairs_noalloc(gsr_context* ctx, uint64_t* keys, uint32_t* values, uint32_t maxCount,
                                  const uint32_t* countDevice, uint32_t keyBits, void* streamArg) {
    if (!ctx || (maxCount && (!keys || !values)))
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_sort_pairs_noalloc: null argument");
    if (keyBits == 0 || keyBits > 64) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_sort_pairs_noalloc: key_bits %u", keyBits);
    if (maxCount > kNoAllocSortMaxKeys)
        return fail(GSR_ERROR_INVALID_ARGUMENT,
                    "gsr_sort_pairs_noalloc: max_count %u exceeds the fixed scratch of %u keys (use gsr_sort_pairs)",
                    maxCount, kNoAllocSortMaxKeys);
    if (maxCount == 0) return GSR_OK;
    if (!ctx->noAllocScratch) return fail(GSR_ERROR_HIP, "gsr_sort_pairs_noalloc: the device scratch was not resolved");
    GSR_HIP(hipSetDevice(ctx->device));
    hipStream_t stream = streamArg ? static_cast<hipStream_t>(streamArg) : ctx->stream;
    char* base = static_cast<char*>(ctx->noAllocScratch);
    uint64_t* keysB = reinterpret_cast<uint64_t*>(base);
    uint32_t* valuesB = reinterpret_cast<uint32_t*>(base + kNoAllocValuesOffset);
    uint32_t* small = reinterpret_cast<uint32_t*>(base + kNoAllocWorkOffset);
    SortWork work{kSortAlgReduceThenScan, small, nullptr, nullptr, nullptr,
                  small + kMaxSortPasses * kRadix + kMaxSortPasses + 1};
    uint32_t epoch = 0;  // unused by the reduce-then-scan passes
    const int which = launchRadixSort(work, keys, values, keysB, valuesB, maxCount, countDevice, 0u, keyBits, &epoch,
                                      stream);
    GSR_HIP(hipGetLastError());
    if (which) {  // odd pass count: the first `count` pairs are in the scratch copy
        launchCopyPairs(keysB, valuesB, keys, values, countDevice, maxCount, stream);
        GSR_HIP(hipGetLastError());
    }
    return GSR_OK;
}

gsr_status gsr_exclusive_scan(gsr_context* ctx, const uint32_t* in, uint32_t* out, uint32_t count,
                              uint32_t* totalDevice, void* streamArg) {
    if (!ctx || (count && (!in || !out))) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_exclusive_scan: null argument");
    GSR_HIP(hipSetDevice(ctx->device));
    hipStream_t stream = streamArg ? static_cast<hipStream_t>(streamArg) : ctx->stream;
    const size_t blocks = (size_t(count) + kScanBlock - 1) / kScanBlock;
    const size_t need = (blocks + 1) * 4;
    if (ctx->work.bytes < need) {
        GSR_HIP(hipDeviceSynchronize());
        GSR_HIP(ctx->work.allocate(need));
    }
    launchExclusiveScan(in, out, count, ctx->work.as<uint32_t>(), totalDevice, stream);
    GSR_HIP(hipGetLastError());
    return GSR_OK;
}

}  // extern "C"

// --- Batched multi-view -------------------------------------------------------------------------
// Views go round-robin to `branches` streams, each with its own renderer; the branches are forked
// from the caller's stream and joined back with events, so the batch is stream-ordered like one
// gsr_render, and the same calls on a capturing stream record one graph with parallel branches.
// Small views do not fill 256 CUs (a 640x360 view is 920 tiles); side by side they do.
struct gsr_batch {
    gsr_context* ctx = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    hipStream_t captureStream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // What the graph was recorded against: each branch renderer, its id and generation, and
    // whether its frames ran the learned-colour pass. A replay refuses when any of them changed.
    struct Captured {
        const gsr_renderer* renderer;
        uint64_t id, generation;
        const gsr_colour_model* colour;
        bool colourPass;
    };
    std::vector<Captured> captured;
    // Persistent mode: one render launch over every view's tiles, after every view's sort. Its
    // table (a TileJob per view, the views' first tiles, the work counter) goes to the device by
    // a stream-ordered copy from pinned host memory: two tables alternate for eager batches (a
    // table is refilled only once its previous copy has run), and a captured graph has its own,
    // which stays as captured.
    bool persistent = false;
    uint32_t workgroups = 0;
    struct Table {
        DeviceBuffer dev;
        void* host = nullptr;
        size_t bytes = 0;
        hipEvent_t copied = nullptr;
        bool pending = false;
    };
    Table eager[2], capturedTable;
    int nextEager = 0;
};

namespace {

void releaseGraph(gsr_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->captured.clear();
}

bool colourPassOn(const gsr_renderer* r) { return r->colour && r->colour->mlpCount > 0; }

// A renderer whose frames the persistent render can draw: the default culled loop on 16-px tiles.
bool persistentReady(const gsr_renderer* r) {
    const gsr_settings& s = r->settings;
    return r->binSize == 16 && s.render_loop == GSR_RENDER_LOOP_CULLED && s.render_pixels_per_thread == 1 &&
           !s.render_exact_alpha && s.sort_window == 0 && !s.differentiable && s.render_batch == 256 && !s.tile_stats;
}

hipError_t tableReserve(gsr_batch::Table& t, size_t bytes) {
    if (t.bytes >= bytes) return hipSuccess;
    if (t.pending) {
        const hipError_t e = hipEventSynchronize(t.copied);
        if (e != hipSuccess) return e;
        t.pending = false;
    }
    if (t.host) (void)hipHostFree(t.host);
    t.host = nullptr;
    t.bytes = 0;
    hipError_t e = hipHostMalloc(&t.host, bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = t.dev.allocate(bytes);
    if (e == hipSuccess && !t.copied) e = hipEventCreateWithFlags(&t.copied, hipEventDisableTiming);
    if (e == hipSuccess) t.bytes = bytes;
    return e;
}

void tableRelease(gsr_batch::Table& t) {
    if (t.host) (void)hipHostFree(t.host);
    if (t.copied) (void)hipEventDestroy(t.copied);
    t.host = nullptr;
    t.copied = nullptr;
    t.dev.release();
    t.bytes = 0;
}

gsr_status checkBatchArgs(gsr_batch* b, gsr_renderer* const* renderers, uint32_t viewCount, const gsr_camera* cams,
                          void* const* outImages, const char* what) {
    if (!b || !renderers || (viewCount && (!cams || !outImages)))
        return fail(GSR_ERROR_INVALID_ARGUMENT, "%s: null argument", what);
    const uint32_t used = b->persistent ? viewCount : std::min<uint32_t>(viewCount, uint32_t(b->streams.size()));
    for (uint32_t i = 0; i < used; ++i) {
        if (!renderers[i]) return fail(GSR_ERROR_INVALID_ARGUMENT, "%s: renderer %u is null", what, i);
        if (b->persistent && !persistentReady(renderers[i]))
            return fail(GSR_ERROR_UNSUPPORTED,
                        "%s: renderer %u cannot join a persistent render (it needs the default culled loop on "
                        "16-px tiles: no StopThePop, differentiable, exact alpha, tile_stats or other batch depth)",
                        what, i);
        if (b->persistent && (renderers[i]->settings.format != renderers[0]->settings.format ||
                              renderers[i]->settings.render_splat != renderers[0]->settings.render_splat))
            return fail(GSR_ERROR_INVALID_ARGUMENT, "%s: a persistent render takes one image format and record", what);
        if (renderers[i]->ctx != b->ctx)
            return fail(GSR_ERROR_INVALID_ARGUMENT, "%s: renderer %u belongs to another context", what, i);
        for (uint32_t j = 0; j < i; ++j)
            if (renderers[j] == renderers[i])
                return fail(GSR_ERROR_INVALID_ARGUMENT, "%s: renderers %u and %u are the same (one per %s)", what, j, i,
                            b->persistent ? "view in persistent mode" : "branch");
    }
    return GSR_OK;
}

// Fork, render view v on branch v % branches, join. Every fork is joined, also after a failure.
gsr_status batchRecord(gsr_batch* b, gsr_renderer* const* renderers, uint32_t viewCount, const gsr_camera* cams,
                       void* const* outImages, hipStream_t stream) {
    const uint32_t branches = std::min<uint32_t>(viewCount, uint32_t(b->streams.size()));
    if (branches == 0) return GSR_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GSR_HIP(hipStreamIsCapturing(stream, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    gsr_batch::Table* table = nullptr;
    if (b->persistent) {  // the table this call fills (an eager table waits for its previous copy)
        table = capturing ? &b->capturedTable : &b->eager[b->nextEager];
        if (!capturing) b->nextEager ^= 1;
        const size_t bytes = size_t(viewCount) * sizeof(TileJob) + (size_t(viewCount) + 2) * 4;
        if (capturing && table->bytes < bytes)
            return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_batch_capture: the persistent table was not reserved");
        if (table->pending) {
            GSR_HIP(hipEventSynchronize(table->copied));
            table->pending = false;
        }
        GSR_HIP(tableReserve(*table, bytes));
    }
    std::vector<TileJob> jobs(b->persistent ? viewCount : 0);
    GSR_HIP(hipEventRecord(b->fork, stream));
    for (uint32_t i = 0; i < branches; ++i) GSR_HIP(hipStreamWaitEvent(b->streams[i], b->fork, 0));
    gsr_status st = GSR_OK;
    for (uint32_t v = 0; v < viewCount && st == GSR_OK; ++v)
        st = b->persistent ? renderFrame(renderers[v], &cams[v], outImages[v], b->streams[v % branches], &jobs[v])
                           : gsr_render(renderers[v % branches], &cams[v], outImages[v], b->streams[v % branches]);
    const std::string err = st == GSR_OK ? std::string() : g_lastError;
    for (uint32_t i = 0; i < branches; ++i) {
        GSR_HIP(hipEventRecord(b->joins[i], b->streams[i]));
        GSR_HIP(hipStreamWaitEvent(stream, b->joins[i], 0));
    }
    if (st != GSR_OK) {
        g_lastError = err;
        return st;
    }
    if (b->persistent) {
        // Every view's lists are complete: the table, then one render over all views' tiles.
        char* host = static_cast<char*>(table->host);
        std::memcpy(host, jobs.data(), jobs.size() * sizeof(TileJob));
        uint32_t* tileStart = reinterpret_cast<uint32_t*>(host + jobs.size() * sizeof(TileJob));
        tileStart[0] = 0;
        for (uint32_t v = 0; v < viewCount; ++v) tileStart[v + 1] = tileStart[v] + cams[v].tiles_x * cams[v].tiles_y;
        const size_t bytes = jobs.size() * sizeof(TileJob) + (size_t(viewCount) + 2) * 4;
        GSR_HIP(hipMemcpyAsync(table->dev.ptr, host, bytes, hipMemcpyHostToDevice, stream));
        if (!capturing) {
            GSR_HIP(hipEventRecord(table->copied, stream));
            table->pending = true;
        }
        char* dev = table->dev.as<char>();
        const uint32_t* dStart = reinterpret_cast<const uint32_t*>(dev + jobs.size() * sizeof(TileJob));
        launchRenderViews(reinterpret_cast<const TileJob*>(dev), dStart, viewCount,
                          std::min(b->workgroups, tileStart[viewCount]), const_cast<uint32_t*>(dStart) + viewCount + 1,
                          renderers[0]->settings.format, renderers[0]->settings.render_splat, stream);
        GSR_HIP(hipGetLastError());
        // The views' renderers are in use until this render ends: later frames wait for it.
        if (!capturing)
            for (uint32_t v = 0; v < viewCount; ++v) {
                GSR_HIP(hipEventRecord(renderers[v]->order, stream));
                renderers[v]->lastStream = stream;
            }
    }
    return GSR_OK;
}

}  // namespace

extern "C" {

gsr_status gsr_batch_create(gsr_context* ctx, uint32_t branches, gsr_batch** out) {
    if (!ctx || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_batch_create: null argument");
    *out = nullptr;
    if (branches == 0 || branches > 16)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_batch_create: branches must be 1..16, got %u", branches);
    GSR_HIP(hipSetDevice(ctx->device));
    gsr_batch* b = new (std::nothrow) gsr_batch();
    if (!b) return fail(GSR_ERROR_OUT_OF_MEMORY, "gsr_batch_create: host allocation");
    b->ctx = ctx;
    hipError_t e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    for (uint32_t i = 0; i < branches && e == hipSuccess; ++i) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);  // default priority
        if (e == hipSuccess) b->streams.push_back(st);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->joins.push_back(ev);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->captureStream, hipStreamNonBlocking);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    b->workgroups = uint32_t(std::max(cus, 1)) * 8u;  // 8 workgroups of the culled render per CU
    if (e != hipSuccess) {
        gsr_batch_destroy(b);
        return hipFail(e, "gsr_batch_create");
    }
    *out = b;
    return GSR_OK;
}

void gsr_batch_destroy(gsr_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    releaseGraph(b);
    for (gsr_batch::Table& t : b->eager) tableRelease(t);
    tableRelease(b->capturedTable);
    for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
    for (hipEvent_t e : b->joins) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->captureStream) (void)hipStreamDestroy(b->captureStream);
    delete b;
}

uint32_t gsr_batch_branches(const gsr_batch* b) { return b ? uint32_t(b->streams.size()) : 0; }

gsr_status gsr_batch_set_persistent(gsr_batch* b, int32_t persistent) {
    if (!b) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_batch_set_persistent: null batch");
    if (b->persistent != (persistent != 0)) {
        GSR_HIP(hipSetDevice(b->ctx->device));
        GSR_HIP(hipDeviceSynchronize());
        releaseGraph(b);  // a graph of the other mode
        b->persistent = persistent != 0;
    }
    return GSR_OK;
}

gsr_status gsr_batch_render(gsr_batch* b, gsr_renderer* const* renderers, uint32_t viewCount, const gsr_camera* cams,
                            void* const* outImages, void* streamArg) {
    gsr_status st = checkBatchArgs(b, renderers, viewCount, cams, outImages, "gsr_batch_render");
    if (st != GSR_OK) return st;
    GSR_HIP(hipSetDevice(b->ctx->device));
    return batchRecord(b, renderers, viewCount, cams, outImages,
                       streamArg ? static_cast<hipStream_t>(streamArg) : b->ctx->stream);
}

gsr_status gsr_batch_capture(gsr_batch* b, gsr_renderer* const* renderers, uint32_t viewCount,
                             const gsr_camera* cams, void* const* outImages) {
    gsr_status st = checkBatchArgs(b, renderers, viewCount, cams, outImages, "gsr_batch_capture");
    if (st != GSR_OK) return st;
    const uint32_t used = b->persistent ? viewCount : std::min<uint32_t>(viewCount, uint32_t(b->streams.size()));
    for (uint32_t i = 0; i < used; ++i) {
        const gsr_renderer* r = renderers[i];
        if (r->statHost || r->settings.profile_stages || r->settings.tile_stats)
            return fail(GSR_ERROR_INVALID_ARGUMENT,
                        "gsr_batch_capture: renderer %u reads back statistics or stage times (adaptive_list, "
                        "frame_stats, profile_stages, tile_stats); a captured batch takes plain renderers", i);
        // The one-sweep sort tags its look-back status words with a host-side epoch, which a
        // graph would freeze: a replay could read the previous replay's words as current.
        if (!r->binning && r->settings.sort_algorithm == GSR_SORT_ONESWEEP)
            return fail(GSR_ERROR_INVALID_ARGUMENT,
                        "gsr_batch_capture: renderer %u uses the one-sweep sort, whose status epochs cannot be "
                        "captured; use the reduce-then-scan sort", i);
    }
    // Every view's size must have been rendered by its branch's renderer before (a new size
    // allocates and synchronises, which a capture cannot record).
    for (uint32_t v = 0; v < viewCount; ++v) {
        const gsr_renderer* r = renderers[v % used];
        const uint32_t tiles = r->binsX(cams[v]) * r->binsY(cams[v]);
        const bool orderReady =
            !r->tileSort || r->findOrder(r->binsX(cams[v]), r->binsY(cams[v])) != nullptr;
        if (tiles > r->rangesTiles || !orderReady || !r->haveFrame)
            return fail(GSR_ERROR_INVALID_ARGUMENT,
                        "gsr_batch_capture: view %u (%ux%u) is a size its renderer has not rendered yet: render the "
                        "batch once (gsr_batch_render) before capturing it", v, cams[v].width, cams[v].height);
    }
    GSR_HIP(hipSetDevice(b->ctx->device));
    GSR_HIP(hipDeviceSynchronize());  // replays of the previous graph may still run (and read its table)
    releaseGraph(b);
    if (b->persistent)  // allocated here: a capture cannot allocate
        GSR_HIP(tableReserve(b->capturedTable, size_t(viewCount) * sizeof(TileJob) + (size_t(viewCount) + 2) * 4));
    GSR_HIP(hipStreamBeginCapture(b->captureStream, hipStreamCaptureModeThreadLocal));
    st = batchRecord(b, renderers, viewCount, cams, outImages, b->captureStream);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b->captureStream, &graph);
    if (st != GSR_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hipFail(e, "gsr_batch_capture: hipStreamEndCapture");
    }
    b->graph = graph;
    const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        releaseGraph(b);
        return hipFail(ei, "gsr_batch_capture: hipGraphInstantiate");
    }
    for (uint32_t i = 0; i < used; ++i) {
        const gsr_renderer* r = renderers[i];
        b->captured.push_back({r, r->id, r->generation, r->colour, colourPassOn(r)});
    }
    return GSR_OK;
}

gsr_status gsr_batch_replay(gsr_batch* b, void* streamArg) {
    if (!b) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_batch_replay: null batch");
    if (!b->exec) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_batch_replay: nothing captured");
    // The graph holds raw pointers into each renderer: refuse (before launching anything) when a
    // renderer was destroyed, reallocated a buffer, or changed its colour binding since the capture.
    for (size_t i = 0; i < b->captured.size(); ++i) {
        const gsr_batch::Captured& c = b->captured[i];
        if (!rendererLive(c.renderer, c.id))
            return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_batch_replay: renderer %zu of the capture was destroyed", i);
        if (c.renderer->generation != c.generation || c.renderer->colour != c.colour ||
            colourPassOn(c.renderer) != c.colourPass)
            return fail(GSR_ERROR_INVALID_ARGUMENT,
                        "gsr_batch_replay: renderer %zu changed since the capture (a larger view reallocated its "
                        "buffers, or its colour model or fallback marks changed): capture the batch again", i);
    }
    GSR_HIP(hipSetDevice(b->ctx->device));
    GSR_HIP(hipGraphLaunch(b->exec, streamArg ? static_cast<hipStream_t>(streamArg) : b->ctx->stream));
    return GSR_OK;
}

/* --- Learned colour --------------------------------------------------------------------------- */

uint32_t gsr_colour_model_param_count(void) { return uint32_t(mlpParamCount()); }

}  // extern "C"

namespace {

// Every model call starts after all work on the device (frames in flight may read the model) and
// ends with its stream drained.
hipError_t modelBegin(gsr_colour_model* m) {
    hipError_t e = hipSetDevice(m->ctx->device);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e;
}

SceneRef modelScene(const gsr_colour_model* m) {
    SceneRef ref = m->scene->ref;
    ref.degree = m->scene->shDegree;  // the target: the scene's full SH colour
    return ref;
}

// After a parameter or latent change: the bf16 weight tiles and latents, and no gaussian marked
// until the next fallback pass.
hipError_t modelRefresh(gsr_colour_model* m, hipStream_t stream) {
    launchMlpPack(m->params.as<float>(), m->packs.as<uint2>(), stream);
    launchToBf16(m->latents.as<float>(), m->latentsBf16.as<uint16_t>(), m->count * 16, stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemsetAsync(m->useMlp.ptr, 0, std::max<size_t>(m->count, 1), stream);
    m->mlpCount = 0;
    return e;
}

// The per-workgroup gradient rows of a training pass of `samples` (grown on use; the caller has
// drained the device).
hipError_t reservePartials(gsr_colour_model* m, uint32_t samples) {
    const size_t bytes = size_t(mlpTrainBlocks(samples)) * size_t(mlpParamCount()) * 4;
    return m->partials.bytes >= bytes ? hipSuccess : m->partials.allocate(bytes);
}

// Eigen-decomposition of the symmetric n x n matrix a (row-major; destroyed) by cyclic Jacobi
// rotations: eigenvalues in w, eigenvectors in the columns of v.
void symmetricEigen(std::vector<double>& a, int n, std::vector<double>& w, std::vector<double>& v) {
    v.assign(size_t(n) * n, 0.0);
    for (int i = 0; i < n; ++i) v[size_t(i) * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) (i == j ? diag : off) += a[size_t(i) * n + j] * a[size_t(i) * n + j];
        if (off <= 1e-24 * (diag + 1e-300)) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[size_t(p) * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (a[size_t(q) * n + q] - a[size_t(p) * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = a[size_t(k) * n + p], akq = a[size_t(k) * n + q];
                    a[size_t(k) * n + p] = c * akp - s * akq;
                    a[size_t(k) * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = a[size_t(p) * n + k], aqk = a[size_t(q) * n + k];
                    a[size_t(p) * n + k] = c * apk - s * aqk;
                    a[size_t(q) * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[size_t(k) * n + p], vkq = v[size_t(k) * n + q];
                    v[size_t(k) * n + p] = c * vkp - s * vkq;
                    v[size_t(k) * n + q] = s * vkp + c * vkq;
                }
            }
    }
    w.resize(size_t(n));
    for (int i = 0; i < n; ++i) w[size_t(i)] = a[size_t(i) * n + i];
}

// The latents' start: every gaussian's SH values (D = 3 (degree + 1)^2 of them) projected on the
// scene's leading principal components (up to 16; each scaled to standard deviation 0.5, the
// others 0), so that a latent starts out carrying its gaussian's colour data and training only
// has to shape the shared network around it.
hipError_t initLatentsPca(gsr_colour_model* m, hipStream_t stream) {
    const SceneRef ref = modelScene(m);
    const int D = 3 * (ref.degree + 1) * (ref.degree + 1), pairs = D * (D + 1) / 2;
    const uint32_t blocks = shMomentBlocks(m->count);
    DeviceBuffer part, projMu;
    hipError_t e = part.allocate(size_t(blocks) * size_t(D + pairs) * 4);
    if (e != hipSuccess) return e;
    launchShMoments(ref, part.as<float>(), stream);
    std::vector<float> rows(size_t(blocks) * size_t(D + pairs));
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(rows.data(), part.ptr, rows.size() * 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    std::vector<double> mean(size_t(D), 0.0), cov(size_t(D) * D, 0.0);
    for (uint32_t b = 0; b < blocks; ++b) {
        const float* r = rows.data() + size_t(b) * size_t(D + pairs);
        for (int i = 0; i < D; ++i) mean[size_t(i)] += r[i];
        int p = D;
        for (int i = 0; i < D; ++i)
            for (int j = i; j < D; ++j) cov[size_t(i) * D + j] += r[p++];
    }
    const double n = double(m->count);
    for (int i = 0; i < D; ++i) mean[size_t(i)] /= n;
    for (int i = 0; i < D; ++i)
        for (int j = i; j < D; ++j) {
            const double c = cov[size_t(i) * D + j] / n - mean[size_t(i)] * mean[size_t(j)];
            cov[size_t(i) * D + j] = cov[size_t(j) * D + i] = c;
        }
    std::vector<double> w, v;
    symmetricEigen(cov, D, w, v);
    std::vector<int> order(static_cast<size_t>(D));
    for (int i = 0; i < D; ++i) order[size_t(i)] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[size_t(a)] > w[size_t(b)]; });
    std::vector<float> pm(16 * 48 + 48, 0.0f);
    const double top = D > 0 ? std::max(w[size_t(order[0])], 0.0) : 0.0;
    for (int r = 0; r < std::min(16, D); ++r) {
        const double lam = w[size_t(order[size_t(r)])];
        // Components below 1e-6 of the largest are the moments' fp32 rounding: left at 0.
        if (!(lam > 1e-6 * top) || !(lam > 0.0)) continue;
        const double s = 0.5 / std::sqrt(lam);
        for (int f = 0; f < D; ++f) pm[size_t(r) * 48 + f] = float(v[size_t(f) * D + order[size_t(r)]] * s);
    }
    for (int f = 0; f < D; ++f) pm[16 * 48 + size_t(f)] = float(mean[size_t(f)]);
    e = projMu.allocate(pm.size() * 4);
    if (e == hipSuccess) e = hipMemcpyAsync(projMu.ptr, pm.data(), pm.size() * 4, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    launchInitLatentsPca(ref, projMu.as<float>(), m->latents.as<float>(), stream);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);  // the temporary buffers are freed on return
    return e;
}

uint32_t trainSeed(const gsr_colour_model* m, uint32_t index) {
    uint32_t x = m->seed ^ (index * 0x9E3779B9u + 0x632BE5ABu);
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    return x;
}

}  // namespace

extern "C" {

gsr_status gsr_colour_model_create(gsr_context* ctx, const gsr_scene* scene, uint32_t seed, gsr_colour_model** out) {
    if (!ctx || !scene || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_create: null argument");
    *out = nullptr;
    if (scene->count == 0) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_create: empty scene");
    if (scene->count > (0xFFFFFFFFu / 16u))
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_create: %u gaussians (at most %u)", scene->count,
                    0xFFFFFFFFu / 16u);
    GSR_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<gsr_colour_model> m(new (std::nothrow) gsr_colour_model());
    if (!m) return fail(GSR_ERROR_OUT_OF_MEMORY, "gsr_colour_model_create: host allocation");
    m->ctx = ctx;
    m->scene = scene;
    m->count = scene->count;
    m->seed = seed;
    const size_t P = size_t(mlpParamCount()), L = size_t(scene->count) * 16;
    hipError_t e = hipSuccess;
    for (DeviceBuffer* b : {&m->params, &m->gradParams, &m->mParams, &m->vParams}) {
        if (e == hipSuccess) e = b->allocate(P * 4);
        if (e == hipSuccess) e = hipMemset(b->ptr, 0, b->bytes);
    }
    for (DeviceBuffer* b : {&m->latents, &m->gradLatents, &m->mLatents, &m->vLatents}) {
        if (e == hipSuccess) e = b->allocate(L * 4);
        if (e == hipSuccess) e = hipMemset(b->ptr, 0, b->bytes);
    }
    if (e == hipSuccess) e = m->latentsBf16.allocate(L * 2);
    if (e == hipSuccess) e = m->packs.allocate(size_t(mlpPackedTiles()) * 64 * 8);
    if (e == hipSuccess) e = m->useMlp.allocate(scene->count);
    if (e == hipSuccess) e = m->errors.allocate(size_t(scene->count) * 4);
    if (e == hipSuccess) e = m->sse.allocate(4);
    if (e != hipSuccess) return hipFail(e, "gsr_colour_model_create: allocation");
    // He-normal weights (fan-in 19 for layer 1: latent + direction; 64 for layers 2 and 3, the
    // last scaled by 0.5), zero biases; latents from the SH's principal components (initLatentsPca).
    std::vector<float> p(P, 0.0f);
    uint64_t state = 0x853C49E6748FEA9Bull ^ (uint64_t(seed) << 1);
    auto uniform = [&]() {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return (double((state >> 11) & ((1ull << 53) - 1)) + 0.5) / double(1ull << 53);
    };
    auto normal = [&]() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); };
    const float s1 = std::sqrt(2.0f / 19.0f), s2 = std::sqrt(2.0f / 64.0f), s3 = 0.5f * std::sqrt(1.0f / 64.0f);
    for (int r = 0; r < 64; ++r)
        for (int c = 0; c < 19; ++c) p[size_t(r) * 32 + c] = s1 * float(normal());
    const size_t w2 = 64 * 32, w3 = w2 + 64 * 64 + 64;
    for (size_t i = 0; i < 64 * 64; ++i) p[w2 + i] = s2 * float(normal());
    for (size_t i = 0; i < 3 * 64; ++i) p[w3 + i] = s3 * float(normal());
    hipStream_t stream = ctx->stream;
    if ((e = hipMemcpy(m->params.ptr, p.data(), P * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return hipFail(e, "gsr_colour_model_create: upload");
    if ((e = initLatentsPca(m.get(), stream)) != hipSuccess) return hipFail(e, "gsr_colour_model_create: latents");
    if ((e = modelRefresh(m.get(), stream)) != hipSuccess) return hipFail(e, "gsr_colour_model_create: init");
    GSR_HIP(hipStreamSynchronize(stream));
    *out = m.release();
    return GSR_OK;
}

void gsr_colour_model_destroy(gsr_colour_model* m) {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipDeviceSynchronize();
    delete m;
}

gsr_status gsr_colour_model_train(gsr_colour_model* m, uint32_t steps, uint32_t samples, float lrWeights,
                                  float lrLatents, float* lastLoss) {
    if (!m) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_train: null model");
    if (samples == 0 || samples % 256 != 0)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_train: samples_per_step %u (a positive multiple of 256)",
                    samples);
    if (!(lrWeights >= 0.0f) || !(lrLatents >= 0.0f))
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_train: negative or NaN learning rate");
    GSR_HIP(modelBegin(m));
    GSR_HIP(reservePartials(m, samples));
    hipStream_t stream = m->ctx->stream;
    const SceneRef ref = modelScene(m);
    const uint32_t P = uint32_t(mlpParamCount()), L = m->count * 16;
    // A step of at most `count` samples touches distinct gaussians: their latents take their Adam
    // step inside the training pass (lazy Adam: untouched latents and moments keep their values).
    // Larger steps (small scenes) add the latent gradients up and take a dense Adam step.
    const bool lazy = samples <= m->count;
    for (uint32_t i = 0; i < steps; ++i) {
        ++m->step;
        GSR_HIP(hipMemsetAsync(m->sse.ptr, 0, 4, stream));
        launchMlpTrain(ref, m->params.as<float>(), m->packs.as<uint2>(), m->latents.as<float>(),
                       m->gradParams.as<float>(), m->partials.as<float>(), m->gradLatents.as<float>(),
                       m->mLatents.as<float>(), m->vLatents.as<float>(), lrLatents, m->step, m->sse.as<float>(),
                       sampleStream(trainSeed(m, m->streamIndex++), m->count), samples,
                       lazy ? kTrainStep : kTrainGrad, stream);
        launchAdam(m->params.as<float>(), m->gradParams.as<float>(), m->mParams.as<float>(), m->vParams.as<float>(), P,
                   lrWeights, m->step, stream);
        if (!lazy)
            launchAdam(m->latents.as<float>(), m->gradLatents.as<float>(), m->mLatents.as<float>(),
                       m->vLatents.as<float>(), L, lrLatents, m->step, stream);
        launchMlpPack(m->params.as<float>(), m->packs.as<uint2>(), stream);
        GSR_HIP(hipGetLastError());
    }
    GSR_HIP(modelRefresh(m, stream));
    float sse = 0.0f;
    GSR_HIP(hipMemcpyAsync(&sse, m->sse.ptr, 4, hipMemcpyDeviceToHost, stream));
    GSR_HIP(hipStreamSynchronize(stream));
    if (lastLoss) *lastLoss = steps ? sse / (3.0f * float(samples)) : 0.0f;
    return GSR_OK;
}

gsr_status gsr_colour_model_loss(gsr_colour_model* m, uint32_t seed, uint32_t samples, float* loss) {
    if (!m || !loss) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_loss: null argument");
    if (samples == 0 || samples % 256 != 0)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_loss: samples %u (a positive multiple of 256)", samples);
    GSR_HIP(modelBegin(m));
    hipStream_t stream = m->ctx->stream;
    GSR_HIP(hipMemsetAsync(m->sse.ptr, 0, 4, stream));
    launchMlpTrain(modelScene(m), m->params.as<float>(), m->packs.as<uint2>(), m->latents.as<float>(), nullptr, nullptr,
                   nullptr, nullptr, nullptr, 0.0f, 1u, m->sse.as<float>(), sampleStream(seed, m->count), samples,
                   kTrainLoss, stream);
    GSR_HIP(hipGetLastError());
    float sse = 0.0f;
    GSR_HIP(hipMemcpyAsync(&sse, m->sse.ptr, 4, hipMemcpyDeviceToHost, stream));
    GSR_HIP(hipStreamSynchronize(stream));
    *loss = sse / (3.0f * float(samples));
    return GSR_OK;
}

gsr_status gsr_colour_model_gradients(gsr_colour_model* m, uint32_t seed, uint32_t samples, float* gradParams,
                                      float* gradLatents, float* loss, uint32_t* gaussians, float* dirs) {
    if (!m) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_gradients: null model");
    if (samples == 0 || samples % 256 != 0)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_gradients: samples %u (a positive multiple of 256)",
                    samples);
    GSR_HIP(modelBegin(m));
    GSR_HIP(reservePartials(m, samples));
    hipStream_t stream = m->ctx->stream;
    const size_t P = size_t(mlpParamCount()), L = size_t(m->count) * 16;
    // The latent gradients are zero between training steps (Adam clears them); clear them anyway.
    // The weight gradient is written whole by the pass.
    GSR_HIP(hipMemsetAsync(m->gradLatents.ptr, 0, L * 4, stream));
    GSR_HIP(hipMemsetAsync(m->sse.ptr, 0, 4, stream));
    const SampleStream st = sampleStream(seed, m->count);
    launchMlpTrain(modelScene(m), m->params.as<float>(), m->packs.as<uint2>(), m->latents.as<float>(),
                   m->gradParams.as<float>(), m->partials.as<float>(), m->gradLatents.as<float>(), nullptr, nullptr,
                   0.0f, 1u, m->sse.as<float>(), st, samples, kTrainGrad, stream);
    GSR_HIP(hipGetLastError());
    if (gradParams) GSR_HIP(hipMemcpyAsync(gradParams, m->gradParams.ptr, P * 4, hipMemcpyDeviceToHost, stream));
    if (gradLatents) GSR_HIP(hipMemcpyAsync(gradLatents, m->gradLatents.ptr, L * 4, hipMemcpyDeviceToHost, stream));
    float sse = 0.0f;
    GSR_HIP(hipMemcpyAsync(&sse, m->sse.ptr, 4, hipMemcpyDeviceToHost, stream));
    if (gaussians || dirs) {
        DeviceBuffer g, d;
        GSR_HIP(g.allocate(size_t(samples) * 4));
        GSR_HIP(d.allocate(size_t(samples) * 12));
        launchMlpSamples(st, 0, samples, g.as<uint32_t>(), d.as<float>(), stream);
        GSR_HIP(hipGetLastError());
        if (gaussians) GSR_HIP(hipMemcpyAsync(gaussians, g.ptr, size_t(samples) * 4, hipMemcpyDeviceToHost, stream));
        if (dirs) GSR_HIP(hipMemcpyAsync(dirs, d.ptr, size_t(samples) * 12, hipMemcpyDeviceToHost, stream));
        GSR_HIP(hipStreamSynchronize(stream));
    }
    GSR_HIP(hipMemsetAsync(m->gradParams.ptr, 0, P * 4, stream));
    GSR_HIP(hipMemsetAsync(m->gradLatents.ptr, 0, L * 4, stream));
    GSR_HIP(hipStreamSynchronize(stream));
    if (loss) *loss = sse / (3.0f * float(samples));
    return GSR_OK;
}

gsr_status gsr_colour_model_get(gsr_colour_model* m, float* params, float* latents) {
    if (!m) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_get: null model");
    GSR_HIP(modelBegin(m));
    if (params) GSR_HIP(hipMemcpy(params, m->params.ptr, size_t(mlpParamCount()) * 4, hipMemcpyDeviceToHost));
    if (latents) GSR_HIP(hipMemcpy(latents, m->latents.ptr, size_t(m->count) * 64, hipMemcpyDeviceToHost));
    return GSR_OK;
}

gsr_status gsr_colour_model_set(gsr_colour_model* m, const float* params, const float* latents) {
    if (!m) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_set: null model");
    GSR_HIP(modelBegin(m));
    hipStream_t stream = m->ctx->stream;
    if (params) GSR_HIP(hipMemcpy(m->params.ptr, params, size_t(mlpParamCount()) * 4, hipMemcpyHostToDevice));
    if (latents) GSR_HIP(hipMemcpy(m->latents.ptr, latents, size_t(m->count) * 64, hipMemcpyHostToDevice));
    // A new starting point: fresh Adam moments.
    for (DeviceBuffer* b : {&m->mParams, &m->vParams, &m->mLatents, &m->vLatents, &m->gradParams, &m->gradLatents})
        GSR_HIP(hipMemsetAsync(b->ptr, 0, b->bytes, stream));
    m->step = 0;
    GSR_HIP(modelRefresh(m, stream));
    GSR_HIP(hipStreamSynchronize(stream));
    return GSR_OK;
}

gsr_status gsr_colour_model_evaluate(gsr_colour_model* m, const uint32_t* gaussians, const float* dirs, uint32_t n,
                                     float* rgb) {
    if (!m || (n && (!gaussians || !dirs || !rgb)))
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_evaluate: null argument");
    for (uint32_t i = 0; i < n; ++i)
        if (gaussians[i] >= m->count)
            return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_evaluate: gaussian %u of %u", gaussians[i],
                        m->count);
    if (n == 0) return GSR_OK;
    GSR_HIP(modelBegin(m));
    hipStream_t stream = m->ctx->stream;
    DeviceBuffer g, d, out;
    GSR_HIP(g.allocate(size_t(n) * 4));
    GSR_HIP(d.allocate(size_t(n) * 12));
    GSR_HIP(out.allocate(size_t(n) * 12));
    GSR_HIP(hipMemcpy(g.ptr, gaussians, size_t(n) * 4, hipMemcpyHostToDevice));
    GSR_HIP(hipMemcpy(d.ptr, dirs, size_t(n) * 12, hipMemcpyHostToDevice));
    launchMlpEvaluate(m->params.as<float>(), m->packs.as<uint2>(), m->latents.as<float>(), g.as<uint32_t>(),
                      d.as<float>(), n, out.as<float>(), stream);
    GSR_HIP(hipGetLastError());
    GSR_HIP(hipMemcpyAsync(rgb, out.ptr, size_t(n) * 12, hipMemcpyDeviceToHost, stream));
    GSR_HIP(hipStreamSynchronize(stream));
    return GSR_OK;
}

gsr_status gsr_colour_model_fallback(gsr_colour_model* m, float threshold, uint32_t* mlpCount, float* errors) {
    if (!m) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_colour_model_fallback: null model");
    GSR_HIP(modelBegin(m));
    hipStream_t stream = m->ctx->stream;
    launchMlpFallback(modelScene(m), m->params.as<float>(), m->packs.as<uint2>(), m->latentsBf16.as<uint16_t>(),
                      threshold, m->useMlp.as<uint8_t>(), m->errors.as<float>(), stream);
    GSR_HIP(hipGetLastError());
    std::vector<uint8_t> marks(m->count);
    GSR_HIP(hipMemcpyAsync(marks.data(), m->useMlp.ptr, m->count, hipMemcpyDeviceToHost, stream));
    if (errors) GSR_HIP(hipMemcpyAsync(errors, m->errors.ptr, size_t(m->count) * 4, hipMemcpyDeviceToHost, stream));
    GSR_HIP(hipStreamSynchronize(stream));
    uint32_t c = 0;
    for (uint8_t v : marks) c += v ? 1u : 0u;
    m->mlpCount = c;
    if (mlpCount) *mlpCount = c;
    return GSR_OK;
}

gsr_status gsr_renderer_set_colour_model(gsr_renderer* r, gsr_colour_model* m) {
    if (!r) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_set_colour_model: null renderer");
    if (m && m->scene != r->scene)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_renderer_set_colour_model: the model belongs to another scene");
    if (m && r->settings.differentiable)
        return fail(GSR_ERROR_UNSUPPORTED, "gsr_renderer_set_colour_model: a differentiable renderer colours from SH");
    if (r->colour != m) ++r->generation;
    r->colour = m;
    return GSR_OK;
}

gsr_status gsr_backward(gsr_renderer* r, const void* gradImage, const gsr_gradients* out, void* streamArg) {
    if (!r || !gradImage || !out) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_backward: null argument");
    if (!r->settings.differentiable)
        return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_backward: the renderer was not created differentiable");
    if (!r->haveFrame) return fail(GSR_ERROR_INVALID_ARGUMENT, "gsr_backward: no frame rendered yet");
    GSR_HIP(hipSetDevice(r->ctx->device));
    hipStream_t stream = streamArg ? static_cast<hipStream_t>(streamArg) : r->ctx->stream;
    const gsr_scene* sc = r->scene;
    const gsr_camera& cam = r->lastCam;
    DevCamera dc{};
    std::memcpy(dc.view, cam.view_inria, sizeof(dc.view));
    std::memcpy(dc.pos, cam.position, sizeof(dc.pos));
    dc.tanX = cam.tan_half_fov_x;
    dc.tanY = cam.tan_half_fov_y;
    dc.focalX = cam.focal_x;
    dc.focalY = cam.focal_y;
    dc.width = cam.width;
    dc.height = cam.height;
    dc.tilesX = cam.tiles_x;
    dc.tilesY = cam.tiles_y;
    dc.binSize = r->binSize;
    dc.binsX = r->binsX(cam);
    dc.binsY = r->binsY(cam);
    SceneRef ref = sc->ref;
    ref.degree = r->degree;
    // The frame's stream may differ from this one: order after it.
    GSR_HIP(orderAfterLastFrame(r, stream, false));
    GSR_HIP(hipMemsetAsync(r->grad2D.ptr, 0, size_t(sc->count) * 4 * size_t(grad2DComponents()), stream));
    const uint32_t* values = r->sortedIn ? r->valuesB.as<uint32_t>() : r->valuesA.as<uint32_t>();
    launchRenderBackward(r->splats.as<float4>(), values, r->ranges.as<uint2>(), dc, static_cast<const float4*>(gradImage),
                         r->pixelState.as<uint2>(), r->grad2D.as<float>(), sc->count, stream);
    GSR_HIP(hipGetLastError());
    launchPreprocessBackward(ref, dc, r->splats.as<float4>(), r->grad2D.as<float>(), out->means, out->log_scales,
                             out->quats, out->opacity_logits, out->sh, stream);
    GSR_HIP(hipGetLastError());
    return GSR_OK;
}

}  // extern "C"
