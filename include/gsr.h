/* gsr.h: C ABI of the MI355X (gfx950) splat-tiles forward renderer, libgsr.so.
 *
 * Drop-in boundary for the reference's `splat-tiles` path. Each entry point names the
 * reference interface it replaces (paths are in netolcc06/gaussian-splat-renderer):
 *
 *   gsr_context_*          VulkanContext for this path: one device, one stream (core/context.h:9-10)
 *   gsr_scene_*            GpuScene::upload / destroy (render/gpu_scene.h:42-78) for the f32 SoA
 *                          layout (scene/gaussian_scene.h:1-36) and the fp16 records
 *                          (scene/scene_compression.h:28-35: G16 geometry + S16 colour)
 *   gsr_camera_*           makeSplatCamera / splatCameraFromViewer (splat/splat_math.cpp:9-42)
 *   gsr_renderer_create    SplatTileRenderer::init + onRenderTargetsChanged
 *                          (render/splat_tile_renderer.h:145-147, splat_tile_renderer.cpp:126-570)
 *   gsr_render             SplatTileRenderer::render (splat_tile_renderer.cpp:624-833): the whole
 *                          frame, asynchronous on the given stream, no host synchronisation
 *   gsr_read_*             readBack / readBackSplats / readBackImage / lastStats
 *                          (splat_tile_renderer.h:156-196); blocking, for tests and tools
 *
 * Conventions: plain pointers and sizes, no C++ or torch types. Every call returns a
 * gsr_status; on failure gsr_last_error() (per thread) says why. Handles are not thread
 * safe; one thread records a renderer's frames, as in the reference (application.h:55).
 * Device pointers are HIP device addresses of the context's device.
 */
#ifndef GSR_H
#define GSR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gsr_status {
    GSR_OK = 0,
    GSR_ERROR_INVALID_ARGUMENT = 1, /* std::runtime_error on bad settings/formats (splat_tile_renderer.cpp:139-163) */
    GSR_ERROR_HIP = 2,              /* a HIP runtime call failed (VK_CHECK, vk_common.h:28-35) */
    GSR_ERROR_OUT_OF_MEMORY = 3,    /* device allocation failed; nothing is left allocated (gpu_scene.cpp:108-114) */
    GSR_ERROR_UNSUPPORTED = 4,      /* a mode the library does not provide */
    GSR_ERROR_FORMAT = 5,           /* a bad scene file (PlyError, scene/ply_loader.h:29-32) */
    GSR_ERROR_IO = 6                /* a file could not be opened or read (std::runtime_error, ply_loader.h:86-88) */
} gsr_status;

typedef struct gsr_context gsr_context;
typedef struct gsr_scene gsr_scene;
typedef struct gsr_renderer gsr_renderer;
typedef struct gsr_buffer gsr_buffer;

/* Scene formats (scene_compression.h:6-35). */
typedef enum gsr_scene_format {
    GSR_SCENE_F32_SOA = 0, /* posAlpha f32x4, scales f32x4 (sigma, 0), rotations f32x4 (w,x,y,z), sh f32[N*K*3] */
    GSR_SCENE_F16 = 1      /* geometry 8 words (pos f32 x3, h(alpha,sx), h(sy,sz), h(qw,qx), h(qy,qz), 0) +
                              colour K*3 fp16 padded to a multiple of 8 halves; q renormalised on decode */
} gsr_scene_format;

/* TileBounds (splat/tile_coverage.h:21). */
typedef enum gsr_bounds { GSR_BOUNDS_RECT = 0, GSR_BOUNDS_TIGHT = 1, GSR_BOUNDS_CUTOFF_ONLY = 2 } gsr_bounds;

/* SplatTargetFormat for this path: RGBA8 (the product) or RGBA32F (float image, tests). */
typedef enum gsr_target_format { GSR_TARGET_RGBA8 = 0, GSR_TARGET_RGBA32F = 1 } gsr_target_format;

/* CullReason (splat/splat_math.h:110). */
typedef enum gsr_cull_reason {
    GSR_CULL_VISIBLE = 0,
    GSR_CULL_NEAR_PLANE = 1,
    GSR_CULL_SINGULAR_COV = 2,
    GSR_CULL_OFF_SCREEN = 3,
    GSR_CULL_LOW_OPACITY = 4
} gsr_cull_reason;

/* SplatCamera (splat/splat_math.h:43-50). view_inria is column-major (m[col*4+row]). */
typedef struct gsr_camera {
    float view_inria[16];
    float position[3];
    uint32_t width, height;
    float tan_half_fov_x, tan_half_fov_y;
    float focal_x, focal_y;
    uint32_t tiles_x, tiles_y;
} gsr_camera;

/* ProjectedSplat / GLSL Splat, 48 bytes (splat/splat_math.h:115-123). */
typedef struct gsr_projected_splat {
    float center[2];
    float depth;
    uint32_t cull_reason;
    float conic_opacity[4];
    float color[3];
    float radius;
} gsr_projected_splat;

/* TileControl, 80 bytes (render/splat_tile_renderer.h:116-128). */
typedef struct gsr_control {
    uint32_t gaussian_count;   /* N */
    uint32_t visible;          /* splats with a non-zero tile count */
    uint32_t duplicates;       /* M, unclamped */
    uint32_t entries;          /* min(M, capacity) */
    uint32_t cull_counts[8];   /* per gsr_cull_reason, [0] = visible */
    uint32_t overflow;         /* M - entries */
    uint32_t ranges_groups[3]; /* grid of the tile-ranges pass (informational) */
    uint32_t coverage_mismatch;/* splats whose written tiles != their count (always 0) */
    uint32_t sort_timeouts;    /* radix-sort look-back waits that gave up (always 0) */
    uint32_t pad[2];
} gsr_control;

/* Scene description. Host pointers are copied at creation; with device_pointers != 0 the
 * 16-B-per-gaussian arrays (pos_alpha, scales, rotations) are bound in place (zero copy) and
 * must outlive the scene (gpu_scene.h:4-6). The wider records (sh; f16 geometry and colour)
 * are always re-laid out at creation into library-owned planes of 16-B quads. */
typedef struct gsr_scene_desc {
    uint32_t count;
    int32_t sh_degree; /* 0..3; K = (sh_degree + 1)^2 coefficients per channel */
    gsr_scene_format format;
    int32_t device_pointers;
    const float* pos_alpha;   /* F32_SOA */
    const float* scales;      /* F32_SOA */
    const float* rotations;   /* F32_SOA */
    const float* sh;          /* F32_SOA */
    const uint32_t* geometry; /* F16: 8 words per gaussian */
    const uint16_t* color;    /* F16: round_up(K*3, 8) halves per gaussian */
} gsr_scene_desc;

/* SplatSettings + TileSettings for this path (splat_tile_renderer.h:73-114). */
typedef struct gsr_settings {
    gsr_bounds bounds;           /* default TIGHT (cli.h:44) */
    uint32_t budget_mb;          /* list memory: 16.5 (binning) or 24.5 (global) B per entry, in blocks of
                                    2048 entries (splat_tile_renderer.cpp:117-124); 0 = 256 */
    int32_t max_sh_degree;       /* evaluated bands = min(scene degree, this) */
    gsr_target_format format;    /* output image format */
    int32_t profile_stages;      /* 1: hipEvent pairs around every stage (gsr_read_stage_times); 2: also an
                                    event after every kernel (gsr_read_kernel_times) */
    int32_t render_pixels_per_thread; /* tile_render: 1, 2 or 4 pixels per thread (0 = default 1) */
    int32_t render_exact_alpha;  /* 1: alpha = min(0.99, o expf(power)) in the reference's operation order for
                                    every pair; 0 (default): log2-domain alpha with v_exp_f32 and a cutoff
                                    test before the exponential (same image within 1/255) */
    int32_t sort_algorithm;      /* GSR_SORT_*: radix-sort pass structure (same result, bit for bit) */
    int32_t render_loop;         /* GSR_RENDER_LOOP_*: loop form of the fast one-pixel render */
    int32_t sort_mode;           /* GSR_SORT_MODE_*: how the tile list is built (same list without overflow) */
    int32_t full_key_sort;       /* global mode: 0 (default) = the radix sort orders 32-bit tile ids only and a
                                    per-tile sort orders each tile's range by (depth, index); 1 = the radix sort
                                    orders all bits of 64-bit (tile, depth) keys (so does the one-sweep sort).
                                    Same list, bit for bit. */
    int32_t render_splat;        /* GSR_RENDER_SPLAT_*: the record the render pass gathers per list entry */
    uint32_t sort_window;        /* StopThePop per-pixel re-sorting window K (TileSettings::sortWindow,
                                    splat_tile_renderer.h:108): 0 = INRIA order; 1..64 = blend in the order
                                    of a K-slot window sorted by the depth of each splat's maximum along the
                                    pixel's ray (tile_render.comp:201-279), reference alpha; K <= 16 in
                                    registers, 17..64 in per-pixel LDS heaps (same order); K > 64:
                                    GSR_ERROR_INVALID_ARGUMENT (splat_tile_renderer.cpp:142) */
    int32_t adaptive_list;       /* TileSettings::adaptiveList (splat_tile_renderer.h:104): size the list from the
                                    measured entry count (adaptCapacity, splat_tile_renderer.cpp:444-486), within
                                    the budget; resizes never stall the frames in flight */
    uint32_t adaptive_initial_entries; /* first guess (0 = max(4 N, 2^20)) */
    int32_t frame_stats;         /* 1: read every frame's control block back asynchronously
                                    (gsr_renderer_last_stats); implied by adaptive_list */
    int32_t differentiable;      /* 1: frames keep what gsr_backward needs (no counterpart in the reference):
                                    the render pass is the training forward (exact alpha, per-pixel final
                                    transmittance and last contributor). Needs format RGBA32F, render_splat
                                    FULL, sort_window 0 and an f32 scene */
    int32_t tile_stats;          /* 1: every frame also counts the per-tile render statistics of the reference's
                                    stats build (TileStat, splat_tile_renderer.h:130-141; gsr_read_tile_stats).
                                    Culled render loop only; the counting slows the render pass */
    int32_t tile_size;           /* TileSettings::tileSize (splat_tile_renderer.h:107): the lists' bins, 16 or 8
                                    px (0 = 16). INRIA's getRect stays on 16-px tiles; with 8 the coverage is
                                    taken per 8x8 bin (tile ids, ranges, gsr_read_lists and the tile statistics
                                    follow the bin grid ceil(w / 8) x ceil(h / 8)) and each bin is rendered by one
                                    wave. 8 takes the culled loop only (no StopThePop, differentiable, exact
                                    alpha or tile_stats) */
    uint32_t render_batch;       /* TileSettings::renderBatch (splat_tile_renderer.h:110): splats staged in LDS per
                                    round of the culled loop; 0 = one per thread (256 per 16-px tile, 64 per 8-px
                                    bin). 16-px tiles: 128, 256 or 512; 8-px bins: 32, 64 or 128. Non-default
                                    values need the culled loop (the autotuner's batch depth) */
    uint32_t reserved[1];
} gsr_settings;

/* RenderSplatMode (splat_tile_renderer.h:88-92; render_splat.glsl:1-21), same values. The
 * preprocess writes the packed record next to the 48-byte splat; the render gathers it.
 *   p32 (8 words): cx, cy, a, b | c, opacity, h(r) | h(g) << 16, h(b) | radius << 16
 *   p24 (6 words): cx, cy | h(a 2^12) | h(b 2^12) << 16, h(c 2^12) | h(opacity) << 16 | colour as p32
 * h() = IEEE fp16, round to nearest even; radius = min(radius, 65535) as an integer. In p24 mode
 * the 48-byte splat holds the conic and opacity as the record gives them back, and the tiles are
 * counted with those values (splat_preprocess.comp:296-305). */
typedef enum gsr_render_splat {
    GSR_RENDER_SPLAT_FULL = 0, /* the 48-byte splat */
    GSR_RENDER_SPLAT_P32 = 1,  /* 32 bytes: fp32 centre, conic, opacity; fp16 colour */
    GSR_RENDER_SPLAT_P24 = 2   /* 24 bytes: fp32 centre; fp16 conic (x 2^12), opacity, colour */
} gsr_render_splat;

/* Tile-list construction (TileSortMode, splat_tile_renderer.h:74-78; same values). */
typedef enum gsr_sort_mode {
    GSR_SORT_MODE_BINNING = 0, /* depth sort of the visible splats (32-bit keys), duplicate in depth order,
                                  stable sort of 32-bit tile ids; 16.5 B per list entry (default) */
    GSR_SORT_MODE_GLOBAL = 1   /* duplicate in index order, the list in (tile, depth, index) order as one sort
                                  of 64-bit (tile, depth) keys gives it; 24.5 B per list entry */
} gsr_sort_mode;

/* Loop form of the fast one-pixel tile render (same blend, same image within 1/255). */
typedef enum gsr_render_loop {
    GSR_RENDER_LOOP_CULLED = 0,     /* predicated, one 8x8 quadrant per wave, each wave walks only the splats
                                       whose alpha >= 1/255 box meets its quadrant (default) */
    GSR_RENDER_LOOP_DIVERGENT = 1,  /* per-lane exits, one splat per iteration */
    GSR_RENDER_LOOP_PREDICATED = 2  /* wave-uniform, 4 splats per iteration, branch-free blend, 16x4 rows per
                                       wave, every splat of the tile */
} gsr_render_loop;

/* Radix-sort pass structure. */
typedef enum gsr_sort_algorithm {
    GSR_SORT_REDUCE_THEN_SCAN = 0, /* per digit: block histograms, scan, ballot-ranked scatter (default) */
    GSR_SORT_ONESWEEP = 1          /* one histogram pass, then one scatter per digit with decoupled look-back;
                                      64-bit keys only (global mode: all key bits; gsr_sort_pairs) */
} gsr_sort_algorithm;

/* Stage timings of the last profiled frame, ms (the GpuProfiler scopes, profiler.h:1-87). */
typedef enum gsr_stage {
    GSR_STAGE_PREPROCESS = 0, /* preprocess + cull + tile counts + block sums */
    GSR_STAGE_SCAN = 1,       /* binning: block-sum scan + control (tile_setup) + compaction of the visible;
                                 global: nothing (the duplicate takes its offsets from the preprocess's sums) */
    GSR_STAGE_DEPTH_SORT = 2, /* binning: depth sort of the visible splats (global: 0) */
    GSR_STAGE_DUPLICATE = 3,  /* in-block scan + duplication; binning: + count gather and block-sum scan */
    GSR_STAGE_SORT = 4,       /* all radix passes of the tile-list sort; global mode, full_key_sort 0: + ranges
                                 (+ the per-tile depth sort, unless the culled render loop does it) */
    GSR_STAGE_RANGES = 5,     /* tile ranges (global mode, full_key_sort 0: nothing, counted in SORT) */
    GSR_STAGE_RENDER = 6,     /* tile render incl. RGBA8 resolve (global mode's culled loop: + each tile's
                                 depth sort) */
    GSR_STAGE_TOTAL = 7,
    GSR_STAGE_COUNT = 8
} gsr_stage;

const char* gsr_last_error(void);
const char* gsr_version(void);

/* Context: device ordinal + the stream used when gsr_render gets a null stream. */
gsr_status gsr_context_create(int32_t device, gsr_context** out);
gsr_status gsr_context_synchronize(gsr_context* ctx);
void gsr_context_destroy(gsr_context* ctx);

/* Plain device buffers for callers without their own allocator (the caller-owned colour
 * target of FrameInput, renderer.h:32-35). */
gsr_status gsr_buffer_create(gsr_context* ctx, size_t bytes, gsr_buffer** out);
void* gsr_buffer_device_ptr(gsr_buffer* buf);
gsr_status gsr_buffer_read(gsr_buffer* buf, void* host_dst, size_t bytes);
gsr_status gsr_buffer_write(gsr_buffer* buf, const void* host_src, size_t bytes);
void gsr_buffer_destroy(gsr_buffer* buf);

/* GpuScene::upload (gpu_scene.h:42-78) / destroy. */
gsr_status gsr_scene_create(gsr_context* ctx, const gsr_scene_desc* desc, gsr_scene** out);
void gsr_scene_destroy(gsr_scene* scene);

/* The fp16 encoder of scene_compression.cpp:34-49 (geometry) and the S16 colour record:
 * fp32 -> fp16 round to nearest even (half.h:32-64). Host only. */
gsr_status gsr_encode_f16(uint32_t count, int32_t sh_degree, const float* pos_alpha, const float* scales,
                          const float* rotations, const float* sh, uint32_t* geometry_out, uint16_t* color_out);

/* makeSplatCamera (splat_math.cpp:9-27) and splatCameraFromViewer (splat_math.cpp:29-42).
 * Host only; matrices column-major. */
gsr_status gsr_camera_make(const float view_inria[16], uint32_t width, uint32_t height, float tan_half_fov_x,
                           float tan_half_fov_y, gsr_camera* out);
gsr_status gsr_camera_from_viewer(const float view[16], const float proj[16], const float model[16], uint32_t width,
                                  uint32_t height, gsr_camera* out);

/* SplatTileRenderer::init (+ the per-size buffers of onRenderTargetsChanged, created on
 * first use of a size). The renderer binds the scene; the scene must outlive it. */
gsr_status gsr_renderer_create(gsr_context* ctx, const gsr_scene* scene, const gsr_settings* settings,
                               gsr_renderer** out);
void gsr_renderer_destroy(gsr_renderer* r);
uint32_t gsr_renderer_capacity(const gsr_renderer* r);     /* list capacity, entries */
uint32_t gsr_renderer_max_capacity(const gsr_renderer* r); /* capacity of the full budget (maxCapacity) */
uint32_t gsr_renderer_list_resizes(const gsr_renderer* r); /* adaptive list resizes so far (listResizes) */
/* lastStats (splat_tile_renderer.h:167-176): the control block of the newest frame whose
 * asynchronous read-back has completed, and its index (0 = the first frame rendered). Does not
 * wait for the GPU; fails while no frame has completed. Needs frame_stats or adaptive_list. */
gsr_status gsr_renderer_last_stats(gsr_renderer* r, gsr_control* out, uint64_t* frame_index);
uint64_t gsr_renderer_total_bytes(const gsr_renderer* r);  /* device memory owned by the renderer */

/* SplatTileRenderer::render: the whole frame on `stream` (null = the context's stream),
 * no host synchronisation. out_image: device pointer, width*height pixels of
 * settings.format (RGBA8: 4 bytes; RGBA32F: 16 bytes), row 0 = top, tightly packed. Every
 * pixel is written: C + T * background (black), alpha = 1 - T. A frame on another stream than
 * the renderer's previous frame waits for that frame (the two share the renderer's buffers);
 * frames recorded into a graph capture are ordered by the stream the graph is launched on. */
gsr_status gsr_render(gsr_renderer* r, const gsr_camera* cam, void* out_image, void* stream);

/* Test/debug readbacks of the last rendered frame (blocking; the GPU is synchronised).
 * read_control: the 80-byte control block. read_lists: the sorted list (control.entries
 * entries: tile id, depth bits, gaussian index) and [start, end) per tile (2 per tile);
 * any pointer may be null. read_splats: N projected splats + tile counts (index order). */
gsr_status gsr_read_control(gsr_renderer* r, gsr_control* out);
gsr_status gsr_read_lists(gsr_renderer* r, uint32_t* tile_keys, uint32_t* depth_keys, uint32_t* values,
                          uint32_t* ranges, uint32_t max_entries);
gsr_status gsr_read_splats(gsr_renderer* r, gsr_projected_splat* splats, uint32_t* tile_counts);
/* readBackPackedSplats (splat_tile_renderer.h:184-187), packed modes: the render-splat words of
 * every gaussian (8 or 6 each; culled gaussians' words are not written). The tile rectangles
 * and depths of the reference's call are in the 48-byte splats (gsr_read_splats). */
gsr_status gsr_read_packed_splats(gsr_renderer* r, uint32_t* words);
gsr_status gsr_read_stage_times(gsr_renderer* r, float* ms, uint32_t n);
/* The reference's per-tile render statistics of the last frame (tile_stats = 1; TileStat,
 * splat_tile_renderer.h:130-141, counted as tile_render.comp:78-90, 380-399): GSR_TILE_STAT_COUNT
 * counters per tile, tile-major, for min(tiles, max_tiles) tiles. */
typedef enum gsr_tile_stat {
    GSR_TILE_STAT_LIST_LENGTH = 0, /* entries in the tile's range */
    GSR_TILE_STAT_FETCHED = 1,     /* entries loaded before the whole tile stopped */
    GSR_TILE_STAT_EVALUATED = 2,   /* (pixel, entry) pairs of the reference's loop: a pixel walks every fetched
                                      entry until it terminates (the terminating one included) */
    GSR_TILE_STAT_BLENDED = 3,     /* pairs blended */
    GSR_TILE_STAT_TERMINATED = 4,  /* pixels that reached T < 1e-4 */
    GSR_TILE_STAT_LANE_SLOTS = 5,  /* lane slots this library's render loop spent evaluating pairs (per wave and
                                      batch: the list slots it walked x 64) */
    GSR_TILE_STAT_MAX_DEPTH = 6,   /* the largest evaluated count of one pixel */
    GSR_TILE_STAT_PIXELS = 7,      /* pixels of the tile inside the image */
    GSR_TILE_STAT_COUNT = 8
} gsr_tile_stat;
gsr_status gsr_read_tile_stats(gsr_renderer* r, uint32_t* stats, uint32_t max_tiles);
/* profile_stages = 2: the kernels of the last frame in launch order (static names) and the time of
 * each in ms (hipEvents recorded between consecutive launches on the frame's stream). Fills
 * min(kernels, max) entries (either array may be null); *count (may be null) = the kernels. */
gsr_status gsr_read_kernel_times(gsr_renderer* r, const char** names, float* ms, uint32_t max, uint32_t* count);

/* GpuRadixSort on its own (gpu_radix_sort.h:38-122, test_gpu_sort.cpp): stable LSD sort of
 * (key, value) pairs by the low `key_bits` bits of 64-bit keys, in place, on device
 * pointers. `count_device` (may be null) holds the key count on the device, clamped to
 * max_count; null = max_count keys. Asynchronous on `stream`. */
gsr_status gsr_sort_pairs(gsr_context* ctx, uint64_t* keys, uint32_t* values, uint32_t max_count,
                          const uint32_t* count_device, uint32_t key_bits, gsr_sort_algorithm algorithm, void* stream);

/* gsr_sort_pairs without any allocation. No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments, result and stability as gsr_sort_pairs with the
 * reduce-then-scan algorithm, but the ping-pong copy and the histograms live in a fixed
 * library-owned device array (about 12.5 MiB, held by every GPU the library runs on until the
 * process ends), so max_count is at most 1,048,576 keys (GSR_ERROR_INVALID_ARGUMENT above that)
 * and the call makes no hipMalloc, hipFree or other allocating call. Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams, other threads) race;
 * order them. Asynchronous on `stream` (null = the context's). */
gsr_status gsr_sort_pairs_noalloc(gsr_context* ctx, uint64_t* keys, uint32_t* values, uint32_t max_count,
                                  const uint32_t* count_device, uint32_t key_bits, void* stream);

/* GpuScan on its own (gpu_scan.h:25-76): exclusive prefix sum of u32 values, the total in
 * *total_device. Asynchronous on `stream`. */
gsr_status gsr_exclusive_scan(gsr_context* ctx, const uint32_t* in, uint32_t* out, uint32_t count,
                              uint32_t* total_device, void* stream);

/* --- Batched multi-view ------------------------------------------------------------------------
 * No counterpart in the reference (its benchmark renders one view per frame, application.cpp:
 * 1042-1057); for callers that render many cameras of one scene at once (dataset views, light
 * probes, multi-camera rigs). A batch owns `branches` HIP streams (1..16, default priority).
 * gsr_batch_render renders view v with renderers[v % b] on stream v % b, b = min(branches,
 * view_count), each branch's views in order, forked from `stream` and joined back to it:
 * asynchronous and stream-ordered like gsr_render, and the images equal those of one
 * gsr_render per view. Each branch needs its own renderer of the scene (lists are per
 * renderer); renderers[0 .. b-1] must be distinct. gsr_batch_capture records the same work into
 * a HIP graph held by the batch (it replaces the previous one); gsr_batch_replay launches it on
 * `stream`, one launch for the whole batch, with the cameras, images and renderers of the
 * capture. A capture needs every view's size rendered before by its branch's renderer (render
 * the batch once first) and plain renderers (no adaptive_list, frame_stats or profile_stages;
 * the reduce-then-scan sort). gsr_batch_replay fails with GSR_ERROR_INVALID_ARGUMENT, launching
 * nothing, when a renderer of the capture has been destroyed, has reallocated a buffer (a view
 * larger than it had rendered) or has changed its colour model or fallback marks since the
 * capture: capture again. Replays and other work on the same renderers must be stream-ordered
 * by the caller (replays on one stream are). */
typedef struct gsr_batch gsr_batch;
gsr_status gsr_batch_create(gsr_context* ctx, uint32_t branches, gsr_batch** out);
void gsr_batch_destroy(gsr_batch* batch);
uint32_t gsr_batch_branches(const gsr_batch* batch);
/* Persistent mode (persistent != 0; off by default): every view gets its own renderer
 * (renderers[0 .. view_count-1] distinct), runs its frame up to the tile depth sort on branch
 * v % b, and after the join ONE render launch (about 8 workgroups per CU) draws all views' tiles,
 * each workgroup taking the next (view, tile) from a device counter, with the same tile body as
 * the default culled loop, so the images equal those of stream mode. Renderers must use that
 * loop: 16-px tiles, one pixel per thread, no exact alpha, StopThePop, differentiable,
 * tile_stats or non-default render_batch (GSR_ERROR_UNSUPPORTED), one image format and
 * render_splat for all views. Switching the mode waits for the device and drops a captured graph. */
gsr_status gsr_batch_set_persistent(gsr_batch* batch, int32_t persistent);
gsr_status gsr_batch_render(gsr_batch* batch, gsr_renderer* const* renderers, uint32_t view_count,
                            const gsr_camera* cams, void* const* out_images, void* stream);
gsr_status gsr_batch_capture(gsr_batch* batch, gsr_renderer* const* renderers, uint32_t view_count,
                             const gsr_camera* cams, void* const* out_images);
gsr_status gsr_batch_replay(gsr_batch* batch, void* stream);

/* --- Learned colour ----------------------------------------------------------------------------
 * No counterpart in the reference (its colour is always the SH evaluation, evalSH,
 * splat_math.cpp:129-151). A colour model gives every gaussian of a scene a 16-value latent and
 * shares one MLP among them: [latent, view direction, 1] -> 64 -> ReLU -> 64 -> ReLU -> 3,
 * colour = max(y + 0.5, 0), computed with bf16 MFMA products and fp32 accumulation. It is
 * trained (Adam) towards the scene's own SH colours on random (gaussian, direction) samples; a
 * fallback pass then marks the gaussians whose learned colour is within `threshold` of the SH
 * colour in all 32 test directions, and a renderer bound to the model colours those from the
 * MLP and the rest from SH. Parameters (gsr_colour_model_param_count floats): W1 [64][32]
 * (columns 0..15 the latent, 16..18 the direction, 19 the bias, 20..31 zero), W2 [64][64],
 * b2 [64], W3 [3][64], b3 [3]. Latents: count x 16 floats (rendered through fp16 copies).
 * Host pointers throughout; calls are synchronous unless they say otherwise. The model must
 * outlive every renderer bound to it. */
typedef struct gsr_colour_model gsr_colour_model;
uint32_t gsr_colour_model_param_count(void);
gsr_status gsr_colour_model_create(gsr_context* ctx, const gsr_scene* scene, uint32_t seed, gsr_colour_model** out);
void gsr_colour_model_destroy(gsr_colour_model* model);
/* `steps` Adam steps of `samples_per_step` samples each (a multiple of 256); *last_loss (may be
 * null) = the mean squared error of the last step, per channel, before its update. A sample stream
 * draws its gaussians as a permutation (sample s: (a s + b) mod count, a coprime to count), so a
 * step of at most `count` samples touches distinct gaussians; such a step updates the weights with
 * Adam and only the touched latents (lazy Adam: the other latents and their moments keep their
 * values). A larger step (a scene smaller than the step) adds the latent gradients of repeated
 * gaussians and takes a dense Adam step on every latent. */
gsr_status gsr_colour_model_train(gsr_colour_model* model, uint32_t steps, uint32_t samples_per_step,
                                  float lr_weights, float lr_latents, float* last_loss);
/* The mean squared error over `samples` samples (a multiple of 256) of stream `seed`. */
gsr_status gsr_colour_model_loss(gsr_colour_model* model, uint32_t seed, uint32_t samples, float* loss);
/* The gradient of that error with respect to the parameters and latents (no update), and the
 * samples it used (gaussians: samples values, dirs: samples x 3; either may be null). */
gsr_status gsr_colour_model_gradients(gsr_colour_model* model, uint32_t seed, uint32_t samples, float* grad_params,
                                      float* grad_latents, float* loss, uint32_t* gaussians, float* dirs);
gsr_status gsr_colour_model_get(gsr_colour_model* model, float* params, float* latents);
gsr_status gsr_colour_model_set(gsr_colour_model* model, const float* params, const float* latents);
/* y + 0.5 (unclamped) for n (gaussian, direction) pairs; rgb: n x 3. */
gsr_status gsr_colour_model_evaluate(gsr_colour_model* model, const uint32_t* gaussians, const float* dirs, uint32_t n,
                                     float* rgb);
/* Marks the gaussians that use the MLP (error <= threshold; a negative threshold marks none);
 * *mlp_count (may be null) = how many; errors (may be null): count floats, the largest channel
 * difference of each gaussian over the test directions. */
gsr_status gsr_colour_model_fallback(gsr_colour_model* model, float threshold, uint32_t* mlp_count, float* errors);
/* Binds a model of the renderer's scene (null: plain SH colour again). */
gsr_status gsr_renderer_set_colour_model(gsr_renderer* r, gsr_colour_model* model);

/* --- Differentiable rendering ------------------------------------------------------------------
 * No counterpart in the reference (it renders forward only). For a renderer created with
 * `differentiable`, gsr_backward takes dL/d(image) of its last frame (device, RGBA32F layout of
 * that frame's size; the alpha channel is ignored) and writes the gradients of L with respect to
 * the scene's parameters, as seen through the image out = sum_i c_i a_i T_i (black background):
 *   means           N x 3
 *   log_scales      N x 3   (the scene's scales are exp(log_scales))
 *   quats           N x 4   (w, x, y, z as stored; the f32 path uses them unnormalised)
 *   opacity_logits  N       (the scene's opacities are sigmoid(logits))
 *   sh              N x K x 3 (the scene's layout; bands above the evaluated degree get 0)
 * Device pointers, float32, each may be null; overwritten, not accumulated; culled gaussians get
 * zeros. The alpha clamp at 0.99 passes no gradient, as min() does. Asynchronous on `stream` (or
 * the context's), ordered after the frame when on the frame's stream. */
typedef struct gsr_gradients {
    float* means;
    float* log_scales;
    float* quats;
    float* opacity_logits;
    float* sh;
} gsr_gradients;
gsr_status gsr_backward(gsr_renderer* r, const void* grad_image, const gsr_gradients* out, void* stream);

/* --- PLY scenes (loadGaussianPly, scene/ply_loader.h:1-12 and 86-95) ---------------------------
 * A binary little-endian INRIA 3DGS .ply file decoded and activated on the host into the f32
 * SoA arrays gsr_scene_create takes (format GSR_SCENE_F32_SOA, device_pointers 0):
 *   pos_alpha  (x, y, z, 1 / (1 + exp(-opacity)))      count x 4
 *   scales     (exp(scale_0..2), 0)                      count x 4
 *   rotations  (rot_0..3) / |q|, w first; identity for |q|^2 <= 1e-24 or non-finite
 *   sh         count x K x 3, coefficient-major RGB triplets (the file's f_rest is channel-major)
 * plus the raw file values. The arrays belong to the gsr_ply_scene until gsr_ply_free.
 * Bad files fail with GSR_ERROR_FORMAT and a message naming the problem (the PlyError texts). */
typedef struct gsr_ply_scene {
    uint32_t count;
    int32_t sh_degree;            /* 0..3 from the number of f_rest_* properties (0/9/24/45) */
    const float* pos_alpha;
    const float* scales;
    const float* rotations;
    const float* sh;
    const float* raw_opacity;     /* count: the opacity logits */
    const float* raw_log_scales;  /* count x 3 */
    const float* raw_rotations;   /* count x 4, as stored */
    uint64_t non_finite_values;   /* NaN/Inf values read */
    uint32_t zero_quaternions;    /* replaced by the identity */
    uint32_t warning_count;
    const char* const* warnings;  /* skipped elements, trailing bytes, conversions, the counts above */
    uint32_t comment_count;
    const char* const* comments;  /* the header's comment / obj_info lines */
    void* internal;
} gsr_ply_scene;

gsr_status gsr_ply_load(const char* path, gsr_ply_scene** out);
gsr_status gsr_ply_load_memory(const void* bytes, uint64_t size, const char* name, gsr_ply_scene** out);
void gsr_ply_free(gsr_ply_scene* scene);

#ifdef __cplusplus
}
#endif

#endif /* GSR_H */
