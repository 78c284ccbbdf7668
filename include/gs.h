/* gs.h - C ABI of the MI355X tiled Gaussian-splatting frame pipeline.
 *
 * This is the drop-in boundary for the reference's per-frame render path
 * (vinneyto/3dgs_tile_webgpu). Each entry point cites the reference
 * interface it replaces. All pointers are plain host (or, where stated,
 * device) pointers; no PyTorch or HIP types cross the boundary.
 *
 * Conventions (reference GaussianPass.ts:161-260, 1010-1017;
 * docs/dynamic-intersections.md:25-54):
 *  - every call returns an int status: GS_OK or a negative GS_E_* code;
 *  - configuration errors return GS_E_RANGE / GS_E_TYPE and leave the
 *    context unchanged (the reference throws RangeError / TypeError);
 *  - intersection overflow is NOT an error: the frame publishes nothing,
 *    keeps the previous image and reports the demand via gs_read_state;
 *  - gs_last_error returns a per-context message for the last failure.
 *
 * Threading: a context is single-threaded and owns one HIP stream. Render
 * calls are asynchronous on that stream; gs_read_state, gs_read_image and
 * gs_read_debug synchronise. Independent contexts may run concurrently.
 */
#ifndef GS_H
#define GS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_OK 0
#define GS_E_RANGE (-1)  /* RangeError: a value outside its domain */
#define GS_E_TYPE (-2)   /* TypeError: an unsupported format or variant */
#define GS_E_OOM (-3)    /* allocation-failed */
#define GS_E_DEVICE (-4) /* HIP runtime failure */
#define GS_E_STATE (-5)  /* call order violated (e.g. render before data) */

/* pipeline/types.ts:12-17 */
#define GS_DEPTH_FLOAT32 0u
#define GS_DEPTH_PACKED16 1u
#define GS_AA_COMPENSATED 0u
#define GS_AA_CLASSIC 1u
#define GS_BOUNDARY_CLIP 0u
#define GS_BOUNDARY_REPEAT_X 1u
#define GS_BOUNDARY_REPEAT_Y 2u

/* GaussianData.ts:10-33 formats */
#define GS_GEOMETRY_COMPACT 0u
#define GS_SH_RGB8E8 0u

#define GS_MEMORY_HOST 0u
#define GS_MEMORY_DEVICE 1u

typedef struct gs_ctx gs_ctx;

/* Pass options (pipeline/types.ts:27-82, GaussianPassOptions) that shape
 * the GPU pipeline. Defaults from gs_config_default match GaussianPass.ts
 * :157-260 and IntersectionCapacityPolicy.ts:71-81. */
typedef struct gs_config {
  int32_t device;                    /* HIP device ordinal */
  uint32_t depth_sort_mode;          /* GS_DEPTH_*, default FLOAT32 */
  uint32_t antialias_mode;           /* GS_AA_*, default COMPENSATED */
  uint32_t subpixel_sample_culling;  /* default 1 */
  uint32_t raster_chunk_size;        /* 0 = null (no chunks), default 8192 */
  uint32_t intersection_capacity;    /* initial record capacity, default 262144 */
  uint32_t max_intersection_capacity;/* hard limit, default 1 << 28, at most 2^30 - 1 */
  uint32_t output_depth;             /* also produce the r32f scalar target */
  float transmittance_threshold;     /* (0, 1), default 1e-4 */
  float depth_alpha_threshold;       /* default 0.95 */
  float invalid_depth_value;         /* default 1 */
  float background[4];               /* default (0, 0, 0, 0) */
  uint32_t raster_stats;             /* count checked/blended pairs */
  uint32_t capture_graph;            /* replay frames as a captured hipGraph
                                        (default 1); re-captured when sizes
                                        or buffers change (camera and object
                                        frames are uploaded per launch) */
  /* Launch shape (no reference counterpart; the reference fixes 16-pixel
   * tiles, constants.ts:1, 4-bit radix digits, radix.ts, and 256-record
   * raster batches). Only tile_size 16 yields keys comparable with the
   * reference; digit width and batch depth never change any output. */
  uint32_t tile_size;                /* 8, 16 or 32; default 16 */
  uint32_t radix_bits;               /* 4, 6 or 8; default 8 */
  uint32_t raster_batch;             /* 128 or 256; default 256 */
  uint32_t batch_lanes;              /* concurrent frames of a batch; only 1
                                        is accepted at present (default 1) */
  uint32_t heavy_split_units;        /* heavy raster units (tiles above 1024
                                        records, chunk tasks) above which 16-pixel
                                        tiles are rasterized with two lanes per
                                        pixel instead of four (whole tiles, then
                                        row slabs, see heavy_pair_units);
                                        GS_SPLIT_AUTO (default) = 1.25 per compute
                                        unit; never changes any output */
  uint32_t heavy_fine_split_units;   /* heavy raster units above which those
                                        slabs have 4 rows instead of 8;
                                        GS_SPLIT_AUTO (default) = 16 per compute
                                        unit; never changes any output */
  uint32_t view_group;               /* views of a batch rendered by each
                                        launch (1..8, default 1): every stage
                                        is one kernel over the group's views,
                                        largest views grouped first; each view
                                        of a group has its own working set.
                                        Never changes any output */
  uint32_t heavy_pair_units;         /* heavy raster units above which the
                                        two-lane shape is row slabs rather than
                                        whole tiles (applies above
                                        heavy_split_units); GS_SPLIT_AUTO
                                        (default) = 5.5 per compute unit; never
                                        changes any output */
} gs_config;

#define GS_SPLIT_AUTO 0xffffffffu

/* Packed Gaussian buffers (GaussianData.ts:10-33 / GaussianBuffers), copied
 * during the call (host memory) or device-to-device (device memory). No
 * pointer is retained after return. */
typedef struct gs_data_desc {
  uint32_t count;           /* slots in every buffer */
  uint32_t active_count;    /* active ordinals (<= count) */
  uint32_t geometry_format; /* GS_GEOMETRY_COMPACT */
  uint32_t sh_format;       /* GS_SH_RGB8E8 */
  uint32_t sh_degree;       /* 0..3 */
  uint32_t memory;          /* GS_MEMORY_HOST or GS_MEMORY_DEVICE */
  const float* means;              /* count * 4 f32: xyz, objectId (<0 hole) */
  const uint32_t* scales_opacity;  /* count * 2 u32 (compact) */
  const uint32_t* rotations;       /* count u32 (oct101012) */
  const uint32_t* sh;              /* count * (deg+1)^2 u32 (rgb8e8) */
  const uint32_t* active_slots;    /* active_count u32 source slots, or NULL
                                      for the identity list */
} gs_data_desc;

/* Per-frame camera and domain (FrameUniforms.ts:41-58 / prepareFrame). */
typedef struct gs_frame {
  float projection[16];  /* camera.projectionMatrix, column-major */
  uint32_t width, height;
  float near_plane, far_plane; /* depthRange (default camera near/far) */
  float filter_variance;       /* default 0.3 */
  uint32_t boundary_axis;      /* GS_BOUNDARY_* */
  uint32_t attempt;            /* render(attempt, resetValid) */
  uint32_t reset_valid;
  const float* object_frames;  /* this camera's object table (object_count
                                  * 40 f32, as gs_set_objects), or NULL for
                                  the table set by gs_set_objects */
} gs_frame;

/* Output targets. With memory == GS_MEMORY_HOST the call synchronises and
 * copies; with GS_MEMORY_DEVICE the kernels write the caller's device
 * buffers directly (asynchronous). NULL targets keep the library's own
 * images (see gs_read_image). color: W*H*4 binary16 (rgba16float);
 * depth: W*H f32 (only with output_depth). */
typedef struct gs_image {
  void* color;
  void* depth;
  uint32_t memory;
} gs_image;

/* outputState + dispatch state (intersections.ts:10-41, 114-134;
 * GaussianPassStats, types.ts:94-103). */
typedef struct gs_state {
  uint32_t attempt;
  uint32_t version;            /* last attempt whose image is complete */
  uint32_t requested;          /* demand K (saturates at 0xffffffff) */
  uint32_t overflow;           /* 1 if demand > capacity */
  uint32_t visible;            /* V */
  uint32_t intersections;      /* K published this frame (0 on overflow) */
  uint32_t capacity;
  uint32_t tiles_x, tiles_y;
  uint64_t raster_checked;     /* with raster_stats */
  uint64_t raster_blended;
  uint64_t raster_checked_heavy; /* the part of raster_checked in heavy units
                                    (tiles above 1024 records, chunk tasks) */
  uint32_t error_flags;        /* bit 0: a radix pass, bit 1: a scan exceeded
                                  its bounded wait for earlier partitions (the
                                  frame is invalid) */
} gs_state;

void gs_config_default(gs_config* config);

/* GaussianPass constructor + TiledGaussianPipeline constructor
 * (GaussianPass.ts:144-307, 852-877; TiledGaussianPipeline.ts:65-163). */
int gs_create(const gs_config* config, gs_ctx** out);
void gs_destroy(gs_ctx* ctx);
const char* gs_last_error(const gs_ctx* ctx);

/* TiledGaussianPipeline.updateData (TiledGaussianPipeline.ts:203-218,
 * ProjectionStage.ts:172-212, 298-329). */
int gs_set_data(gs_ctx* ctx, const gs_data_desc* data);

/* Active-slot update for [start, start+n) (GaussianData activation,
 * GaussianData.ts:170-232; ProjectionStage.updateActiveSlots :298-329).
 * new_active_count may shrink or grow the active list (<= count). */
int gs_update_active(gs_ctx* ctx, uint32_t start, uint32_t n,
                     const uint32_t* slots, uint32_t new_active_count);

/* Object frame table: 10 vec4 (40 f32) per object id: matrixWorld,
 * modelView, camera position in object space, visibility flag
 * (ObjectFrameState.ts:97-120). */
int gs_set_objects(gs_ctx* ctx, uint32_t object_count, const float* frames);

/* Fill one object's 40-float frame from f64 matrices the way
 * ObjectFrameState.writeCloud does (ObjectFrameState.ts:97-120):
 * camera_world and object_world are column-major 4x4 f64. */
void gs_object_frame(const double camera_world[16], const double object_world[16],
                     int visible, float out[40]);

/* prepareFrame + render (TiledGaussianPipeline.ts:286-341). */
int gs_render(gs_ctx* ctx, const gs_frame* frame, const gs_image* out);

/* Many cameras of uneven size in one captured launch (no reference
 * counterpart; the reference renders one camera per pass and shares data
 * through dataMode "external", docs/render-extensions.md:5-47). The n
 * frames run inside one hipGraph that is kept across calls while sizes and
 * buffers stay the same. Each view has
 * its own intersection capacity, state and (with outs == NULL or a NULL
 * color) library-owned images. Overflow behaves per view as in gs_render. */
int gs_render_batch(gs_ctx* ctx, uint32_t n, const gs_frame* frames,
                    const gs_image* outs);

/* States of batch views [first, first + n) after the last gs_render_batch
 * (synchronises). */
int gs_read_batch_state(gs_ctx* ctx, uint32_t first, uint32_t n, gs_state* out);

/* resizeIntersections for one batch view. */
int gs_resize_batch_intersections(gs_ctx* ctx, uint32_t view, uint32_t capacity);

/* Copy batch view `view`'s last image to host (sizes as gs_read_image), or
 * return its device images. */
int gs_read_batch_image(gs_ctx* ctx, uint32_t view, void* color, size_t color_bytes,
                        void* depth, size_t depth_bytes);
int gs_batch_device_image(gs_ctx* ctx, uint32_t view, void** color, void** depth);

/* Launch shape of later frames (the gs_config fields above). */
typedef struct gs_tuning {
  uint32_t tile_size;
  uint32_t radix_bits;
  uint32_t raster_batch;
} gs_tuning;
int gs_set_tuning(gs_ctx* ctx, const gs_tuning* tuning);
int gs_get_tuning(gs_ctx* ctx, gs_tuning* tuning);

/* Autotuner (no reference counterpart): renders the batch under every
 * launch shape (tile 8/16/32 x digits 4/6/8 x batch 128/256), growing view
 * capacities on overflow, times `reps` launches of each with HIP events,
 * leaves the fastest shape set on the context and returns it in *best.
 * table (table_len entries, may be NULL) receives each shape's mean ms per
 * batch; *entries the number of shapes tried. */
typedef struct gs_tune_entry {
  gs_tuning tuning;
  double ms;
} gs_tune_entry;
int gs_autotune(gs_ctx* ctx, uint32_t n, const gs_frame* frames, uint32_t reps,
                gs_tuning* best, gs_tune_entry* table, uint32_t table_len,
                uint32_t* entries);

/* readStats / getOutputStatus (TiledGaussianPipeline.ts:353-366;
 * IntersectionStage.ts:197-213; GaussianPass.ts:989-1008). */
int gs_read_state(gs_ctx* ctx, gs_state* out);

/* resizeIntersections (TiledGaussianPipeline.ts:220-231). */
int gs_resize_intersections(gs_ctx* ctx, uint32_t capacity);

/* Copy the last frame's color (W*H*4 binary16) / depth (W*H f32) to host. */
int gs_read_image(gs_ctx* ctx, void* color, size_t color_bytes, void* depth,
                  size_t depth_bytes);

/* getResources as debug copies (TiledGaussianPipeline.ts:391-406): names
 * projected_mean, projected_conic, projected_color, tile_counts,
 * visible_offsets, depth_records, ordered_counts, intersection_offsets,
 * sorted_records, tile_offsets. *size receives the byte size; with
 * dst == NULL only the size is returned. */
int gs_read_debug(gs_ctx* ctx, const char* name, void* dst, size_t bytes,
                  size_t* size);

/* The HIP stream the context renders on (a hipStream_t), and its device
 * output images, for callers that chain their own kernels. */
void* gs_stream(gs_ctx* ctx);
int gs_device_image(gs_ctx* ctx, void** color, void** depth);

/* Per-stage GPU time with HIP events recorded on the context's stream
 * (no reference counterpart; the reference's timestamp queries feed the
 * sandbox's KernelTimingInspector). While profiling, frames launch stage by
 * stage instead of as a graph. Stages: projection, visible scan+compaction,
 * depth sort, ordered scan, emission, tile sort, tile ranges, raster. */
#define GS_STAGE_COUNT 8
int gs_profile_begin(gs_ctx* ctx, uint32_t max_frames);
/* Synchronises; stage_ms receives the per-stage sums over the recorded
 * frames, frame_ms the sum of whole-frame times, heavy_ms the sum of the
 * heavy-unit raster kernel's times (events on its own stream around each of
 * its per-shape launches, the longest of which per frame is the one that
 * ran; the others exit at once), *frames the count. */
int gs_profile_end(gs_ctx* ctx, double stage_ms[GS_STAGE_COUNT], double* frame_ms,
                   double* heavy_ms, uint32_t* frames);

/* Capacity growth rule (IntersectionCapacityPolicy.ts:90-101): returns the
 * next capacity, or 0 when demand exceeds maximum. */
uint32_t gs_capacity_next(uint32_t current, uint32_t demand, uint32_t maximum,
                          double growth, double headroom);

/* Learned cull (no reference counterpart; the reference's nearest
 * mechanisms are the opacity and subpixel culls, ProjectionStage.ts:533,
 * 572-587). A 16-64-64-1 ReLU MLP with bf16 weights, run on the matrix
 * cores inside projection, predicts the log of each visible Gaussian's
 * largest alpha * T; with the cull enabled, Gaussians predicted below
 * log(threshold) are dropped before duplication and sort, in gs_render and
 * gs_render_batch alike. Parameters (GS_CULL_PARAMS f32): W1[64][16],
 * b1[64], W2[64][64], b2[64], w3[64], b3. Features (GS_CULL_FEATURES f32):
 * log pixel radius, log conic a, conic b / sqrt(a c), log conic c, opacity,
 * object-space position . view direction, log view depth, object-space view direction (3), object-space
 * position (3), log SH energy, x / width, y / height. */
#define GS_CULL_PARAMS 5313
#define GS_CULL_FEATURES 16
/* He-uniform weights from seed, identity normalisation, threshold 1e-4,
 * Adam state cleared. Required before the other gs_cull_* calls. */
int gs_cull_init(gs_ctx* ctx, uint64_t seed);
int gs_cull_set_params(gs_ctx* ctx, const float* params);
int gs_cull_get_params(gs_ctx* ctx, float* params);
/* Inputs are standardised as (f - mu) * inv_sigma before the network. */
int gs_cull_set_normalization(gs_ctx* ctx, const float mu[16], const float inv_sigma[16]);
int gs_cull_get_normalization(gs_ctx* ctx, float mu[16], float inv_sigma[16]);
/* Switch the cull in projection on or off; threshold in (0, 1) on alpha*T
 * (also the threshold the training loss weighs under-predictions around). */
int gs_cull_enable(gs_ctx* ctx, int enabled, float threshold);
/* Training set of up to `capacity` samples (cleared). */
int gs_cull_reserve(gs_ctx* ctx, uint32_t capacity);
/* Render `frame` on the single-camera path with the plain pipeline,
 * chunking off and every blended alpha * T recorded per Gaussian, then
 * append a hashed fraction `rate` of its visible Gaussians (features, log
 * max alpha * T clamped to [-16, 0]). *state receives the frame's state;
 * an overflowing frame appends nothing. */
int gs_cull_collect(gs_ctx* ctx, const gs_frame* frame, float rate, uint32_t seed,
                    gs_state* state);
int gs_cull_dataset_size(gs_ctx* ctx, uint32_t* n);
int gs_cull_read_dataset(gs_ctx* ctx, uint32_t first, uint32_t n, float* features,
                         float* targets);
/* Standardisation from the training set's feature means and deviations. */
int gs_cull_fit_normalization(gs_ctx* ctx);
/* `steps` Adam steps (beta 0.9 / 0.999) on minibatches of `batch` samples;
 * loss mean(w (y - t)^2) with w = 1 + under_weight * exp(t) where the
 * target t is above log(threshold) and the prediction y below t, else 1.
 * *last_loss: the last step's. */
int gs_cull_train(gs_ctx* ctx, uint32_t steps, uint32_t batch, float lr, float under_weight,
                  uint64_t seed, float* last_loss);
/* Predictions for n host feature rows (raw features), and the training
 * loss gradient over n host samples (parameter order, no update). */
int gs_cull_eval(gs_ctx* ctx, uint32_t n, const float* features, float* out);
int gs_cull_gradient(gs_ctx* ctx, uint32_t n, const float* features, const float* targets,
                     float under_weight, float* grad, float* loss);
/* Training checkpoint (no counterpart in the reference; training state is
 * the learned cull's own): the f32 parameters, Adam's first and second
 * moments and step count, the input standardisation and the threshold, as
 * one blob of GS_CULL_CHECKPOINT_BYTES bytes (a 32-byte header: magic
 * "GSCULL01", u32 version 1, u32 GS_CULL_PARAMS, u32 GS_CULL_FEATURES, u32
 * Adam steps taken, f32 threshold (gs_cull_enable's last), u32 0; then f32
 * params, m, v, mu, inv_sigma; host byte order). Loading one resumes training where the
 * saved context stopped; a blob of another size, version or network, or
 * with non-finite values, is rejected with GS_E_RANGE and nothing changes.
 * The dataset and the enabled flag are not part of it. */
#define GS_CULL_CHECKPOINT_BYTES (32 + (3 * GS_CULL_PARAMS + 2 * GS_CULL_FEATURES) * 4)
int gs_cull_save(gs_ctx* ctx, void* dst, size_t bytes);
int gs_cull_load(gs_ctx* ctx, const void* src, size_t bytes);

/* One frame without allocating (no counterpart in the reference): for
 * callers that cannot call hipMalloc -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation. It renders what gs_render
 * renders, with the same kernels and bits, but takes the scene
 * (data->memory == GS_MEMORY_DEVICE; active_slots, when given, must hold
 * slots below data->count), this camera's object table
 * (frame->object_frames: object_count * 40 f32 in device memory) and the
 * output images (out->memory == GS_MEMORY_DEVICE; depth with
 * config->output_depth) from the caller, and keeps every working array of
 * the frame in one library-owned device array of GS_NOALLOC_SCRATCH_BYTES
 * per GPU, held from library load until the process ends. config supplies
 * the options, the device and the intersection capacity
 * (config->intersection_capacity); the frame's working set, about
 * 128 B per active Gaussian + 16 B per intersection of capacity + 20 B per
 * tile pixel per raster chunk task, must fit the array, else GS_E_RANGE.
 * Kernels are enqueued on `stream` (a hipStream_t; NULL is the null stream)
 * and nothing is allocated or freed. With state == NULL the call is
 * asynchronous (and can be captured); otherwise it synchronises the stream
 * and fills *state. An overflowing frame draws nothing. Every call on one
 * GPU uses the same array, which keeps the last call's contents: calls that
 * overlap on one GPU race. Returns GS_OK or GS_E_RANGE / GS_E_TYPE /
 * GS_E_DEVICE (no message is kept: there is no context). */
#define GS_NOALLOC_SCRATCH_BYTES (32u << 20)
int gs_render_noalloc(const gs_config* config, const gs_data_desc* data, uint32_t object_count,
                      const gs_frame* frame, const gs_image* out, void* stream, gs_state* state);

#ifdef __cplusplus
}
#endif

#endif /* GS_H */
