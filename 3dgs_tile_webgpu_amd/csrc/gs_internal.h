// gs_internal.h - types shared by the gfx950 kernels and the host driver.
#ifndef GS_INTERNAL_H
#define GS_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gs {

struct CullNet;

constexpr int kBlock = 256;          // threads per block (4 wave64)
constexpr int kScanItems = 8;        // elements per thread in a scan
constexpr int kScanPartition = kBlock * kScanItems;  // 2048 elements per scan block
#ifndef GS_SORT_THREADS
#define GS_SORT_THREADS 512
#endif
constexpr int kSortThreads = GS_SORT_THREADS;  // threads of a radix-pass block
constexpr int kSortItems = 8;        // keys per thread in a radix pass
constexpr int kSortPartition = kSortThreads * kSortItems;  // keys per sort block
constexpr int kMaxRadixBits = 8;     // widest digit of the LSD sorts
constexpr int kMaxRadix = 1 << kMaxRadixBits;
constexpr int kMaxTile = 32;         // widest tile edge (pixels)
constexpr int kViewGroupMax = 8;     // views of one group launch
constexpr int kMaxSortPasses = 8;    // 32-bit keys with 4-bit digits
// Digit totals of the two sorts of a frame (depth, then tile): each sort has
// kMaxSortPasses rows of 256 totals, each row held as kHistShards partial
// copies (a block adds into copy blockIdx % kHistShards), so that no total
// takes more than a few same-address atomics per block count.
constexpr int kHistShards = 8;
constexpr size_t kSortTotalsWords = (size_t)2 * kMaxSortPasses * kHistShards * 256;

// The part of a frame's uniforms that changes with the camera
// (FrameUniforms.ts:41-58): it lives in device memory and is uploaded
// before each launch, so a captured frame graph survives camera motion.
// The view's object frame table (ObjectFrameState.ts:97-120, 40 floats per
// object) follows it in the same allocation.
struct FrameCamera {
  float projection[16];
  float near_plane, far_plane;
  uint32_t attempt, reset_valid;
};
constexpr size_t kCameraBytes = 256;  // FrameCamera slot, padded

// Per-frame constants of a launch (FrameUniforms.ts:41-58 minus the camera,
// plus the stage constants the reference keeps in uniforms).
struct FrameParams {
  const FrameCamera* camera;
  float viewport[2];  // width, height
  float inv_viewport[2];  // 1 / width, 1 / height (IEEE quotients, host-computed)
  float filter_variance;
  float transmittance_threshold;
  float background[4];
  float depth_alpha_threshold;
  float invalid_depth_value;
  uint32_t boundary_axis;
  uint32_t active_count;
  uint32_t tiles_x, tiles_y, tile_count;
  uint32_t width, height;
  uint32_t capacity;
  uint32_t attempt, reset_valid;
  uint32_t object_count;
  uint32_t sh_degree;
  uint32_t depth_mode;             // 0 float32 keys, 1 packed16
  uint32_t antialias_compensated;
  uint32_t subpixel_culling;
  uint32_t chunk_size;             // 0 = no chunking
  uint32_t has_active_slots;
  uint32_t output_depth;
  uint32_t raster_stats;
  uint32_t tile_size;               // tile edge: 8, 16 or 32 (16 matches constants.ts:1)
  uint32_t split_units;             // heavy units above which 16x16 tiles leave 4 lanes per pixel
  uint32_t fine_split_units;        // ... above which row slabs have 4 rows, not 8
  uint32_t pair_units;              // ... above which whole tiles with 2 lanes give way to slabs
};

// Device-side counters of one frame (dispatch state + outputState).
struct DeviceState {
  uint32_t visible;   // V
  uint32_t total;     // requested K (saturating)
  uint32_t count;     // K published (0 on overflow)
  uint32_t overflow;
  uint32_t out_attempt, out_version, out_demand, out_overflow;
  uint32_t chunk_tasks;  // raster chunk tasks of this frame
  uint32_t heavy_units;  // raster heavy work-list length of this frame
  unsigned long long checked, blended;
  unsigned long long checked_heavy;  // the part of `checked` in heavy units
  uint32_t sort_error;  // a look-back exceeded its wait bound: bit 0 in a
                        // radix pass, bit 1 in a scan
  uint32_t heavy_split;  // this frame's heavy shape: 0 whole tiles, 1 8-row, 2 4-row slabs
  uint32_t light_units;  // row-slab frames: light tiles listed after the heavy units
  uint32_t heavy_next[8];  // per XCD: heavy work items taken beyond the first round
  uint32_t depth_in_place;  // the depth sort's last digit was one value for every key: its
                            // pass was skipped and the result is that pass's input buffer
};

struct SceneView {
  const float4* means;
  const uint2* scales_opacity;
  const uint32_t* rotations;
  const uint32_t* sh;
  const uint32_t* active_slots;
  const float* objects;
  // Per slot, the largest rgb8e8 exponent byte of the higher SH bands (a
  // learned-cull feature; static with the scene), or null to read the SH.
  const uint8_t* sh_top;
};

// A visible Gaussian's projection in one 64-byte line, so that the raster's
// gather of a key (by active ordinal, in depth order within a tile, hence
// scattered) reads one line instead of three: mean (cx, cy, depth,
// opacity), conic (a, b, c, rX + 512 rY), colour (r, g, b, slot bits) and
// aux = (ln(255 opacity), the raster's per-Gaussian cutoff
// (TileRasterizer.ts:579-581); the row and column terms and the shared
// slack coefficient of the raster's support bound), computed once by the
// projection.
struct alignas(64) Splat {
  float4 mean, conic, color, aux;
};
static_assert(sizeof(Splat) == 64, "one cache line per splat");

struct ProjectedView {
  float4* mean;           // per ordinal, every active Gaussian (visibility, depth)
  Splat* splat;           // per ordinal, written for visible Gaussians only
  uint32_t* tile_counts;  // per ordinal
  uint2* hits;            // per ordinal: 8-byte tile-hit record (see kHitMasked)
};

// Length of a device-sized sequence: (*ptr rounded up to a multiple of
// 2^log2_div, divided by it) * mul, or `constant` when ptr is null.
struct LenSpec {
  const uint32_t* ptr;
  uint32_t log2_div;  // 0 = no division
  uint32_t mul;
  uint32_t constant;
};

enum ScanMode : uint32_t {
  kScanPlain = 0,
  kScanVisible = 1,        // total -> state.visible
  kScanIntersections = 2,  // total -> state.total, count, overflow, output
};

struct FrameBuffers {
  FrameParams params;
  DeviceState* state;
  SceneView scene;
  ProjectedView proj;
  uint32_t* visible_offsets;
  uint2* depth_a;
  uint2* depth_b;
  uint32_t* ordered_counts;
  uint2* ordered_hits;  // tile-hit records in depth order
  uint32_t* intersection_offsets;
  uint2* records_a;
  uint2* records_b;
  uint32_t* tile_offsets;
  uint32_t* hist;       // sort workspace: kSortTotalsWords digit totals
                        // (zero before the first frame), then 2 x
                        // sort_parts x 256 partition status words
  uint32_t sort_parts;  // partitions the status rows are sized for
  uint32_t* partials;   // scan workspace: one 64-bit status word per scan
                        // partition (zero between scans)
  uint32_t* chunk_first;  // per tile: first chunk task (tile capacity)
  uint2* tasks;           // raster units: (tile, chunk or ~0u); heavy units,
                          // then (row-slab frames) the non-empty light tiles
  float4* chunk_partials; // per task and pixel: (C.rgb, T)
  float* partial_depth;   // per task and pixel: weighted depth
  uint32_t task_capacity;  // chunk tasks (partials slots)
  uint32_t unit_capacity;  // raster units (chunk tasks + heavy tiles + tiles)
  uint16_t* color_out;  // W*H*4 binary16
  float* depth_out;     // W*H f32 or null
  uint32_t max_n;       // slot capacity of the per-Gaussian arrays
  uint32_t max_k;       // intersection capacity
  uint32_t max_tiles;
  uint32_t depth_passes;
  uint32_t tile_passes;
  uint32_t radix_bits;   // digit width: 4, 6 or 8
  uint32_t raster_batch; // records staged per light-tile round: 128 or 256
  uint32_t cu_count;     // compute units of the device (persistent grids)
  // Learned cull: the network run inside projection (or null), and for
  // training-set collection the per-Gaussian features written by projection
  // and the per-Gaussian max alpha * T recorded by the raster (or null).
  const CullNet* cull_net;
  float* cull_feats;
  uint32_t* record;
};

// Frame stages, in launch order; launch_frame can record an event before
// each stage and one after the last (kStageCount + 1 events), then four on
// the raster's second stream: before the heavy-unit launches and after each
// of the three (kMarks in all).
enum Stage : int {
  kStageProject = 0,
  kStageVisible,
  kStageDepthSort,
  kStageScan,
  kStageEmit,
  kStageTileSort,
  kStageTileOffsets,
  kStageRaster,
  kStageCount
};
// Then a pair (before, after) around each heavy shape's launch, on the
// stream it runs on.
constexpr int kMarkHeavyBegin = kStageCount + 1;
constexpr int kMarkHeavyShapes = 4;
constexpr int kMarks = kMarkHeavyBegin + 2 * kMarkHeavyShapes;

// The raster's second branch: the heavy-unit kernels run on aux, forked
// from the frame's stream at `fork` and joined back into it at `join`,
// concurrently with the light-tile kernel on the frame's stream.
struct RasterStreams {
  hipStream_t aux;
  hipEvent_t fork, join;
};

// Enqueue one frame on `stream` (all kernels; sizes read on the device).
// `marks` is null or kMarks events (stage boundaries, heavy-unit kernels).
hipError_t launch_frame(const FrameBuffers& b, hipStream_t stream, const RasterStreams& rs,
                        hipEvent_t* marks);
// Enqueue the frames of a group of up to kViewGroupMax views on `stream`,
// each stage one launch over all of them. The views share the launch shape,
// the mode (plain, cull, collect), the scene and their sort pass counts
// (hipErrorInvalidValue otherwise); each has its own working buffers,
// state and outputs. launch_frame is the group of one.
// The per-slot SH exponent bytes of SceneView::sh_top for n slots.
hipError_t launch_sh_top(const uint32_t* sh, uint32_t n, uint32_t sh_degree, uint8_t* out,
                         hipStream_t stream);
hipError_t launch_group(const FrameBuffers* views, uint32_t n, hipStream_t stream,
                        const RasterStreams& rs, hipEvent_t* marks);
// Location of the sorted depth / tile records after launch_frame
// (depth_in_place: DeviceState::depth_in_place of the frame).
const uint2* sorted_depth_records(const FrameBuffers& b, uint32_t depth_in_place);
const uint2* sorted_tile_records(const FrameBuffers& b);

// Learned-cull training and evaluation launches (gs_cull.hip).
hipError_t cull_collect(const float4* pmean, const float* feats, const uint32_t* record,
                        uint32_t active, uint32_t keep, uint32_t seed, float* ds_feats,
                        float* ds_targets, uint32_t* count, uint32_t capacity, hipStream_t s);
hipError_t cull_moments(const float* feats, uint32_t n, double* out, hipStream_t s);
hipError_t cull_train_step(const float* feats, const float* targets, uint32_t n, uint32_t perm_a,
                           uint32_t perm_b, uint32_t start, uint32_t batch, const CullNet* net,
                           float under_weight, float* grad, float* loss, hipStream_t s);
hipError_t cull_adam(float* params, float* m, float* v, float* grad, float lr, float beta1,
                     float beta2, float eps, float corr1, float corr2, CullNet* net,
                     hipStream_t s);
hipError_t cull_eval(const float* feats, uint32_t n, const CullNet* net, float* out,
                     hipStream_t s);

}  // namespace gs

#endif
