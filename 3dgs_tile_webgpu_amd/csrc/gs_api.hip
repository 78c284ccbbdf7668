// gs_api.hip - context, device memory and the C ABI of include/gs.h.
//
// The context plays the role of the reference's TiledGaussianPipeline
// (pipeline/TiledGaussianPipeline.ts:39-484): it owns every device buffer,
// grows them geometrically and never shrinks them (AttributePool.ts:41-67),
// and enqueues frames on its own HIP stream.
//
// Frames run on lanes: a lane is one frame's working set (projected arrays,
// records, histograms, tile ranges, raster partials) plus a stream pair.
// Lane 0 uses the context's stream. A view set is a group of frames
// launched together -- the single camera of gs_render, or the cameras of
// gs_render_batch spread over lanes -- with one DeviceState per view, one
// device table of per-view camera uniforms and object frames, and one
// captured graph that is replayed while the launch description is
// unchanged. The camera table is uploaded before every launch, so moving a
// camera never re-captures.
#include <math.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gs.h"
#include "gs_cull.h"
#include "gs_internal.h"

using gs::DeviceState;
using gs::FrameBuffers;
using gs::FrameCamera;
using gs::FrameParams;

namespace {

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct Lane {
  hipStream_t stream = nullptr;  // lane 0: the context stream (not owned)
  hipStream_t aux = nullptr;     // second raster branch, forked per frame
  hipEvent_t fork = nullptr, join = nullptr, done = nullptr;
  // Per-Gaussian frame arrays, sized for slot_capacity Gaussians.
  uint32_t slot_capacity = 0;
  DeviceBuffer pmean, splat, tile_counts, hits, visible_offsets, depth_a, depth_b,
      ordered_counts, ordered_hits, intersection_offsets;
  // Intersection records, sized for `capacity`.
  uint32_t capacity = 0;
  DeviceBuffer records_a, records_b;
  // Radix sort workspace and scan partials, sized for max(slots, capacity).
  uint32_t sort_capacity = 0, sort_parts = 0;
  DeviceBuffer hist, partials;
  // Tile ranges, sized for tile_capacity + 1.
  uint32_t tile_capacity = 0;
  DeviceBuffer tile_offsets, chunk_first;
  // Raster heavy units and chunk partials (RasterChunkWorkspace.ts:12-59).
  DeviceBuffer tasks, partials_px, partial_depth;
};

struct ViewSet {
  uint32_t n = 0;
  DeviceBuffer state;    // n DeviceState
  DeviceBuffer dynamic;  // n * stride bytes: FrameCamera, then object frames
  size_t stride = 0;
  std::vector<uint32_t> capacity;  // intersection capacity per view
  std::vector<uint32_t> lane_of;
  // View groups (gs_config.view_group): the views in launch order, groups
  // being consecutive runs of view_group of them; each view's working set
  // (0: its lane's, j > 0: the context's group slot j - 1).
  uint32_t group = 1;
  std::vector<uint32_t> order, slot_of;
  std::vector<DeviceBuffer> color, depth;  // library-owned images per view
  std::vector<FrameBuffers> frames;        // last launch description
  std::vector<uint32_t> width, height;
  // Captured graph and the description it was captured from.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<FrameBuffers> key;
  std::vector<uint32_t> key_lanes, key_order, key_slots;
  bool graph_valid = false;
  bool rendered = false;
};

// The learned cull's device state: the network the kernels read, the fp32
// parameters with their Adam moments and gradient, the instrumented frame's
// features and contribution record, and the training set.
struct CullState {
  DeviceBuffer net, params, m, v, grad, loss, moments;
  DeviceBuffer feats, record;
  DeviceBuffer ds_feats, ds_targets, ds_count;
  DeviceBuffer tmp_in, tmp_target, tmp_out;
  uint32_t ds_capacity = 0;
  uint32_t step = 0;
  float threshold = 1e-4f;  // the last gs_cull_enable's (alpha * T)
  bool ready = false;
  bool enabled = false;
  bool collecting = false;
};

}  // namespace

struct gs_ctx {
  gs_config cfg;
  uint32_t cu_count = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  hipEvent_t batch_fork = nullptr;
  std::string error;
  // Scene (GaussianData contract).
  uint32_t count = 0, active_count = 0, sh_degree = 0;
  bool has_data = false, has_active = false;
  DeviceBuffer means, scales_opacity, rotations, sh, active;
  DeviceBuffer sh_top;  // per slot: the higher SH bands' largest exponent byte
  // Object frame table set by gs_set_objects (host copy).
  std::vector<float> objects_host;
  uint32_t object_count = 0;
  // Intersection capacity of the single-camera path.
  uint32_t capacity = 0;
  std::vector<Lane> lanes;
  // Working sets of the second and later views of a view group (no streams:
  // a group's views are launched together on the context stream).
  std::vector<Lane> slots;
  ViewSet single, batch;
  // Camera-table staging: a ring of pinned slots.
  static constexpr int kRing = 8;
  void* ring[kRing] = {};
  hipEvent_t ring_done[kRing] = {};
  size_t ring_bytes = 0;
  int ring_next = 0;
  std::vector<unsigned char> staging;
  // Stage profiling (single-camera path).
  std::vector<hipEvent_t> marks;
  uint32_t profile_capacity = 0, profile_frames = 0;
  CullState cull;
};

namespace {

int fail(gs_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return code;
}

int hip_fail(gs_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, GS_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Wait for everything this context enqueued: the context stream, each
// lane's stream and each raster branch stream.
hipError_t sync_streams(gs_ctx* ctx) {
  hipError_t e = hipStreamSynchronize(ctx->stream);
  for (const Lane& l : ctx->lanes) {
    if (e == hipSuccess && l.stream && l.stream != ctx->stream) e = hipStreamSynchronize(l.stream);
    if (e == hipSuccess && l.aux) e = hipStreamSynchronize(l.aux);
  }
  return e;
}

// Grow-only device allocation (AttributePool.ts:41-67 semantics). A buffer
// is freed only once every stream of the context has drained.
int ensure(gs_ctx* ctx, DeviceBuffer& b, size_t bytes, const char* what) {
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return GS_OK;
  if (b.ptr) {
    hipError_t e = sync_streams(ctx);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.ptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, GS_E_OOM, std::string("allocation-failed: ") + what);
  }
  b.bytes = bytes;
  return GS_OK;
}

void release(DeviceBuffer& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t next_pow2(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

uint32_t bits_for(uint32_t v) {  // bits needed to represent v
  uint32_t b = 0;
  while (b < 32 && (v >> b) != 0) b++;
  return b;
}

uint32_t grow_to(uint32_t current, uint32_t need) {
  uint32_t cap = current ? current : 1024u;
  while (cap < need) cap = cap > 0x7fffffffu ? need : cap * 2u;
  return cap;
}

int lane_slots(gs_ctx* ctx, Lane& l, uint32_t n) {
  if (n <= l.slot_capacity && l.pmean.ptr) return GS_OK;
  uint32_t cap = grow_to(l.slot_capacity, n);
  int rc;
  size_t c = cap;
  if ((rc = ensure(ctx, l.pmean, c * 16, "projected-mean")) ||
      (rc = ensure(ctx, l.splat, c * sizeof(gs::Splat), "projected-splats")) ||
      (rc = ensure(ctx, l.tile_counts, c * 4, "tile-counts")) ||
      (rc = ensure(ctx, l.hits, c * 8, "tile-hits")) ||
      (rc = ensure(ctx, l.ordered_hits, c * 8, "ordered-tile-hits")) ||
      (rc = ensure(ctx, l.visible_offsets, c * 4, "visible-offsets")) ||
      (rc = ensure(ctx, l.depth_a, c * 8, "depth-records-a")) ||
      (rc = ensure(ctx, l.depth_b, c * 8, "depth-records-b")) ||
      (rc = ensure(ctx, l.ordered_counts, c * 4, "ordered-tile-counts")) ||
      (rc = ensure(ctx, l.intersection_offsets, c * 4, "intersection-offsets")))
    return rc;
  l.slot_capacity = cap;
  return GS_OK;
}

int lane_records(gs_ctx* ctx, Lane& l, uint32_t capacity) {
  if (capacity <= l.capacity && l.records_a.ptr) return GS_OK;
  int rc;
  if ((rc = ensure(ctx, l.records_a, (size_t)capacity * 8, "records-a")) ||
      (rc = ensure(ctx, l.records_b, (size_t)capacity * 8, "records-b")))
    return rc;
  l.capacity = capacity;
  return GS_OK;
}

int lane_sort(gs_ctx* ctx, Lane& l) {
  uint32_t m = std::max(l.slot_capacity, l.capacity);
  if (m <= l.sort_capacity && l.hist.ptr) return GS_OK;
  uint32_t parts = div_up(m, gs::kSortPartition);
  if (parts == 0) parts = 1;
  uint32_t scan_parts = div_up(m, gs::kScanPartition);
  int rc;
  // Sort workspace: the digit totals of the depth and tile sorts, then two
  // sets of partition status rows (passes alternate between them). The
  // totals start at zero; afterwards each sort's counting kernel zeroes the
  // other sort's totals.
  size_t sort_words = gs::kSortTotalsWords + (size_t)2 * parts * 256;
  // Scan workspace: a 64-bit status word per partition. It starts at zero;
  // the radix sort's counting kernel zeroes it again after each scan.
  size_t scan_bytes = (size_t)8 * (scan_parts + 1);
  if ((rc = ensure(ctx, l.hist, sort_words * 4, "radix-sort-workspace")) ||
      (rc = ensure(ctx, l.partials, scan_bytes, "scan-workspace")))
    return rc;
  // On the context stream: every lane's work forks from it after this.
  hipError_t e = hipMemsetAsync(l.partials.ptr, 0, l.partials.bytes, ctx->stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(l.hist.ptr, 0, gs::kSortTotalsWords * 4, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipMemsetAsync");
  l.sort_capacity = m;
  l.sort_parts = parts;
  return GS_OK;
}

int lane_create(gs_ctx* ctx, Lane& l, bool own_stream) {
  // Default stream priority only.
  if (own_stream) {
    if (hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) != hipSuccess)
      return fail(ctx, GS_E_DEVICE, "hipStreamCreate");
  } else {
    l.stream = ctx->stream;
  }
  if (hipStreamCreateWithFlags(&l.aux, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&l.fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&l.join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&l.done, hipEventDisableTiming) != hipSuccess)
    return fail(ctx, GS_E_DEVICE, "lane stream/event creation");
  return GS_OK;
}

void lane_destroy(gs_ctx* ctx, Lane& l) {
  DeviceBuffer* all[] = {&l.pmean,     &l.splat,          &l.tile_counts,  &l.hits,
                         &l.ordered_hits,
                         &l.visible_offsets, &l.depth_a,  &l.depth_b,      &l.ordered_counts,
                         &l.intersection_offsets, &l.records_a, &l.records_b, &l.hist,
                         &l.partials,  &l.tile_offsets,   &l.chunk_first,  &l.tasks,
                         &l.partials_px, &l.partial_depth};
  for (DeviceBuffer* b : all) release(*b);
  if (l.fork) (void)hipEventDestroy(l.fork);
  if (l.join) (void)hipEventDestroy(l.join);
  if (l.done) (void)hipEventDestroy(l.done);
  if (l.aux) (void)hipStreamDestroy(l.aux);
  if (l.stream && l.stream != ctx->stream) (void)hipStreamDestroy(l.stream);
  l = Lane();
}

void set_release(ViewSet& v) {
  release(v.state);
  release(v.dynamic);
  for (auto& b : v.color) release(b);
  for (auto& b : v.depth) release(b);
  if (v.exec) (void)hipGraphExecDestroy(v.exec);
  if (v.graph) (void)hipGraphDestroy(v.graph);
  v = ViewSet();
}

int copy_in(gs_ctx* ctx, DeviceBuffer& dst, const void* src, size_t bytes, uint32_t memory,
            const char* what) {
  int rc = ensure(ctx, dst, bytes, what);
  if (rc) return rc;
  if (bytes == 0) return GS_OK;
  hipError_t e;
  if (memory == GS_MEMORY_DEVICE) {
    e = hipMemcpyAsync(dst.ptr, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
  } else {
    // Host arrays are copied during the call; nothing is retained.
    e = hipMemcpyAsync(dst.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  return GS_OK;
}

bool finite(float v) { return v == v && fabsf(v) <= 3.402823466e38f; }

bool valid_tuning(uint32_t tile, uint32_t bits, uint32_t batch) {
  return (tile == 8 || tile == 16 || tile == 32) && (bits == 4 || bits == 6 || bits == 8) &&
         (batch == 128 || batch == 256);
}

}  // namespace

extern "C" void gs_config_default(gs_config* c) {
  memset(c, 0, sizeof(*c));
  c->device = 0;
  c->depth_sort_mode = GS_DEPTH_FLOAT32;
  c->antialias_mode = GS_AA_COMPENSATED;
  c->subpixel_sample_culling = 1;
  c->raster_chunk_size = 8192;
  c->intersection_capacity = 262144;
  c->max_intersection_capacity = 1u << 28;
  c->output_depth = 0;
  c->transmittance_threshold = 1e-4f;
  c->depth_alpha_threshold = 0.95f;
  c->invalid_depth_value = 1.0f;
  c->raster_stats = 0;
  c->capture_graph = 1;
  c->tile_size = 16;
  c->radix_bits = 8;
  c->raster_batch = 256;
  c->batch_lanes = 1;
  c->view_group = 1;
  c->heavy_split_units = GS_SPLIT_AUTO;
  c->heavy_fine_split_units = GS_SPLIT_AUTO;
  c->heavy_pair_units = GS_SPLIT_AUTO;
}

// GaussianPass.ts:157-260 option validation.
static int validate_config(const gs_config* c, const char*& msg) {
  if (c->depth_sort_mode > GS_DEPTH_PACKED16) {
    msg = "depthSortMode must be float32 or packed16";
    return GS_E_TYPE;
  }
  if (c->antialias_mode > GS_AA_CLASSIC) {
    msg = "antialiasMode must be compensated or classic";
    return GS_E_TYPE;
  }
  if (!(c->transmittance_threshold > 0.0f && c->transmittance_threshold < 1.0f)) {
    msg = "rasterTransmittanceThreshold must be finite and in (0, 1)";
    return GS_E_RANGE;
  }
  // rasterChunks.ts:80-99 (validateRasterChunkSize).
  if (c->raster_chunk_size != 0 &&
      (c->raster_chunk_size < 256 || c->raster_chunk_size % 256 != 0)) {
    msg = "rasterChunkSize must be a multiple of 256 and at least 256";
    return GS_E_RANGE;
  }
  if (c->intersection_capacity < 1 || c->max_intersection_capacity < 1) {
    msg = "intersection capacity must be a positive integer";
    return GS_E_RANGE;
  }
  if (c->intersection_capacity > c->max_intersection_capacity) {
    msg = "initialIntersectionCapacity exceeds maximum capacity";
    return GS_E_RANGE;
  }
  if (c->max_intersection_capacity > 0x3fffffffu) {
    msg = "Intersection capacity exceeds device buffer or dispatch limits";
    return GS_E_RANGE;
  }
  if (!(c->depth_alpha_threshold >= 0.0f && c->depth_alpha_threshold <= 1.0f)) {
    msg = "depthAlphaThreshold must be in [0, 1]";
    return GS_E_RANGE;
  }
  for (int i = 0; i < 4; i++)
    if (!finite(c->background[i])) {
      msg = "background must be finite";
      return GS_E_RANGE;
    }
  if (!valid_tuning(c->tile_size, c->radix_bits, c->raster_batch)) {
    msg = "tile size must be 8, 16 or 32, radix bits 4, 6 or 8, raster batch 128 or 256";
    return GS_E_RANGE;
  }
  // Concurrent lanes inside one captured batch graph are switched off: the
  // four-lane batch ended in a host segmentation fault inside
  // gs_render_batch whose cause is not established.
  if (c->batch_lanes != 1) {
    msg = "batch lanes above 1 are disabled";
    return GS_E_RANGE;
  }
  if (c->view_group < 1 || c->view_group > (uint32_t)gs::kViewGroupMax) {
    msg = "viewGroup must be an integer in [1, 8]";
    return GS_E_RANGE;
  }
  return GS_OK;
}

extern "C" int gs_create(const gs_config* config, gs_ctx** out) {
  if (!out) return GS_E_RANGE;
  *out = nullptr;
  gs_config c;
  if (config) c = *config;
  else gs_config_default(&c);
  const char* msg = nullptr;
  int rc = validate_config(&c, msg);
  if (rc) return rc;
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
    (void)hipGetLastError();
    return GS_E_DEVICE;
  }
  if (c.device < 0 || c.device >= devices) return GS_E_RANGE;
  if (hipSetDevice(c.device) != hipSuccess) return GS_E_DEVICE;
  gs_ctx* ctx = new gs_ctx();
  ctx->cfg = c;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) ==
          hipSuccess &&
      cus > 0)
    ctx->cu_count = (uint32_t)cus;
  // Default stream priority only.
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->batch_fork, hipEventDisableTiming) != hipSuccess) {
    gs_destroy(ctx);
    return GS_E_DEVICE;
  }
  ctx->lanes.resize(c.batch_lanes);
  for (uint32_t i = 0; i < c.batch_lanes; i++)
    if ((rc = lane_create(ctx, ctx->lanes[i], i != 0))) {
      gs_destroy(ctx);
      return GS_E_DEVICE;
    }
  ctx->capacity = c.intersection_capacity;
  if ((rc = lane_records(ctx, ctx->lanes[0], ctx->capacity))) {
    gs_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return GS_OK;
}

extern "C" void gs_destroy(gs_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->cfg.device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DeviceBuffer* all[] = {&ctx->means, &ctx->scales_opacity, &ctx->rotations, &ctx->sh,
                         &ctx->active, &ctx->sh_top};
  for (DeviceBuffer* b : all) release(*b);
  set_release(ctx->single);
  set_release(ctx->batch);
  {
    CullState& c = ctx->cull;
    DeviceBuffer* cb[] = {&c.net,  &c.params,   &c.m,          &c.v,        &c.grad,
                          &c.loss, &c.moments,  &c.feats,      &c.record,   &c.ds_feats,
                          &c.ds_targets, &c.ds_count, &c.tmp_in, &c.tmp_target, &c.tmp_out};
    for (DeviceBuffer* b : cb) release(*b);
  }
  for (Lane& l : ctx->lanes) lane_destroy(ctx, l);
  for (Lane& l : ctx->slots) lane_destroy(ctx, l);
  for (int i = 0; i < gs_ctx::kRing; i++) {
    if (ctx->ring[i]) (void)hipHostFree(ctx->ring[i]);
    if (ctx->ring_done[i]) (void)hipEventDestroy(ctx->ring_done[i]);
  }
  for (hipEvent_t e : ctx->marks) (void)hipEventDestroy(e);
  if (ctx->batch_fork) (void)hipEventDestroy(ctx->batch_fork);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" const char* gs_last_error(const gs_ctx* ctx) {
  return ctx ? ctx->error.c_str() : "no context";
}

extern "C" int gs_set_data(gs_ctx* ctx, const gs_data_desc* d) {
  if (!ctx || !d) return GS_E_RANGE;
  if (d->geometry_format != GS_GEOMETRY_COMPACT)
    return fail(ctx, GS_E_TYPE, "only the compact geometry format is supported");
  if (d->sh_format != GS_SH_RGB8E8)
    return fail(ctx, GS_E_TYPE, "only the rgb8e8 SH format is supported");
  if (d->sh_degree > 3) return fail(ctx, GS_E_RANGE, "shDegree must be 0..3");
  if (d->active_count > d->count)
    return fail(ctx, GS_E_RANGE, "activeCount exceeds the slot count");
  if (d->memory > GS_MEMORY_DEVICE) return fail(ctx, GS_E_TYPE, "unknown memory kind");
  if (d->count > 0 && (!d->means || !d->scales_opacity || !d->rotations || !d->sh))
    return fail(ctx, GS_E_RANGE, "missing packed buffer");
  (void)hipSetDevice(ctx->cfg.device);
  size_t n = d->count;
  size_t ncoef = (size_t)(d->sh_degree + 1) * (d->sh_degree + 1);
  int rc;
  if ((rc = copy_in(ctx, ctx->means, d->means, n * 16, d->memory, "means")) ||
      (rc = copy_in(ctx, ctx->scales_opacity, d->scales_opacity, n * 8, d->memory,
                    "scalesOpacity")) ||
      (rc = copy_in(ctx, ctx->rotations, d->rotations, n * 4, d->memory, "rotations")) ||
      (rc = copy_in(ctx, ctx->sh, d->sh, n * ncoef * 4, d->memory, "shCoefficients")) ||
      (rc = ensure(ctx, ctx->sh_top, n, "sh-exponents")))
    return rc;
  hipError_t se = gs::launch_sh_top((const uint32_t*)ctx->sh.ptr, (uint32_t)n, d->sh_degree,
                                    (uint8_t*)ctx->sh_top.ptr, ctx->stream);
  if (se != hipSuccess) return hip_fail(ctx, se, "sh exponents");
  if (d->active_slots && d->memory == GS_MEMORY_HOST)
    for (uint32_t i = 0; i < d->active_count; i++)
      if (d->active_slots[i] >= d->count)
        return fail(ctx, GS_E_RANGE, "active slot out of range");
  ctx->has_active = d->active_slots != nullptr;
  if (ctx->has_active &&
      (rc = copy_in(ctx, ctx->active, d->active_slots, (size_t)d->active_count * 4, d->memory,
                    "activeSlots")))
    return rc;
  ctx->count = d->count;
  ctx->active_count = d->active_count;
  ctx->sh_degree = d->sh_degree;
  ctx->has_data = true;
  ctx->single.rendered = false;
  ctx->batch.rendered = false;
  return GS_OK;
}

extern "C" int gs_update_active(gs_ctx* ctx, uint32_t start, uint32_t n, const uint32_t* slots,
                                uint32_t new_active_count) {
  if (!ctx) return GS_E_RANGE;
  if (!ctx->has_data) return fail(ctx, GS_E_STATE, "no data");
  if (new_active_count > ctx->count)
    return fail(ctx, GS_E_RANGE, "activeCount exceeds the slot count");
  if ((uint64_t)start + n > new_active_count)
    return fail(ctx, GS_E_RANGE, "active range exceeds activeCount");
  if (n > 0 && !slots) return fail(ctx, GS_E_RANGE, "missing slots");
  for (uint32_t i = 0; i < n; i++)
    if (slots[i] >= ctx->count) return fail(ctx, GS_E_RANGE, "active slot out of range");
  (void)hipSetDevice(ctx->cfg.device);
  int rc = GS_OK;
  if (!ctx->has_active) {
    // Materialise the identity list before a partial update.
    std::vector<uint32_t> ident(ctx->count);
    for (uint32_t i = 0; i < ctx->count; i++) ident[i] = i;
    if ((rc = copy_in(ctx, ctx->active, ident.data(), ident.size() * 4, GS_MEMORY_HOST,
                      "activeSlots")))
      return rc;
    ctx->has_active = true;
  } else if (ctx->active.bytes < (size_t)ctx->count * 4) {
    DeviceBuffer grown;
    if ((rc = ensure(ctx, grown, (size_t)ctx->count * 4, "activeSlots"))) return rc;
    hipError_t e = hipMemcpyAsync(grown.ptr, ctx->active.ptr, ctx->active.bytes,
                                  hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "activeSlots");
    release(ctx->active);
    ctx->active = grown;
  }
  if (n > 0) {
    hipError_t e = hipMemcpyAsync((uint32_t*)ctx->active.ptr + start, slots, (size_t)n * 4,
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "activeSlots");
  }
  ctx->active_count = new_active_count;
  return GS_OK;
}

extern "C" int gs_set_objects(gs_ctx* ctx, uint32_t object_count, const float* frames) {
  if (!ctx) return GS_E_RANGE;
  if (object_count > 0 && !frames) return fail(ctx, GS_E_RANGE, "missing object frames");
  ctx->objects_host.assign(frames, frames + (size_t)object_count * 40);
  ctx->object_count = object_count;
  return GS_OK;
}

// f64 4x4 helpers, column-major like Three.js Matrix4.elements.
static void mat4_mul(const double* a, const double* b, double* o) {
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 4; r++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += a[k * 4 + r] * b[c * 4 + k];
      o[c * 4 + r] = s;
    }
}

// General inverse by cofactors (Three.js Matrix4.invert); a singular matrix
// yields zeros, as Three.js does.
static void mat4_invert(const double* m, double* o) {
  double n11 = m[0], n21 = m[1], n31 = m[2], n41 = m[3];
  double n12 = m[4], n22 = m[5], n32 = m[6], n42 = m[7];
  double n13 = m[8], n23 = m[9], n33 = m[10], n43 = m[11];
  double n14 = m[12], n24 = m[13], n34 = m[14], n44 = m[15];
  double t11 = n23 * n34 * n42 - n24 * n33 * n42 + n24 * n32 * n43 - n22 * n34 * n43 -
               n23 * n32 * n44 + n22 * n33 * n44;
  double t12 = n14 * n33 * n42 - n13 * n34 * n42 - n14 * n32 * n43 + n12 * n34 * n43 +
               n13 * n32 * n44 - n12 * n33 * n44;
  double t13 = n13 * n24 * n42 - n14 * n23 * n42 + n14 * n22 * n43 - n12 * n24 * n43 -
               n13 * n22 * n44 + n12 * n23 * n44;
  double t14 = n14 * n23 * n32 - n13 * n24 * n32 - n14 * n22 * n33 + n12 * n24 * n33 +
               n13 * n22 * n34 - n12 * n23 * n34;
  double det = n11 * t11 + n21 * t12 + n31 * t13 + n41 * t14;
  if (det == 0.0) {
    for (int i = 0; i < 16; i++) o[i] = 0.0;
    return;
  }
  double d = 1.0 / det;
  o[0] = t11 * d;
  o[1] = (n24 * n33 * n41 - n23 * n34 * n41 - n24 * n31 * n43 + n21 * n34 * n43 +
          n23 * n31 * n44 - n21 * n33 * n44) * d;
  o[2] = (n22 * n34 * n41 - n24 * n32 * n41 + n24 * n31 * n42 - n21 * n34 * n42 -
          n22 * n31 * n44 + n21 * n32 * n44) * d;
  o[3] = (n23 * n32 * n41 - n22 * n33 * n41 - n23 * n31 * n42 + n21 * n33 * n42 +
          n22 * n31 * n43 - n21 * n32 * n43) * d;
  o[4] = t12 * d;
  o[5] = (n13 * n34 * n41 - n14 * n33 * n41 + n14 * n31 * n43 - n11 * n34 * n43 -
          n13 * n31 * n44 + n11 * n33 * n44) * d;
  o[6] = (n14 * n32 * n41 - n12 * n34 * n41 - n14 * n31 * n42 + n11 * n34 * n42 +
          n12 * n31 * n44 - n11 * n32 * n44) * d;
  o[7] = (n12 * n33 * n41 - n13 * n32 * n41 + n13 * n31 * n42 - n11 * n33 * n42 -
          n12 * n31 * n43 + n11 * n32 * n43) * d;
  o[8] = t13 * d;
  o[9] = (n14 * n23 * n41 - n13 * n24 * n41 - n14 * n21 * n43 + n11 * n24 * n43 +
          n13 * n21 * n44 - n11 * n23 * n44) * d;
  o[10] = (n12 * n24 * n41 - n14 * n22 * n41 + n14 * n21 * n42 - n11 * n24 * n42 -
           n12 * n21 * n44 + n11 * n22 * n44) * d;
  o[11] = (n13 * n22 * n41 - n12 * n23 * n41 - n13 * n21 * n42 + n11 * n23 * n42 +
           n12 * n21 * n43 - n11 * n22 * n43) * d;
  o[12] = t14 * d;
  o[13] = (n13 * n24 * n31 - n14 * n23 * n31 + n14 * n21 * n33 - n11 * n24 * n33 -
           n13 * n21 * n34 + n11 * n23 * n34) * d;
  o[14] = (n14 * n22 * n31 - n12 * n24 * n31 - n14 * n21 * n32 + n11 * n24 * n32 +
           n12 * n21 * n34 - n11 * n22 * n34) * d;
  o[15] = (n12 * n23 * n31 - n13 * n22 * n31 + n13 * n21 * n32 - n11 * n23 * n32 -
           n12 * n21 * n33 + n11 * n22 * n33) * d;
}

extern "C" void gs_object_frame(const double camera_world[16], const double object_world[16],
                                int visible, float out[40]) {
  double view[16], model_view[16], inverse_model[16];
  mat4_invert(camera_world, view);  // camera.matrixWorldInverse
  mat4_mul(view, object_world, model_view);
  mat4_invert(object_world, inverse_model);
  double cx = camera_world[12], cy = camera_world[13], cz = camera_world[14];
  // Vector3.applyMatrix4 (with the projective divide).
  const double* e = inverse_model;
  double w = 1.0 / (e[3] * cx + e[7] * cy + e[11] * cz + e[15]);
  double lx = (e[0] * cx + e[4] * cy + e[8] * cz + e[12]) * w;
  double ly = (e[1] * cx + e[5] * cy + e[9] * cz + e[13]) * w;
  double lz = (e[2] * cx + e[6] * cy + e[10] * cz + e[14]) * w;
  for (int i = 0; i < 16; i++) out[i] = (float)object_world[i];
  for (int i = 0; i < 16; i++) out[16 + i] = (float)model_view[i];
  out[32] = (float)lx;
  out[33] = (float)ly;
  out[34] = (float)lz;
  out[35] = 1.0f;
  out[36] = visible ? 1.0f : 0.0f;
  out[37] = out[38] = out[39] = 0.0f;
}

extern "C" uint32_t gs_capacity_next(uint32_t current, uint32_t demand, uint32_t maximum,
                                     double growth, double headroom) {
  if (demand > maximum) return 0;
  double desired = ceil((double)demand * headroom);
  double grown = ceil((double)current * growth);
  if (grown > desired) desired = grown;
  double rounded = ceil(desired / 256.0) * 256.0;
  if (rounded > (double)maximum) rounded = (double)maximum;
  return (uint32_t)rounded;
}

extern "C" int gs_resize_intersections(gs_ctx* ctx, uint32_t capacity) {
  if (!ctx) return GS_E_RANGE;
  if (capacity < 1 || capacity > ctx->cfg.max_intersection_capacity)
    return fail(ctx, GS_E_RANGE, "Intersection capacity exceeds device buffer or dispatch limits");
  (void)hipSetDevice(ctx->cfg.device);
  int rc = lane_records(ctx, ctx->lanes[0], capacity);
  if (rc) return rc;
  ctx->capacity = capacity;
  return GS_OK;
}

extern "C" int gs_resize_batch_intersections(gs_ctx* ctx, uint32_t view, uint32_t capacity) {
  if (!ctx) return GS_E_RANGE;
  if (view >= ctx->batch.n) return fail(ctx, GS_E_RANGE, "batch view out of range");
  if (capacity < 1 || capacity > ctx->cfg.max_intersection_capacity)
    return fail(ctx, GS_E_RANGE, "Intersection capacity exceeds device buffer or dispatch limits");
  ctx->batch.capacity[view] = capacity;
  return GS_OK;
}

namespace {

// TiledGaussianPipeline.ts:286-300 checks on one camera.
int check_frame(uint32_t ts, const gs_frame* f, const char*& msg) {
  msg = nullptr;
  if (f->width < 1 || f->height < 1) msg = "empty render size";
  else if (div_up(f->width, ts) > 65535 || div_up(f->height, ts) > 65535)
    msg = "Render size exceeds the tile dispatch limit";
  else if (!(f->near_plane > 0.0f && f->far_plane > f->near_plane && finite(f->far_plane)))
    msg = "depthRange must be positive and increasing";
  else if (!(f->filter_variance >= 0.0f && finite(f->filter_variance)))
    msg = "filterVariance must be finite and nonnegative";
  else if (f->boundary_axis > GS_BOUNDARY_REPEAT_Y)
    msg = "unknown boundaryPolicy";
  else
    for (int i = 0; i < 16; i++)
      if (!finite(f->projection[i])) msg = "projection must be finite";
  return msg ? GS_E_RANGE : GS_OK;
}

int validate_frame(gs_ctx* ctx, const gs_frame* f) {
  const char* msg = nullptr;
  int rc = check_frame(ctx->cfg.tile_size, f, msg);
  return rc ? fail(ctx, rc, msg) : GS_OK;
}

// rasterChunks.ts:73-78 (maxRasterChunkTasks): every chunked tile holds
// more than chunk_size records, so ceil(2K / chunk_size) tasks suffice.
uint32_t task_capacity_for(const gs_config& c, uint32_t capacity) {
  if (!c.raster_chunk_size) return 1;
  uint32_t t = div_up(2ull * capacity, c.raster_chunk_size);
  return t < 1 ? 1 : t;
}

// Raster units: the chunk tasks, tiles above 1024 records and (row-slab
// frames) the light tiles.
uint32_t unit_capacity_for(uint32_t task_capacity, uint32_t capacity, uint32_t tiles) {
  return task_capacity + div_up(capacity, 1024) + 1 + tiles;
}

int lane_frame_buffers(gs_ctx* ctx, Lane& l, uint32_t capacity, uint32_t tiles) {
  int rc;
  if ((rc = lane_slots(ctx, l, ctx->active_count)) || (rc = lane_records(ctx, l, capacity)) ||
      (rc = lane_sort(ctx, l)))
    return rc;
  if (tiles > l.tile_capacity || !l.tile_offsets.ptr) {
    uint32_t cap = next_pow2(tiles);
    if ((rc = ensure(ctx, l.tile_offsets, ((size_t)cap + 1) * 4, "tile-offsets")) ||
        (rc = ensure(ctx, l.chunk_first, ((size_t)cap + 1) * 4, "raster-chunk-offsets")))
      return rc;
    l.tile_capacity = cap;
  }
  uint32_t tasks = task_capacity_for(ctx->cfg, l.capacity);
  uint32_t units = unit_capacity_for(tasks, l.capacity, l.tile_capacity);
  size_t px = (size_t)ctx->cfg.tile_size * ctx->cfg.tile_size;
  if ((rc = ensure(ctx, l.tasks, (size_t)units * 8, "raster-heavy-units")) ||
      (rc = ensure(ctx, l.partials_px, (size_t)tasks * px * 16, "raster-chunk-partials")) ||
      (rc = ensure(ctx, l.partial_depth, (size_t)tasks * px * 4, "raster-chunk-depth")))
    return rc;
  return GS_OK;
}

// Camera-table bytes per view: the padded FrameCamera, then the objects.
size_t view_stride(uint32_t object_count) {
  size_t obj = (size_t)object_count * 40 * sizeof(float);
  return gs::kCameraBytes + (obj + 255) / 256 * 256;
}

// The frame uniforms of one camera (FrameUniforms.ts:41-58 plus the stage
// constants) under configuration c on a device with cu_count compute units.
void fill_params(const gs_config& c, uint32_t cu_count, const gs_frame* f,
                 uint32_t active_count, uint32_t object_count, uint32_t sh_degree,
                 bool has_active, uint32_t capacity, const FrameCamera* camera, FrameParams& p) {
  memset(&p, 0, sizeof(p));
  uint32_t ts = c.tile_size;
  uint32_t tiles_x = div_up(f->width, ts), tiles_y = div_up(f->height, ts);
  p.camera = camera;
  p.viewport[0] = (float)f->width;
  p.viewport[1] = (float)f->height;
  p.inv_viewport[0] = 1.0f / p.viewport[0];
  p.inv_viewport[1] = 1.0f / p.viewport[1];
  p.filter_variance = f->filter_variance;
  p.transmittance_threshold = c.transmittance_threshold;
  memcpy(p.background, c.background, sizeof(p.background));
  p.depth_alpha_threshold = c.depth_alpha_threshold;
  p.invalid_depth_value = c.invalid_depth_value;
  p.boundary_axis = f->boundary_axis;
  p.active_count = active_count;
  p.tiles_x = tiles_x;
  p.tiles_y = tiles_y;
  p.tile_count = tiles_x * tiles_y;
  p.width = f->width;
  p.height = f->height;
  p.capacity = capacity;
  p.object_count = object_count;
  p.sh_degree = sh_degree;
  p.depth_mode = c.depth_sort_mode;
  p.antialias_compensated = c.antialias_mode == GS_AA_COMPENSATED ? 1u : 0u;
  p.subpixel_culling = c.subpixel_sample_culling ? 1u : 0u;
  p.chunk_size = c.raster_chunk_size;
  p.has_active_slots = has_active ? 1u : 0u;
  p.output_depth = c.output_depth ? 1u : 0u;
  p.raster_stats = c.raster_stats ? 1u : 0u;
  p.tile_size = ts;
  // Heavy units run as row slabs above 1.25 units per CU: below that, a
  // whole-tile block per unit already covers the CUs and slabs mostly
  // repeat the per-record staging. Measured on this GPU: the 1M/1080p frame
  // (275 units) rasterizes its heavy units in 0.21 ms whole, 0.30 ms as
  // 8-row slabs; batched culled 3M/1080p views run 1,483 frames/s with this
  // threshold, 1,458 at 1.75 units per CU and 1,430 at 2 (the 2M
  // mixed-resolution batch and plain 3M views of 906-1,907 units do not
  // move).
  p.split_units = c.heavy_split_units == GS_SPLIT_AUTO ? cu_count * 5 / 4 : c.heavy_split_units;
  // Row slabs have 8 rows up to 16 units per CU, 4 above: the 3M/1080p
  // views (906-1,907 units) rasterize 9-13 % faster with 8, the 6M/4K
  // frame (17,773 units) 25 % faster with 4.
  p.fine_split_units =
      c.heavy_fine_split_units == GS_SPLIT_AUTO ? cu_count * 16 : c.heavy_fine_split_units;
  // Between the two, up to 5.5 units per CU, whole tiles with two lanes per
  // pixel (three 512-thread blocks to a CU) instead of 8-row slabs, which
  // stage every record twice: plain 3M/1080p views of 857-1,337 units
  // rasterize 1-9 % faster, 1,546 units 1.5 % slower, and 2560x1440 views of
  // 1,148-1,354 units about 1 % slower (the 64-view 2M batch: 1,667 ->
  // 1,733 frames/s with the launch order of launch_raster).
  p.pair_units = c.heavy_pair_units == GS_SPLIT_AUTO ? cu_count * 11 / 2 : c.heavy_pair_units;
}

// Launch description of view v of `set` on lane l. All lane buffers are
// already sized.
void fill_frame(gs_ctx* ctx, ViewSet& set, uint32_t v, const gs_frame* f, const Lane& l,
                void* color_dst, void* depth_dst, FrameBuffers& b) {
  memset(&b, 0, sizeof(b));
  uint32_t ts = ctx->cfg.tile_size;
  uint32_t tiles = div_up(f->width, ts) * div_up(f->height, ts);
  uint32_t capacity = set.capacity[v];
  const unsigned char* dyn = (const unsigned char*)set.dynamic.ptr + (size_t)v * set.stride;
  FrameParams& p = b.params;
  fill_params(ctx->cfg, ctx->cu_count, f, ctx->active_count, ctx->object_count, ctx->sh_degree,
              ctx->has_active, capacity, (const FrameCamera*)dyn, p);
  // Training-set frames: chunking off, so every recorded alpha * T uses the
  // true transmittance in front of it.
  if (ctx->cull.collecting) p.chunk_size = 0;
  b.state = (DeviceState*)set.state.ptr + v;
  b.scene.means = (const float4*)ctx->means.ptr;
  b.scene.scales_opacity = (const uint2*)ctx->scales_opacity.ptr;
  b.scene.rotations = (const uint32_t*)ctx->rotations.ptr;
  b.scene.sh = (const uint32_t*)ctx->sh.ptr;
  b.scene.sh_top = (const uint8_t*)ctx->sh_top.ptr;
  b.scene.active_slots = (const uint32_t*)ctx->active.ptr;
  b.scene.objects = (const float*)(dyn + gs::kCameraBytes);
  b.proj.mean = (float4*)l.pmean.ptr;
  b.proj.splat = (gs::Splat*)l.splat.ptr;
  b.proj.tile_counts = (uint32_t*)l.tile_counts.ptr;
  b.proj.hits = (uint2*)l.hits.ptr;
  b.ordered_hits = (uint2*)l.ordered_hits.ptr;
  b.visible_offsets = (uint32_t*)l.visible_offsets.ptr;
  b.depth_a = (uint2*)l.depth_a.ptr;
  b.depth_b = (uint2*)l.depth_b.ptr;
  b.ordered_counts = (uint32_t*)l.ordered_counts.ptr;
  b.intersection_offsets = (uint32_t*)l.intersection_offsets.ptr;
  b.records_a = (uint2*)l.records_a.ptr;
  b.records_b = (uint2*)l.records_b.ptr;
  b.tile_offsets = (uint32_t*)l.tile_offsets.ptr;
  b.chunk_first = (uint32_t*)l.chunk_first.ptr;
  b.tasks = (uint2*)l.tasks.ptr;
  b.chunk_partials = (float4*)l.partials_px.ptr;
  b.partial_depth = (float*)l.partial_depth.ptr;
  b.task_capacity = task_capacity_for(ctx->cfg, capacity);
  b.unit_capacity = unit_capacity_for(b.task_capacity, capacity, l.tile_capacity);
  b.hist = (uint32_t*)l.hist.ptr;
  b.sort_parts = l.sort_parts;
  b.partials = (uint32_t*)l.partials.ptr;
  b.color_out = (uint16_t*)(color_dst ? color_dst : set.color[v].ptr);
  b.depth_out = ctx->cfg.output_depth ? (float*)(depth_dst ? depth_dst : set.depth[v].ptr)
                                      : nullptr;
  b.max_n = ctx->active_count;
  b.max_k = capacity;
  b.max_tiles = tiles;
  b.radix_bits = ctx->cfg.radix_bits;
  b.raster_batch = ctx->cfg.raster_batch;
  b.cu_count = ctx->cu_count;
  uint32_t key_bits = ctx->cfg.depth_sort_mode == GS_DEPTH_FLOAT32 ? 32u : 16u;
  b.depth_passes = div_up(key_bits, b.radix_bits);
  b.tile_passes = div_up(bits_for(tiles), b.radix_bits);
  if (b.tile_passes == 0) b.tile_passes = 1;
  if (ctx->cull.collecting) {
    b.cull_feats = (float*)ctx->cull.feats.ptr;
    b.record = (uint32_t*)ctx->cull.record.ptr;
  } else if (ctx->cull.enabled) {
    b.cull_net = (const gs::CullNet*)ctx->cull.net.ptr;
  }
}

// Stage the per-view camera uniforms and object frames through the next
// pinned ring slot (once the copy that last used it has completed) and
// enqueue their upload on the context stream.
int upload_cameras(gs_ctx* ctx, ViewSet& set, uint32_t n, const gs_frame* frames) {
  size_t bytes = (size_t)n * set.stride;
  ctx->staging.assign(bytes, 0);
  size_t obj_bytes = (size_t)ctx->object_count * 40 * sizeof(float);
  for (uint32_t v = 0; v < n; v++) {
    unsigned char* dst = ctx->staging.data() + (size_t)v * set.stride;
    FrameCamera cam;
    memset(&cam, 0, sizeof(cam));
    memcpy(cam.projection, frames[v].projection, sizeof(cam.projection));
    cam.near_plane = frames[v].near_plane;
    cam.far_plane = frames[v].far_plane;
    cam.attempt = frames[v].attempt;
    cam.reset_valid = frames[v].reset_valid;
    memcpy(dst, &cam, sizeof(cam));
    const float* objects = frames[v].object_frames ? frames[v].object_frames
                                                   : ctx->objects_host.data();
    if (obj_bytes) memcpy(dst + gs::kCameraBytes, objects, obj_bytes);
  }
  if (bytes > ctx->ring_bytes) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    for (int i = 0; i < gs_ctx::kRing; i++) {
      if (ctx->ring[i]) (void)hipHostFree(ctx->ring[i]);
      ctx->ring[i] = nullptr;
      if (hipHostMalloc(&ctx->ring[i], bytes, hipHostMallocDefault) != hipSuccess)
        return fail(ctx, GS_E_OOM, "allocation-failed: camera staging");
      if (!ctx->ring_done[i] &&
          hipEventCreateWithFlags(&ctx->ring_done[i], hipEventDisableTiming) != hipSuccess)
        return fail(ctx, GS_E_DEVICE, "hipEventCreate");
    }
    ctx->ring_bytes = bytes;
  }
  int slot = ctx->ring_next;
  ctx->ring_next = (slot + 1) % gs_ctx::kRing;
  hipError_t e = hipEventSynchronize(ctx->ring_done[slot]);
  if (e != hipSuccess) return hip_fail(ctx, e, "camera staging wait");
  memcpy(ctx->ring[slot], ctx->staging.data(), bytes);
  e = hipMemcpyAsync(set.dynamic.ptr, ctx->ring[slot], bytes, hipMemcpyHostToDevice,
                     ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ctx->ring_done[slot], ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "camera upload");
  return GS_OK;
}

// Enqueue every frame of the set: lane l > 0 forks from the context stream
// and joins it at the end, frames of one lane run in order.
hipError_t enqueue_set(gs_ctx* ctx, ViewSet& set, hipEvent_t* marks) {
  hipStream_t s = ctx->stream;
  if (set.group > 1) {
    // View groups: each stage one launch over the group's views, on the
    // context stream (the raster's second branch on lane 0's aux stream).
    const Lane& l0 = ctx->lanes[0];
    FrameBuffers grp[gs::kViewGroupMax];
    for (uint32_t first = 0; first < set.n; first += set.group) {
      uint32_t m = std::min(set.group, set.n - first);
      for (uint32_t k = 0; k < m; k++) grp[k] = set.frames[set.order[first + k]];
      hipError_t e = gs::launch_group(grp, m, s, gs::RasterStreams{l0.aux, l0.fork, l0.join},
                                      marks);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  uint32_t used = 0;
  for (uint32_t v = 0; v < set.n; v++) used = std::max(used, set.lane_of[v] + 1);
  hipError_t e = hipSuccess;
  if (used > 1) {
    e = hipEventRecord(ctx->batch_fork, s);
    for (uint32_t l = 1; l < used && e == hipSuccess; l++)
      e = hipStreamWaitEvent(ctx->lanes[l].stream, ctx->batch_fork, 0);
    if (e != hipSuccess) return e;
  }
  for (uint32_t v = 0; v < set.n; v++) {
    Lane& l = ctx->lanes[set.lane_of[v]];
    e = gs::launch_frame(set.frames[v], l.stream,
                         gs::RasterStreams{l.aux, l.fork, l.join}, marks);
    if (e != hipSuccess) return e;
  }
  for (uint32_t l = 1; l < used; l++) {
    e = hipEventRecord(ctx->lanes[l].done, ctx->lanes[l].stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, ctx->lanes[l].done, 0);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

bool same_key(const ViewSet& set) {
  if (!set.graph_valid || set.key.size() != set.n || set.key_lanes != set.lane_of ||
      set.key_order != set.order || set.key_slots != set.slot_of)
    return false;
  return memcmp(set.key.data(), set.frames.data(), sizeof(FrameBuffers) * set.n) == 0;
}

int launch_set(gs_ctx* ctx, ViewSet& set, bool profile) {
  hipStream_t s = ctx->stream;
  if (profile) {
    if (ctx->profile_frames >= ctx->profile_capacity)
      return fail(ctx, GS_E_RANGE, "profiling capacity exhausted");
    hipEvent_t* marks = &ctx->marks[(size_t)ctx->profile_frames * gs::kMarks];
    hipError_t e = enqueue_set(ctx, set, marks);
    if (e != hipSuccess) return hip_fail(ctx, e, "frame launch");
    ctx->profile_frames++;
    return GS_OK;
  }
  if (!ctx->cfg.capture_graph) {
    hipError_t e = enqueue_set(ctx, set, nullptr);
    if (e != hipSuccess) return hip_fail(ctx, e, "frame launch");
    return GS_OK;
  }
  if (!same_key(set)) {
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamBeginCapture");
    hipError_t le = enqueue_set(ctx, set, nullptr);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (le != hipSuccess) e = le;
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      set.graph_valid = false;
      return hip_fail(ctx, e, "frame capture");
    }
    // An update in place only for graphs of one lane (its frames and their
    // raster branches, a topology the re-capture tests cover); any other
    // graph, or a failed update, is instantiated afresh.
    bool updated = false;
    bool one_lane = std::all_of(set.lane_of.begin(), set.lane_of.end(),
                                [](uint32_t l) { return l == 0; });
    if (set.exec && !one_lane) {
      (void)hipGraphExecDestroy(set.exec);
      set.exec = nullptr;
    }
    if (set.exec) {
      hipGraphExecUpdateResult result;
      hipGraphNode_t error_node = nullptr;
      updated = hipGraphExecUpdate(set.exec, g, &error_node, &result) == hipSuccess &&
                result == hipGraphExecUpdateSuccess;
      if (!updated) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(set.exec);
        set.exec = nullptr;
      }
    }
    if (!updated) {
      e = hipGraphInstantiate(&set.exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        set.graph_valid = false;
        return hip_fail(ctx, e, "hipGraphInstantiate");
      }
    }
    if (set.graph) (void)hipGraphDestroy(set.graph);
    set.graph = g;
    set.key = set.frames;
    set.key_lanes = set.lane_of;
    set.key_order = set.order;
    set.key_slots = set.slot_of;
    set.graph_valid = true;
  }
  hipError_t e = hipGraphLaunch(set.exec, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphLaunch");
  return GS_OK;
}

// Spread views over lanes, largest first onto the least loaded lane.
void assign_lanes(gs_ctx* ctx, ViewSet& set, const gs_frame* frames, bool single) {
  set.lane_of.assign(set.n, 0);
  if (single) return;
  uint32_t lanes = std::min<uint32_t>((uint32_t)ctx->lanes.size(), set.n);
  std::vector<uint32_t> order(set.n);
  for (uint32_t v = 0; v < set.n; v++) order[v] = v;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return (uint64_t)frames[a].width * frames[a].height >
           (uint64_t)frames[b].width * frames[b].height;
  });
  std::vector<uint64_t> load(lanes, 0);
  for (uint32_t v : order) {
    uint32_t best = 0;
    for (uint32_t l = 1; l < lanes; l++)
      if (load[l] < load[best]) best = l;
    set.lane_of[v] = best;
    load[best] += (uint64_t)frames[v].width * frames[v].height;
  }
}

// View groups of a batch: the views by area, largest first (stable), in
// groups of view_group; view k of a group uses working set k.
void assign_groups(gs_ctx* ctx, ViewSet& set, const gs_frame* frames, bool single) {
  set.group = single ? 1u : std::min<uint32_t>(ctx->cfg.view_group, set.n);
  if (set.group == 0) set.group = 1;
  set.order.resize(set.n);
  for (uint32_t v = 0; v < set.n; v++) set.order[v] = v;
  set.slot_of.assign(set.n, 0);
  if (set.group == 1) return;
  std::stable_sort(set.order.begin(), set.order.end(), [&](uint32_t a, uint32_t b) {
    return (uint64_t)frames[a].width * frames[a].height >
           (uint64_t)frames[b].width * frames[b].height;
  });
  for (uint32_t i = 0; i < set.n; i++) set.slot_of[set.order[i]] = i % set.group;
}

// Working set of view v: its lane's, or a group slot's.
Lane& working_set(gs_ctx* ctx, const ViewSet& set, uint32_t v) {
  uint32_t j = set.slot_of[v];
  return j == 0 ? ctx->lanes[set.lane_of[v]] : ctx->slots[j - 1];
}

// prepareFrame + render for the views of a set (TiledGaussianPipeline.ts
// :286-341), one launch.
int render_set(gs_ctx* ctx, ViewSet& set, uint32_t n, const gs_frame* frames,
               const gs_image* outs, bool single) {
  if (!ctx->has_data) return fail(ctx, GS_E_STATE, "render before data");
  int rc;
  for (uint32_t v = 0; v < n; v++)
    if ((rc = validate_frame(ctx, &frames[v]))) return rc;
  (void)hipSetDevice(ctx->cfg.device);
  // Views, their capacities and states.
  if (set.capacity.size() < n) set.capacity.resize(n, ctx->cfg.intersection_capacity);
  if (single) set.capacity[0] = ctx->capacity;
  set.n = n;
  size_t state_bytes = (size_t)std::max<uint32_t>(n, 1) * sizeof(DeviceState);
  if (set.state.bytes < state_bytes) {
    if ((rc = ensure(ctx, set.state, state_bytes, "state"))) return rc;
    hipError_t e = hipMemsetAsync(set.state.ptr, 0, set.state.bytes, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "state clear");
  }
  set.stride = view_stride(ctx->object_count);
  if ((rc = ensure(ctx, set.dynamic, (size_t)n * set.stride, "camera-table"))) return rc;
  if (set.color.size() < n) {
    set.color.resize(n);
    set.depth.resize(n);
  }
  set.width.resize(n);
  set.height.resize(n);
  assign_lanes(ctx, set, frames, single);
  assign_groups(ctx, set, frames, single);
  if (ctx->slots.size() + 1 < set.group) ctx->slots.resize(set.group - 1);
  // Working sets (lanes, then group slots): the largest capacity and tile
  // grid of their views.
  uint32_t ts = ctx->cfg.tile_size;
  const size_t n_sets = ctx->lanes.size() + ctx->slots.size();
  auto set_index = [&](uint32_t v) {
    return set.slot_of[v] == 0 ? (size_t)set.lane_of[v]
                               : ctx->lanes.size() + set.slot_of[v] - 1;
  };
  std::vector<uint32_t> lane_cap(n_sets, 0), lane_tiles(n_sets, 0);
  for (uint32_t v = 0; v < n; v++) {
    size_t l = set_index(v);
    uint32_t tiles = div_up(frames[v].width, ts) * div_up(frames[v].height, ts);
    lane_cap[l] = std::max(lane_cap[l], set.capacity[v]);
    lane_tiles[l] = std::max(lane_tiles[l], tiles);
  }
  for (size_t l = 0; l < n_sets; l++) {
    Lane& w = l < ctx->lanes.size() ? ctx->lanes[l] : ctx->slots[l - ctx->lanes.size()];
    if (lane_cap[l] && (rc = lane_frame_buffers(ctx, w, lane_cap[l], lane_tiles[l]))) return rc;
  }
  // Outputs.
  std::vector<void*> color_dst(n, nullptr), depth_dst(n, nullptr);
  for (uint32_t v = 0; v < n; v++) {
    const gs_image* o = outs ? &outs[v] : nullptr;
    if (o && o->memory == GS_MEMORY_DEVICE) {
      color_dst[v] = o->color;
      depth_dst[v] = o->depth;
    }
    size_t pixels = (size_t)frames[v].width * frames[v].height;
    if (!color_dst[v] && (rc = ensure(ctx, set.color[v], pixels * 8, "color-target")))
      return rc;
    if (ctx->cfg.output_depth && !depth_dst[v] &&
        (rc = ensure(ctx, set.depth[v], pixels * 4, "depth-target")))
      return rc;
  }
  // Every view's lane exists and its working set covers the view.
  for (uint32_t v = 0; v < n; v++) {
    uint32_t l = set.lane_of[v];
    if (l >= ctx->lanes.size()) return fail(ctx, GS_E_STATE, "view assigned to no lane");
    if (set.slot_of[v] > ctx->slots.size())
      return fail(ctx, GS_E_STATE, "view assigned to no group slot");
    const Lane& lane = working_set(ctx, set, v);
    uint32_t tiles = div_up(frames[v].width, ts) * div_up(frames[v].height, ts);
    if (lane.capacity < set.capacity[v] || lane.tile_capacity < tiles ||
        lane.slot_capacity < ctx->active_count)
      return fail(ctx, GS_E_STATE, "lane working set smaller than its view");
  }
  set.frames.resize(n);
  for (uint32_t v = 0; v < n; v++) {
    fill_frame(ctx, set, v, &frames[v], working_set(ctx, set, v), color_dst[v], depth_dst[v],
               set.frames[v]);
    set.width[v] = frames[v].width;
    set.height[v] = frames[v].height;
  }
  // The views of a group sort their tile keys with the same passes: a pass
  // above a view's own tile bits finds one digit for every key and copies.
  for (uint32_t first = 0; set.group > 1 && first < n; first += set.group) {
    uint32_t m = std::min(set.group, n - first), passes = 0;
    for (uint32_t k = 0; k < m; k++)
      passes = std::max(passes, set.frames[set.order[first + k]].tile_passes);
    for (uint32_t k = 0; k < m; k++) set.frames[set.order[first + k]].tile_passes = passes;
  }
  if ((rc = upload_cameras(ctx, set, n, frames))) return rc;
  if ((rc = launch_set(ctx, set, single && ctx->profile_capacity != 0))) return rc;
  set.rendered = true;
  return GS_OK;
}

int read_image(gs_ctx* ctx, const ViewSet& set, uint32_t v, void* color, size_t color_bytes,
               void* depth, size_t depth_bytes) {
  if (!set.rendered || v >= set.n) return fail(ctx, GS_E_STATE, "nothing rendered");
  (void)hipSetDevice(ctx->cfg.device);
  const FrameBuffers& b = set.frames[v];
  size_t pixels = (size_t)set.width[v] * set.height[v];
  hipError_t e = hipSuccess;
  if (color) {
    if (color_bytes != pixels * 8) return fail(ctx, GS_E_RANGE, "color buffer size mismatch");
    e = hipMemcpyAsync(color, b.color_out, pixels * 8, hipMemcpyDeviceToHost, ctx->stream);
  }
  if (e == hipSuccess && depth) {
    if (!b.depth_out) return fail(ctx, GS_E_STATE, "outputDepth is disabled");
    if (depth_bytes != pixels * 4) return fail(ctx, GS_E_RANGE, "depth buffer size mismatch");
    e = hipMemcpyAsync(depth, b.depth_out, pixels * 4, hipMemcpyDeviceToHost, ctx->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "image readback");
  return GS_OK;
}

int read_states(gs_ctx* ctx, const ViewSet& set, uint32_t first, uint32_t n, gs_state* out) {
  if (!set.rendered || (uint64_t)first + n > set.n)
    return fail(ctx, GS_E_STATE, "nothing rendered");
  (void)hipSetDevice(ctx->cfg.device);
  std::vector<DeviceState> s(n);
  hipError_t e = hipSuccess;
  if (n)
    e = hipMemcpyAsync(s.data(), (const DeviceState*)set.state.ptr + first,
                       n * sizeof(DeviceState), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "state readback");
  for (uint32_t i = 0; i < n; i++) {
    const FrameParams& p = set.frames[first + i].params;
    gs_state& o = out[i];
    memset(&o, 0, sizeof(o));
    o.attempt = s[i].out_attempt;
    o.version = s[i].out_version;
    o.requested = s[i].out_demand;
    o.overflow = s[i].out_overflow;
    o.visible = s[i].visible;
    o.intersections = s[i].count;
    o.capacity = set.capacity[first + i];
    o.tiles_x = p.tiles_x;
    o.tiles_y = p.tiles_y;
    o.raster_checked = s[i].checked;
    o.raster_blended = s[i].blended;
    o.raster_checked_heavy = s[i].checked_heavy;
    o.error_flags = s[i].sort_error;
  }
  return GS_OK;
}

}  // namespace

extern "C" int gs_render(gs_ctx* ctx, const gs_frame* f, const gs_image* out) {
  if (!ctx || !f) return GS_E_RANGE;
  int rc = render_set(ctx, ctx->single, 1, f, out, true);
  if (rc) return rc;
  if (out && out->memory == GS_MEMORY_HOST)
    return gs_read_image(ctx, out->color, out->color ? (size_t)f->width * f->height * 8 : 0,
                         out->depth, out->depth ? (size_t)f->width * f->height * 4 : 0);
  return GS_OK;
}

extern "C" int gs_render_batch(gs_ctx* ctx, uint32_t n, const gs_frame* frames,
                               const gs_image* outs) {
  if (!ctx || (n > 0 && !frames)) return GS_E_RANGE;
  if (n == 0) return GS_OK;
  int rc = render_set(ctx, ctx->batch, n, frames, outs, false);
  if (rc) return rc;
  for (uint32_t v = 0; v < n; v++)
    if (outs && outs[v].memory == GS_MEMORY_HOST) {
      size_t px = (size_t)frames[v].width * frames[v].height;
      if ((rc = read_image(ctx, ctx->batch, v, outs[v].color, outs[v].color ? px * 8 : 0,
                           outs[v].depth, outs[v].depth ? px * 4 : 0)))
        return rc;
    }
  return GS_OK;
}

extern "C" int gs_profile_begin(gs_ctx* ctx, uint32_t max_frames) {
  if (!ctx || max_frames == 0) return GS_E_RANGE;
  (void)hipSetDevice(ctx->cfg.device);
  size_t need = (size_t)max_frames * gs::kMarks;
  while (ctx->marks.size() < need) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return fail(ctx, GS_E_DEVICE, "hipEventCreate");
    ctx->marks.push_back(e);
  }
  ctx->profile_capacity = max_frames;
  ctx->profile_frames = 0;
  return GS_OK;
}

extern "C" int gs_profile_end(gs_ctx* ctx, double stage_ms[GS_STAGE_COUNT], double* frame_ms,
                              double* heavy_ms, uint32_t* frames) {
  if (!ctx) return GS_E_RANGE;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
  double sums[gs::kStageCount] = {0};
  double total = 0.0, heavy_total = 0.0;
  for (uint32_t f = 0; f < ctx->profile_frames; f++) {
    hipEvent_t* m = &ctx->marks[(size_t)f * gs::kMarks];
    for (int st = 0; st < gs::kStageCount; st++) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, m[st], m[st + 1]) != hipSuccess)
        return fail(ctx, GS_E_DEVICE, "hipEventElapsedTime");
      sums[st] += ms;
    }
    // The heavy-unit kernel: of the four shapes' launches, all but the
    // frame's exit at once; its time is the longest of their intervals
    // (each taken between the events around it on its own stream).
    float whole = 0.0f, heavy = 0.0f;
    if (hipEventElapsedTime(&whole, m[0], m[gs::kStageCount]) != hipSuccess)
      return fail(ctx, GS_E_DEVICE, "hipEventElapsedTime");
    for (int k = 0; k < gs::kMarkHeavyShapes; k++) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, m[gs::kMarkHeavyBegin + 2 * k],
                              m[gs::kMarkHeavyBegin + 2 * k + 1]) != hipSuccess)
        return fail(ctx, GS_E_DEVICE, "hipEventElapsedTime");
      heavy = ms > heavy ? ms : heavy;
    }
    total += whole;
    heavy_total += heavy;
  }
  if (stage_ms)
    for (int st = 0; st < gs::kStageCount; st++) stage_ms[st] = sums[st];
  if (frame_ms) *frame_ms = total;
  if (heavy_ms) *heavy_ms = heavy_total;
  if (frames) *frames = ctx->profile_frames;
  ctx->profile_capacity = 0;
  ctx->profile_frames = 0;
  return GS_OK;
}

extern "C" int gs_read_state(gs_ctx* ctx, gs_state* out) {
  if (!ctx || !out) return GS_E_RANGE;
  if (!ctx->single.rendered) {
    // Before the first frame: an empty state with the configured capacity.
    memset(out, 0, sizeof(*out));
    out->capacity = ctx->capacity;
    return GS_OK;
  }
  int rc = read_states(ctx, ctx->single, 0, 1, out);
  if (rc == GS_OK) out->capacity = ctx->capacity;
  return rc;
}

extern "C" int gs_read_batch_state(gs_ctx* ctx, uint32_t first, uint32_t n, gs_state* out) {
  if (!ctx || (n > 0 && !out)) return GS_E_RANGE;
  return read_states(ctx, ctx->batch, first, n, out);
}

extern "C" int gs_read_image(gs_ctx* ctx, void* color, size_t color_bytes, void* depth,
                             size_t depth_bytes) {
  if (!ctx) return GS_E_RANGE;
  return read_image(ctx, ctx->single, 0, color, color_bytes, depth, depth_bytes);
}

extern "C" int gs_read_batch_image(gs_ctx* ctx, uint32_t view, void* color, size_t color_bytes,
                                   void* depth, size_t depth_bytes) {
  if (!ctx) return GS_E_RANGE;
  return read_image(ctx, ctx->batch, view, color, color_bytes, depth, depth_bytes);
}

extern "C" int gs_batch_device_image(gs_ctx* ctx, uint32_t view, void** color, void** depth) {
  if (!ctx) return GS_E_RANGE;
  if (view >= ctx->batch.n) return fail(ctx, GS_E_RANGE, "batch view out of range");
  if (color) *color = ctx->batch.frames[view].color_out;
  if (depth) *depth = ctx->batch.frames[view].depth_out;
  return GS_OK;
}

// The debug copies describe the last gs_render; a later render of either
// kind reuses lane 0's buffers.
extern "C" int gs_read_debug(gs_ctx* ctx, const char* name, void* dst, size_t bytes,
                             size_t* size) {
  if (!ctx || !name) return GS_E_RANGE;
  if (!ctx->single.rendered) return fail(ctx, GS_E_STATE, "nothing rendered");
  gs_state st;
  int rc = gs_read_state(ctx, &st);
  if (rc) return rc;
  const FrameBuffers& b = ctx->single.frames[0];
  std::string n(name);
  const void* src = nullptr;
  size_t len = 0;
  size_t a = b.params.active_count;
  size_t pitch = 0;  // nonzero: 16-byte fields of the splat records, this far apart
  if (n == "projected_mean") { src = b.proj.mean; len = a * 16; }
  else if (n == "projected_conic") { src = &b.proj.splat[0].conic; len = a * 16; pitch = sizeof(gs::Splat); }
  else if (n == "projected_color") { src = &b.proj.splat[0].color; len = a * 16; pitch = sizeof(gs::Splat); }
  else if (n == "tile_counts") { src = b.proj.tile_counts; len = a * 4; }
  else if (n == "visible_offsets") { src = b.visible_offsets; len = a * 4; }
  else if (n == "depth_records") {
    // Which buffer holds them depends on whether the last depth pass ran.
    uint32_t in_place = 0u;
    hipError_t e = hipMemcpy(&in_place, &b.state->depth_in_place, sizeof(in_place),
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "state readback");
    src = gs::sorted_depth_records(b, in_place);
    len = (size_t)st.visible * 8;
  }
  else if (n == "ordered_counts") { src = b.ordered_counts; len = (size_t)st.visible * 4; }
  else if (n == "intersection_offsets") { src = b.intersection_offsets; len = (size_t)st.visible * 4; }
  else if (n == "sorted_records") { src = gs::sorted_tile_records(b); len = (size_t)st.intersections * 8; }
  else if (n == "tile_offsets") { src = b.tile_offsets; len = ((size_t)b.params.tile_count + 1) * 4; }
  else if (n == "cull_features" && b.cull_feats) { src = b.cull_feats; len = a * 64; }
  else if (n == "contribution_record" && b.record) { src = b.record; len = a * 4; }
  else return fail(ctx, GS_E_RANGE, "unknown debug buffer: " + n);
  if (size) *size = len;
  if (!dst) return GS_OK;
  if (bytes != len) return fail(ctx, GS_E_RANGE, "debug buffer size mismatch for " + n);
  if (len == 0) return GS_OK;
  hipError_t e = pitch ? hipMemcpy2DAsync(dst, 16, src, pitch, 16, len / 16,
                                          hipMemcpyDeviceToHost, ctx->stream)
                       : hipMemcpyAsync(dst, src, len, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "debug readback");
  return GS_OK;
}

extern "C" void* gs_stream(gs_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int gs_device_image(gs_ctx* ctx, void** color, void** depth) {
  if (!ctx) return GS_E_RANGE;
  const ViewSet& s = ctx->single;
  if (color) *color = s.color.empty() ? nullptr : s.color[0].ptr;
  if (depth) *depth = s.depth.empty() ? nullptr : s.depth[0].ptr;
  return GS_OK;
}

extern "C" int gs_set_tuning(gs_ctx* ctx, const gs_tuning* t) {
  if (!ctx || !t) return GS_E_RANGE;
  if (!valid_tuning(t->tile_size, t->radix_bits, t->raster_batch))
    return fail(ctx, GS_E_RANGE,
                "tile size must be 8, 16 or 32, radix bits 4, 6 or 8, raster batch 128 or 256");
  ctx->cfg.tile_size = t->tile_size;
  ctx->cfg.radix_bits = t->radix_bits;
  ctx->cfg.raster_batch = t->raster_batch;
  return GS_OK;
}

extern "C" int gs_get_tuning(gs_ctx* ctx, gs_tuning* t) {
  if (!ctx || !t) return GS_E_RANGE;
  t->tile_size = ctx->cfg.tile_size;
  t->radix_bits = ctx->cfg.radix_bits;
  t->raster_batch = ctx->cfg.raster_batch;
  return GS_OK;
}

// Render the batch until no view overflows, growing each overflowing
// view's capacity by the reference's policy (IntersectionCapacityPolicy.ts
// :90-101 with its default growth 1.5 and headroom 1.25).
static int render_batch_fit(gs_ctx* ctx, uint32_t n, const gs_frame* frames) {
  std::vector<gs_state> st(n);
  for (int attempt = 0; attempt < 8; attempt++) {
    int rc = gs_render_batch(ctx, n, frames, nullptr);
    if (rc == GS_OK) rc = gs_read_batch_state(ctx, 0, n, st.data());
    if (rc) return rc;
    bool grown = false;
    for (uint32_t v = 0; v < n; v++)
      if (st[v].overflow) {
        uint32_t next = gs_capacity_next(ctx->batch.capacity[v], st[v].requested,
                                         ctx->cfg.max_intersection_capacity, 1.5, 1.25);
        if (next == 0) return fail(ctx, GS_E_RANGE, "intersection demand exceeds maximum");
        ctx->batch.capacity[v] = next;
        grown = true;
      }
    if (!grown) return GS_OK;
  }
  return fail(ctx, GS_E_STATE, "batch capacity did not settle");
}

extern "C" int gs_autotune(gs_ctx* ctx, uint32_t n, const gs_frame* frames, uint32_t reps,
                           gs_tuning* best, gs_tune_entry* table, uint32_t table_len,
                           uint32_t* entries) {
  if (!ctx || !frames || n == 0 || reps == 0) return GS_E_RANGE;
  (void)hipSetDevice(ctx->cfg.device);
  static const uint32_t tiles[] = {8, 16, 32}, digits[] = {4, 6, 8}, batches[] = {128, 256};
  gs_tuning saved;
  gs_get_tuning(ctx, &saved);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)
    return fail(ctx, GS_E_DEVICE, "hipEventCreate");
  gs_tuning winner = saved;
  double winner_ms = 1e300;
  uint32_t count = 0;
  int rc = GS_OK;
  for (uint32_t tile : tiles)
    for (uint32_t bits : digits)
      for (uint32_t batch : batches) {
        gs_tuning t = {tile, bits, batch};
        if ((rc = gs_set_tuning(ctx, &t)) || (rc = render_batch_fit(ctx, n, frames))) break;
        hipError_t e = hipEventRecord(t0, ctx->stream);
        for (uint32_t r = 0; r < reps && e == hipSuccess && rc == GS_OK; r++)
          rc = gs_render_batch(ctx, n, frames, nullptr);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(t1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (e != hipSuccess) {
          rc = hip_fail(ctx, e, "autotune timing");
          break;
        }
        double per = (double)ms / reps;
        if (table && count < table_len) table[count] = {t, per};
        count++;
        if (per < winner_ms) {
          winner_ms = per;
          winner = t;
        }
      }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc) {
    gs_set_tuning(ctx, &saved);
    return rc;
  }
  gs_set_tuning(ctx, &winner);
  if (best) *best = winner;
  if (entries) *entries = count;
  return GS_OK;
}


// ------------------------------------------------------ no-allocation path ----
//
// gs_render_noalloc renders one frame like gs_render, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops): the scene and
// the output images are the caller's device buffers, and every working array
// of the frame is carved out of one file-scope device array, `scratch`. Each
// GPU the library runs on holds its own copy of it from module load until the
// process ends. Calls on one GPU share that copy: calls that overlap on one
// GPU (on different streams, or from different threads) race.

namespace {

static_assert(GS_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size is a multiple of 256");
static __device__ __align__(256) unsigned char scratch[GS_NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` per device: the host-side name is only a
// stand-in, so kernels get the address hipGetSymbolAddress returns for the
// current device, taken once per device on its first call.
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

int scratch_address(int device, unsigned char** out) {
  if (device < 0 || device >= kNoAllocMaxDevices) return GS_E_RANGE;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return GS_E_DEVICE;
    }
    g_scratch_ptr[device] = ptr;
  }
  *out = (unsigned char*)g_scratch_ptr[device];
  return GS_OK;
}

// Offsets of one frame's working arrays inside the scratch (256-byte
// aligned), as the context path sizes them for one lane.
struct NoAllocLayout {
  size_t state, camera, pmean, splat, tile_counts, hits, ordered_hits, visible_offsets, depth_a,
      depth_b, ordered_counts, intersection_offsets, records_a, records_b, hist, partials,
      tile_offsets, chunk_first, tasks, partials_px, partial_depth;
  size_t total;
  uint32_t sort_parts, scan_parts, task_capacity, unit_capacity;
};

NoAllocLayout noalloc_layout(const gs_config& c, uint32_t n, uint32_t capacity, uint32_t tiles) {
  NoAllocLayout L;
  size_t at = 0;
  auto take = [&](size_t bytes) {
    size_t off = at;
    at += (std::max<size_t>(bytes, 16) + 255) / 256 * 256;
    return off;
  };
  uint32_t m = std::max(n, capacity);
  L.sort_parts = std::max<uint32_t>(1u, div_up(m, gs::kSortPartition));
  L.scan_parts = div_up(m, gs::kScanPartition);
  L.task_capacity = task_capacity_for(c, capacity);
  L.unit_capacity = unit_capacity_for(L.task_capacity, capacity, tiles);
  size_t px = (size_t)c.tile_size * c.tile_size;
  L.state = take(sizeof(DeviceState));
  L.camera = take(gs::kCameraBytes);
  L.pmean = take((size_t)n * 16);
  L.splat = take((size_t)n * sizeof(gs::Splat));
  L.tile_counts = take((size_t)n * 4);
  L.hits = take((size_t)n * 8);
  L.ordered_hits = take((size_t)n * 8);
  L.visible_offsets = take((size_t)n * 4);
  L.depth_a = take((size_t)n * 8);
  L.depth_b = take((size_t)n * 8);
  L.ordered_counts = take((size_t)n * 4);
  L.intersection_offsets = take((size_t)n * 4);
  L.records_a = take((size_t)capacity * 8);
  L.records_b = take((size_t)capacity * 8);
  L.hist = take((gs::kSortTotalsWords + (size_t)2 * L.sort_parts * 256) * 4);
  L.partials = take((size_t)8 * (L.scan_parts + 1));
  L.tile_offsets = take(((size_t)tiles + 1) * 4);
  L.chunk_first = take(((size_t)tiles + 1) * 4);
  L.tasks = take((size_t)L.unit_capacity * 8);
  L.partials_px = take((size_t)L.task_capacity * px * 16);
  L.partial_depth = take((size_t)L.task_capacity * px * 4);
  L.total = at;
  return L;
}

// The frame's camera uniforms (passed by value, so a captured launch keeps
// them) into the scratch, and the parts of the scratch that must start at
// zero: the frame state, the depth sort's digit totals (the depth sort's
// counting kernel zeroes the tile sort's) and the scan look-back status
// words.
__global__ __launch_bounds__(256) void k_noalloc_begin(FrameCamera cam, FrameCamera* cam_out,
                                                        uint32_t* state_words,
                                                        uint32_t state_count,
                                                        uint32_t* depth_totals,
                                                        unsigned long long* scan_status,
                                                        uint32_t scan_count) {
  uint32_t t = threadIdx.x;
  if (t == 0) *cam_out = cam;
  for (uint32_t i = t; i < state_count; i += blockDim.x) state_words[i] = 0u;
  for (uint32_t i = t; i < (uint32_t)(gs::kSortTotalsWords / 2); i += blockDim.x)
    depth_totals[i] = 0u;
  for (uint32_t i = t; i < scan_count; i += blockDim.x) scan_status[i] = 0ull;
}

}  // namespace

extern "C" int gs_render_noalloc(const gs_config* config, const gs_data_desc* data,
                                 uint32_t object_count, const gs_frame* frame,
                                 const gs_image* out, void* stream, gs_state* state) {
  if (!config || !data || !frame || !out) return GS_E_RANGE;
  const char* msg = nullptr;  // checks shared with the context path; no message is kept
  int rc = validate_config(config, msg);
  if (rc) return rc;
  const gs_config& c = *config;
  if (data->geometry_format != GS_GEOMETRY_COMPACT || data->sh_format != GS_SH_RGB8E8 ||
      data->memory != GS_MEMORY_DEVICE || out->memory != GS_MEMORY_DEVICE)
    return GS_E_TYPE;
  if (data->sh_degree > 3 || data->active_count > data->count) return GS_E_RANGE;
  if (data->active_count > 0 &&
      (!data->means || !data->scales_opacity || !data->rotations || !data->sh))
    return GS_E_RANGE;
  if (!out->color || (c.output_depth && !out->depth)) return GS_E_RANGE;
  if (object_count > 0 && !frame->object_frames) return GS_E_RANGE;
  if ((rc = check_frame(c.tile_size, frame, msg))) return rc;
  uint32_t tiles = div_up(frame->width, c.tile_size) * div_up(frame->height, c.tile_size);
  uint32_t n = data->active_count, capacity = c.intersection_capacity;
  NoAllocLayout L = noalloc_layout(c, n, capacity, tiles);
  if (L.total > (size_t)GS_NOALLOC_SCRATCH_BYTES) return GS_E_RANGE;
  if (hipSetDevice(c.device) != hipSuccess) {
    (void)hipGetLastError();
    return GS_E_DEVICE;
  }
  unsigned char* base = nullptr;
  if ((rc = scratch_address(c.device, &base))) return rc;
  int cus = 0;
  uint32_t cu_count = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) ==
          hipSuccess &&
      cus > 0)
    cu_count = (uint32_t)cus;
  FrameBuffers b;
  memset(&b, 0, sizeof(b));
  FrameCamera* camera = (FrameCamera*)(base + L.camera);
  fill_params(c, cu_count, frame, n, object_count, data->sh_degree, data->active_slots != nullptr,
              capacity, camera, b.params);
  b.state = (DeviceState*)(base + L.state);
  b.scene.means = (const float4*)data->means;
  b.scene.scales_opacity = (const uint2*)data->scales_opacity;
  b.scene.rotations = data->rotations;
  b.scene.sh = data->sh;
  b.scene.active_slots = data->active_slots;
  b.scene.objects = frame->object_frames;
  b.proj.mean = (float4*)(base + L.pmean);
  b.proj.splat = (gs::Splat*)(base + L.splat);
  b.proj.tile_counts = (uint32_t*)(base + L.tile_counts);
  b.proj.hits = (uint2*)(base + L.hits);
  b.ordered_hits = (uint2*)(base + L.ordered_hits);
  b.visible_offsets = (uint32_t*)(base + L.visible_offsets);
  b.depth_a = (uint2*)(base + L.depth_a);
  b.depth_b = (uint2*)(base + L.depth_b);
  b.ordered_counts = (uint32_t*)(base + L.ordered_counts);
  b.intersection_offsets = (uint32_t*)(base + L.intersection_offsets);
  b.records_a = (uint2*)(base + L.records_a);
  b.records_b = (uint2*)(base + L.records_b);
  b.tile_offsets = (uint32_t*)(base + L.tile_offsets);
  b.chunk_first = (uint32_t*)(base + L.chunk_first);
  b.tasks = (uint2*)(base + L.tasks);
  b.chunk_partials = (float4*)(base + L.partials_px);
  b.partial_depth = (float*)(base + L.partial_depth);
  b.task_capacity = L.task_capacity;
  b.unit_capacity = L.unit_capacity;
  b.hist = (uint32_t*)(base + L.hist);
  b.sort_parts = L.sort_parts;
  b.partials = (uint32_t*)(base + L.partials);
  b.color_out = (uint16_t*)out->color;
  b.depth_out = c.output_depth ? (float*)out->depth : nullptr;
  b.max_n = n;
  b.max_k = capacity;
  b.max_tiles = tiles;
  b.radix_bits = c.radix_bits;
  b.raster_batch = c.raster_batch;
  b.cu_count = cu_count;
  b.depth_passes = div_up(c.depth_sort_mode == GS_DEPTH_FLOAT32 ? 32u : 16u, c.radix_bits);
  b.tile_passes = std::max<uint32_t>(1u, div_up(bits_for(tiles), c.radix_bits));
  FrameCamera cam;
  memset(&cam, 0, sizeof(cam));
  memcpy(cam.projection, frame->projection, sizeof(cam.projection));
  cam.near_plane = frame->near_plane;
  cam.far_plane = frame->far_plane;
  cam.attempt = frame->attempt;
  cam.reset_valid = frame->reset_valid;
  hipStream_t s = (hipStream_t)stream;
  k_noalloc_begin<<<1, 256, 0, s>>>(cam, camera, (uint32_t*)b.state,
                                    (uint32_t)(sizeof(DeviceState) / 4), b.hist,
                                    (unsigned long long*)b.partials, L.scan_parts + 1);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = gs::launch_frame(b, s, gs::RasterStreams{nullptr, nullptr, nullptr},
                                            nullptr);
  if (e != hipSuccess) return GS_E_DEVICE;
  if (state) {
    DeviceState d;
    e = hipMemcpyAsync(&d, b.state, sizeof(d), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return GS_E_DEVICE;
    memset(state, 0, sizeof(*state));
    state->attempt = d.out_attempt;
    state->version = d.out_version;
    state->requested = d.out_demand;
    state->overflow = d.out_overflow;
    state->visible = d.visible;
    state->intersections = d.count;
    state->capacity = capacity;
    state->tiles_x = b.params.tiles_x;
    state->tiles_y = b.params.tiles_y;
    state->raster_checked = d.checked;
    state->raster_blended = d.blended;
    state->raster_checked_heavy = d.checked_heavy;
    state->error_flags = d.sort_error;
  }
  return GS_OK;
}

// ------------------------------------------------------------ learned cull ----

namespace {

int cull_require(gs_ctx* ctx) {
  return ctx->cull.ready ? GS_OK : fail(ctx, GS_E_STATE, "learned cull not initialised");
}

// Host image of the network from fp32 parameters (bf16 weights).
void pack_net(const float* p, gs::CullNet& n) {
  for (int k = 0; k < gs::kCullHidden * gs::kCullIn; k++) n.w1[k] = gs::gs_f2bf(p[gs::kOffW1 + k]);
  for (int k = 0; k < gs::kCullHidden * gs::kCullHidden; k++)
    n.w2[k] = gs::gs_f2bf(p[gs::kOffW2 + k]);
  for (int k = 0; k < gs::kCullHidden; k++) {
    n.w3[k] = gs::gs_f2bf(p[gs::kOffW3 + k]);
    n.b1[k] = p[gs::kOffB1 + k];
    n.b2[k] = p[gs::kOffB2 + k];
  }
  n.b3 = p[gs::kOffB3];
}

int cull_sync_copy(gs_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind,
                   const char* what) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? GS_OK : hip_fail(ctx, e, what);
}

// Replace the weights and biases of the device network, keeping its
// threshold and normalisation.
int cull_upload_params(gs_ctx* ctx, const float* params) {
  CullState& c = ctx->cull;
  gs::CullNet n;
  int rc = cull_sync_copy(ctx, &n, c.net.ptr, sizeof(n), hipMemcpyDeviceToHost, "cull network");
  if (rc) return rc;
  pack_net(params, n);
  if ((rc = cull_sync_copy(ctx, c.net.ptr, &n, sizeof(n), hipMemcpyHostToDevice,
                           "cull network")))
    return rc;
  return cull_sync_copy(ctx, c.params.ptr, params, gs::kCullParams * sizeof(float),
                        hipMemcpyHostToDevice, "cull parameters");
}

}  // namespace

extern "C" int gs_cull_init(gs_ctx* ctx, uint64_t seed) {
  if (!ctx) return GS_E_RANGE;
  (void)hipSetDevice(ctx->cfg.device);
  CullState& c = ctx->cull;
  size_t pb = gs::kCullParams * sizeof(float);
  int rc;
  if ((rc = ensure(ctx, c.net, sizeof(gs::CullNet), "cull network")) ||
      (rc = ensure(ctx, c.params, pb, "cull parameters")) ||
      (rc = ensure(ctx, c.m, pb, "cull adam m")) || (rc = ensure(ctx, c.v, pb, "cull adam v")) ||
      (rc = ensure(ctx, c.grad, pb, "cull gradient")) ||
      (rc = ensure(ctx, c.loss, sizeof(float), "cull loss")) ||
      (rc = ensure(ctx, c.moments, 2 * gs::kCullIn * sizeof(double), "cull moments")))
    return rc;
  // He-uniform weights from splitmix64, zero biases, output bias at the
  // middle of the target range.
  std::vector<float> p(gs::kCullParams, 0.0f);
  uint64_t st = seed;
  auto uniform = [&]() {
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  double a1 = sqrt(6.0 / gs::kCullIn), a2 = sqrt(6.0 / gs::kCullHidden);
  for (int k = 0; k < gs::kCullHidden * gs::kCullIn; k++) p[gs::kOffW1 + k] = (float)(a1 * uniform());
  for (int k = 0; k < gs::kCullHidden * gs::kCullHidden; k++)
    p[gs::kOffW2 + k] = (float)(a2 * uniform());
  for (int k = 0; k < gs::kCullHidden; k++) p[gs::kOffW3 + k] = (float)(0.1 * a2 * uniform());
  p[gs::kOffB3] = -8.0f;
  gs::CullNet n;
  memset(&n, 0, sizeof(n));
  pack_net(p.data(), n);
  n.log_threshold = logf(1e-4f);
  for (int k = 0; k < gs::kCullIn; k++) {
    n.mu[k] = 0.0f;
    n.inv_sigma[k] = 1.0f;
  }
  if ((rc = cull_sync_copy(ctx, c.net.ptr, &n, sizeof(n), hipMemcpyHostToDevice, "cull network")) ||
      (rc = cull_sync_copy(ctx, c.params.ptr, p.data(), pb, hipMemcpyHostToDevice,
                           "cull parameters")))
    return rc;
  hipError_t e = hipMemsetAsync(c.m.ptr, 0, pb, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c.v.ptr, 0, pb, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c.grad.ptr, 0, pb, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cull state clear");
  c.step = 0;
  c.threshold = 1e-4f;
  c.ready = true;
  return GS_OK;
}

extern "C" int gs_cull_set_params(gs_ctx* ctx, const float* params) {
  if (!ctx || !params) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  for (int k = 0; k < gs::kCullParams; k++)
    if (!finite(params[k])) return fail(ctx, GS_E_RANGE, "cull parameters must be finite");
  return cull_upload_params(ctx, params);
}

extern "C" int gs_cull_get_params(gs_ctx* ctx, float* params) {
  if (!ctx || !params) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  return cull_sync_copy(ctx, params, ctx->cull.params.ptr, gs::kCullParams * sizeof(float),
                        hipMemcpyDeviceToHost, "cull parameters");
}

extern "C" int gs_cull_set_normalization(gs_ctx* ctx, const float mu[16],
                                         const float inv_sigma[16]) {
  if (!ctx || !mu || !inv_sigma) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  for (int k = 0; k < gs::kCullIn; k++)
    if (!finite(mu[k]) || !finite(inv_sigma[k]))
      return fail(ctx, GS_E_RANGE, "normalisation must be finite");
  char* base = (char*)ctx->cull.net.ptr;
  if ((rc = cull_sync_copy(ctx, base + offsetof(gs::CullNet, mu), mu, 16 * sizeof(float),
                           hipMemcpyHostToDevice, "cull normalisation")))
    return rc;
  return cull_sync_copy(ctx, base + offsetof(gs::CullNet, inv_sigma), inv_sigma,
                        16 * sizeof(float), hipMemcpyHostToDevice, "cull normalisation");
}

extern "C" int gs_cull_get_normalization(gs_ctx* ctx, float mu[16], float inv_sigma[16]) {
  if (!ctx || !mu || !inv_sigma) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  gs::CullNet n;
  if ((rc = cull_sync_copy(ctx, &n, ctx->cull.net.ptr, sizeof(n), hipMemcpyDeviceToHost,
                           "cull network")))
    return rc;
  memcpy(mu, n.mu, sizeof(n.mu));
  memcpy(inv_sigma, n.inv_sigma, sizeof(n.inv_sigma));
  return GS_OK;
}

// Training checkpoint: a 32-byte header (magic, version, parameter and
// feature counts, Adam steps taken, threshold) and the f32 arrays
// params, Adam m, Adam v, mu, inv_sigma, little-endian as the host stores
// them.
namespace {
constexpr char kCheckpointMagic[8] = {'G', 'S', 'C', 'U', 'L', 'L', '0', '1'};
constexpr uint32_t kCheckpointVersion = 1;
struct CheckpointHeader {
  char magic[8];
  uint32_t version, params, features, step;
  float threshold;  // alpha * T, as gs_cull_enable takes it
  uint32_t reserved;
};
static_assert(sizeof(CheckpointHeader) == 32, "checkpoint header layout");
constexpr size_t kCheckpointBytes =
    sizeof(CheckpointHeader) + (3 * (size_t)gs::kCullParams + 2 * gs::kCullIn) * sizeof(float);
static_assert(kCheckpointBytes == GS_CULL_CHECKPOINT_BYTES, "checkpoint size");
}  // namespace

extern "C" int gs_cull_save(gs_ctx* ctx, void* dst, size_t bytes) {
  if (!ctx || !dst) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  if (bytes != kCheckpointBytes) return fail(ctx, GS_E_RANGE, "checkpoint buffer size mismatch");
  CullState& c = ctx->cull;
  gs::CullNet n;
  if ((rc = cull_sync_copy(ctx, &n, c.net.ptr, sizeof(n), hipMemcpyDeviceToHost, "cull network")))
    return rc;
  CheckpointHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, kCheckpointMagic, sizeof(h.magic));
  h.version = kCheckpointVersion;
  h.params = gs::kCullParams;
  h.features = gs::kCullIn;
  h.step = c.step;
  h.threshold = c.threshold;
  char* out = (char*)dst;
  memcpy(out, &h, sizeof(h));
  float* f = (float*)(out + sizeof(h));
  const size_t pb = gs::kCullParams * sizeof(float);
  if ((rc = cull_sync_copy(ctx, f, c.params.ptr, pb, hipMemcpyDeviceToHost, "cull parameters")) ||
      (rc = cull_sync_copy(ctx, f + gs::kCullParams, c.m.ptr, pb, hipMemcpyDeviceToHost,
                           "cull adam m")) ||
      (rc = cull_sync_copy(ctx, f + 2 * gs::kCullParams, c.v.ptr, pb, hipMemcpyDeviceToHost,
                           "cull adam v")))
    return rc;
  memcpy(f + 3 * gs::kCullParams, n.mu, sizeof(n.mu));
  memcpy(f + 3 * gs::kCullParams + gs::kCullIn, n.inv_sigma, sizeof(n.inv_sigma));
  return GS_OK;
}

extern "C" int gs_cull_load(gs_ctx* ctx, const void* src, size_t bytes) {
  if (!ctx || !src) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  if (bytes != kCheckpointBytes) return fail(ctx, GS_E_RANGE, "checkpoint size mismatch");
  CheckpointHeader h;
  memcpy(&h, src, sizeof(h));
  if (memcmp(h.magic, kCheckpointMagic, sizeof(h.magic)) != 0 || h.version != kCheckpointVersion ||
      h.params != (uint32_t)gs::kCullParams || h.features != (uint32_t)gs::kCullIn)
    return fail(ctx, GS_E_RANGE, "not a learned-cull checkpoint of this network");
  // Everything is checked before anything is replaced.
  std::vector<float> f(3 * gs::kCullParams + 2 * gs::kCullIn);
  memcpy(f.data(), (const char*)src + sizeof(h), f.size() * sizeof(float));
  if (!(h.threshold > 0.0f && h.threshold < 1.0f))
    return fail(ctx, GS_E_RANGE, "checkpoint threshold out of range");
  for (size_t k = 0; k < f.size(); k++)
    if (!finite(f[k])) return fail(ctx, GS_E_RANGE, "checkpoint values must be finite");
  for (int k = 0; k < gs::kCullParams; k++)
    if (f[2 * gs::kCullParams + k] < 0.0f)
      return fail(ctx, GS_E_RANGE, "checkpoint second moments must be non-negative");
  CullState& c = ctx->cull;
  gs::CullNet n;
  if ((rc = cull_sync_copy(ctx, &n, c.net.ptr, sizeof(n), hipMemcpyDeviceToHost, "cull network")))
    return rc;
  pack_net(f.data(), n);
  n.log_threshold = logf(h.threshold);  // as gs_cull_enable sets it
  memcpy(n.mu, f.data() + 3 * gs::kCullParams, sizeof(n.mu));
  memcpy(n.inv_sigma, f.data() + 3 * gs::kCullParams + gs::kCullIn, sizeof(n.inv_sigma));
  const size_t pb = gs::kCullParams * sizeof(float);
  if ((rc = cull_sync_copy(ctx, c.net.ptr, &n, sizeof(n), hipMemcpyHostToDevice, "cull network")) ||
      (rc = cull_sync_copy(ctx, c.params.ptr, f.data(), pb, hipMemcpyHostToDevice,
                           "cull parameters")) ||
      (rc = cull_sync_copy(ctx, c.m.ptr, f.data() + gs::kCullParams, pb, hipMemcpyHostToDevice,
                           "cull adam m")) ||
      (rc = cull_sync_copy(ctx, c.v.ptr, f.data() + 2 * gs::kCullParams, pb,
                           hipMemcpyHostToDevice, "cull adam v")))
    return rc;
  c.step = h.step;
  c.threshold = h.threshold;
  return GS_OK;
}

extern "C" int gs_cull_enable(gs_ctx* ctx, int enabled, float threshold) {
  if (!ctx) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  if (!(threshold > 0.0f && threshold < 1.0f))
    return fail(ctx, GS_E_RANGE, "cull threshold must be in (0, 1)");
  float lt = logf(threshold);
  char* base = (char*)ctx->cull.net.ptr;
  if ((rc = cull_sync_copy(ctx, base + offsetof(gs::CullNet, log_threshold), &lt, sizeof(lt),
                           hipMemcpyHostToDevice, "cull threshold")))
    return rc;
  ctx->cull.enabled = enabled != 0;
  ctx->cull.threshold = threshold;
  return GS_OK;
}

extern "C" int gs_cull_reserve(gs_ctx* ctx, uint32_t capacity) {
  if (!ctx || capacity == 0) return GS_E_RANGE;
  (void)hipSetDevice(ctx->cfg.device);
  CullState& c = ctx->cull;
  int rc;
  if ((rc = ensure(ctx, c.ds_feats, (size_t)capacity * gs::kCullIn * sizeof(float),
                   "cull training features")) ||
      (rc = ensure(ctx, c.ds_targets, (size_t)capacity * sizeof(float), "cull training targets")) ||
      (rc = ensure(ctx, c.ds_count, sizeof(uint32_t), "cull training count")))
    return rc;
  hipError_t e = hipMemsetAsync(c.ds_count.ptr, 0, sizeof(uint32_t), ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cull training count");
  c.ds_capacity = capacity;
  return GS_OK;
}

extern "C" int gs_cull_dataset_size(gs_ctx* ctx, uint32_t* n) {
  if (!ctx || !n) return GS_E_RANGE;
  if (!ctx->cull.ds_capacity) {
    *n = 0;
    return GS_OK;
  }
  uint32_t v = 0;
  int rc = cull_sync_copy(ctx, &v, ctx->cull.ds_count.ptr, sizeof(v), hipMemcpyDeviceToHost,
                          "cull training count");
  if (rc) return rc;
  *n = v < ctx->cull.ds_capacity ? v : ctx->cull.ds_capacity;
  return GS_OK;
}

// An instrumented plain frame of the single-camera path, then a hashed
// sample (fraction `rate`) of its visible Gaussians appended to the set.
// An overflowing frame appends nothing; *state reports it.
extern "C" int gs_cull_collect(gs_ctx* ctx, const gs_frame* f, float rate, uint32_t seed,
                               gs_state* state) {
  if (!ctx || !f) return GS_E_RANGE;
  if (!ctx->cull.ds_capacity) return fail(ctx, GS_E_STATE, "no training set reserved");
  if (!(rate > 0.0f && rate <= 1.0f)) return fail(ctx, GS_E_RANGE, "rate must be in (0, 1]");
  if (!ctx->has_data) return fail(ctx, GS_E_STATE, "render before data");
  (void)hipSetDevice(ctx->cfg.device);
  CullState& c = ctx->cull;
  size_t a = ctx->active_count ? ctx->active_count : 1;
  int rc;
  if ((rc = ensure(ctx, c.feats, a * gs::kCullIn * sizeof(float), "cull features")) ||
      (rc = ensure(ctx, c.record, a * sizeof(uint32_t), "contribution record")))
    return rc;
  c.collecting = true;
  rc = render_set(ctx, ctx->single, 1, f, nullptr, true);
  c.collecting = false;
  if (rc) return rc;
  gs_state st;
  if ((rc = gs_read_state(ctx, &st))) return rc;
  if (state) *state = st;
  if (st.overflow) return GS_OK;
  double keep = (double)rate * 4294967296.0;
  uint32_t k = keep >= 4294967295.0 ? 0xffffffffu : (uint32_t)keep;
  hipError_t e = gs::cull_collect((const float4*)ctx->single.frames[0].proj.mean,
                                  (const float*)c.feats.ptr, (const uint32_t*)c.record.ptr,
                                  ctx->active_count, k, seed, (float*)c.ds_feats.ptr,
                                  (float*)c.ds_targets.ptr, (uint32_t*)c.ds_count.ptr,
                                  c.ds_capacity, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cull collect");
  return GS_OK;
}

extern "C" int gs_cull_read_dataset(gs_ctx* ctx, uint32_t first, uint32_t n, float* features,
                                    float* targets) {
  if (!ctx) return GS_E_RANGE;
  uint32_t size = 0;
  int rc = gs_cull_dataset_size(ctx, &size);
  if (rc) return rc;
  if ((uint64_t)first + n > size) return fail(ctx, GS_E_RANGE, "range exceeds the training set");
  if (n == 0) return GS_OK;
  const CullState& c = ctx->cull;
  if (features &&
      (rc = cull_sync_copy(ctx, features, (const float*)c.ds_feats.ptr + (size_t)first * gs::kCullIn,
                           (size_t)n * gs::kCullIn * sizeof(float), hipMemcpyDeviceToHost,
                           "cull training features")))
    return rc;
  if (targets &&
      (rc = cull_sync_copy(ctx, targets, (const float*)c.ds_targets.ptr + first,
                           (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                           "cull training targets")))
    return rc;
  return GS_OK;
}

// Feature standardisation from the training set: mean and 1 / std (std
// floored at 1e-3).
extern "C" int gs_cull_fit_normalization(gs_ctx* ctx) {
  if (!ctx) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  uint32_t n = 0;
  if ((rc = gs_cull_dataset_size(ctx, &n))) return rc;
  if (n < 2) return fail(ctx, GS_E_STATE, "training set too small");
  CullState& c = ctx->cull;
  hipError_t e = hipMemsetAsync(c.moments.ptr, 0, 2 * gs::kCullIn * sizeof(double), ctx->stream);
  if (e == hipSuccess)
    e = gs::cull_moments((const float*)c.ds_feats.ptr, n, (double*)c.moments.ptr, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cull moments");
  double mom[2 * gs::kCullIn];
  if ((rc = cull_sync_copy(ctx, mom, c.moments.ptr, sizeof(mom), hipMemcpyDeviceToHost,
                           "cull moments")))
    return rc;
  float mu[gs::kCullIn], inv[gs::kCullIn];
  for (int k = 0; k < gs::kCullIn; k++) {
    double m = mom[k] / n;
    double var = mom[gs::kCullIn + k] / n - m * m;
    double sd = sqrt(var > 0.0 ? var : 0.0);
    mu[k] = (float)m;
    inv[k] = (float)(1.0 / (sd > 1e-3 ? sd : 1e-3));
  }
  return gs_cull_set_normalization(ctx, mu, inv);
}

namespace {

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace

// `steps` Adam steps on minibatches of `batch` samples drawn by a fresh
// affine permutation of the training set each step.
extern "C" int gs_cull_train(gs_ctx* ctx, uint32_t steps, uint32_t batch, float lr,
                             float under_weight, uint64_t seed, float* last_loss) {
  if (!ctx || batch == 0) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  if (!(lr > 0.0f && finite(lr)) || !(under_weight >= 0.0f && finite(under_weight)))
    return fail(ctx, GS_E_RANGE, "learning rate must be positive, under-weight nonnegative");
  uint32_t n = 0;
  if ((rc = gs_cull_dataset_size(ctx, &n))) return rc;
  if (n == 0) return fail(ctx, GS_E_STATE, "empty training set");
  CullState& c = ctx->cull;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  uint64_t st = seed;
  auto next = [&]() {
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  hipError_t e = hipSuccess;
  for (uint32_t s = 0; s < steps && e == hipSuccess; s++) {
    uint64_t a = 1;
    if (n > 1)
      do a = next() % n; while (a == 0 || gcd64(a, n) != 1);
    uint32_t b = (uint32_t)(next() % n);
    e = hipMemsetAsync(c.loss.ptr, 0, sizeof(float), ctx->stream);
    if (e == hipSuccess)
      e = gs::cull_train_step((const float*)c.ds_feats.ptr, (const float*)c.ds_targets.ptr, n,
                              (uint32_t)a, b, 0u, batch, (const gs::CullNet*)c.net.ptr,
                              under_weight, (float*)c.grad.ptr, (float*)c.loss.ptr, ctx->stream);
    c.step++;
    float corr1 = 1.0f - powf(beta1, (float)c.step), corr2 = 1.0f - powf(beta2, (float)c.step);
    if (e == hipSuccess)
      e = gs::cull_adam((float*)c.params.ptr, (float*)c.m.ptr, (float*)c.v.ptr,
                        (float*)c.grad.ptr, lr, beta1, beta2, eps, corr1, corr2,
                        (gs::CullNet*)c.net.ptr, ctx->stream);
  }
  if (e != hipSuccess) return hip_fail(ctx, e, "cull training");
  float loss = 0.0f;
  if ((rc = cull_sync_copy(ctx, &loss, c.loss.ptr, sizeof(loss), hipMemcpyDeviceToHost,
                           "cull loss")))
    return rc;
  if (last_loss) *last_loss = loss / (float)batch;
  return GS_OK;
}

// The network's predictions for n host feature rows (raw features).
extern "C" int gs_cull_eval(gs_ctx* ctx, uint32_t n, const float* features, float* out) {
  if (!ctx || (n > 0 && (!features || !out))) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc || n == 0) return rc;
  CullState& c = ctx->cull;
  if ((rc = ensure(ctx, c.tmp_in, (size_t)n * gs::kCullIn * sizeof(float), "cull eval input")) ||
      (rc = ensure(ctx, c.tmp_out, (size_t)n * sizeof(float), "cull eval output")) ||
      (rc = cull_sync_copy(ctx, c.tmp_in.ptr, features, (size_t)n * gs::kCullIn * sizeof(float),
                           hipMemcpyHostToDevice, "cull eval input")))
    return rc;
  hipError_t e = gs::cull_eval((const float*)c.tmp_in.ptr, n, (const gs::CullNet*)c.net.ptr,
                               (float*)c.tmp_out.ptr, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cull eval");
  return cull_sync_copy(ctx, out, c.tmp_out.ptr, (size_t)n * sizeof(float),
                        hipMemcpyDeviceToHost, "cull eval output");
}

// Gradient of the training loss over n host samples, in parameter order
// (no parameter update), and the loss.
extern "C" int gs_cull_gradient(gs_ctx* ctx, uint32_t n, const float* features,
                                const float* targets, float under_weight, float* grad,
                                float* loss) {
  if (!ctx || n == 0 || !features || !targets || !grad) return GS_E_RANGE;
  int rc = cull_require(ctx);
  if (rc) return rc;
  CullState& c = ctx->cull;
  size_t pb = gs::kCullParams * sizeof(float);
  if ((rc = ensure(ctx, c.tmp_in, (size_t)n * gs::kCullIn * sizeof(float), "cull grad input")) ||
      (rc = ensure(ctx, c.tmp_target, (size_t)n * sizeof(float), "cull grad target")) ||
      (rc = cull_sync_copy(ctx, c.tmp_in.ptr, features, (size_t)n * gs::kCullIn * sizeof(float),
                           hipMemcpyHostToDevice, "cull grad input")) ||
      (rc = cull_sync_copy(ctx, c.tmp_target.ptr, targets, (size_t)n * sizeof(float),
                           hipMemcpyHostToDevice, "cull grad target")))
    return rc;
  hipError_t e = hipMemsetAsync(c.grad.ptr, 0, pb, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c.loss.ptr, 0, sizeof(float), ctx->stream);
  if (e == hipSuccess)
    e = gs::cull_train_step((const float*)c.tmp_in.ptr, (const float*)c.tmp_target.ptr, n, 1u, 0u,
                            0u, n, (const gs::CullNet*)c.net.ptr, under_weight,
                            (float*)c.grad.ptr, (float*)c.loss.ptr, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "cull gradient");
  if ((rc = cull_sync_copy(ctx, grad, c.grad.ptr, pb, hipMemcpyDeviceToHost, "cull gradient")))
    return rc;
  float l = 0.0f;
  if ((rc = cull_sync_copy(ctx, &l, c.loss.ptr, sizeof(l), hipMemcpyDeviceToHost, "cull loss")))
    return rc;
  if (loss) *loss = l / (float)n;
  e = hipMemsetAsync(c.grad.ptr, 0, pb, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? GS_OK : hip_fail(ctx, e, "cull gradient clear");
}
