// gs_api.hip - context, device memory and the C ABI of include/gs.h.
//
// The context plays the role of the reference's TiledGaussianPipeline
// (pipeline/TiledGaussianPipeline.ts:39-484): it owns every device buffer,
// grows them geometrically and never shrinks them (AttributePool.ts:41-67),
// and enqueues one frame per gs_render on its own HIP stream.
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/gs.h"
#include "gs_internal.h"

using gs::DeviceState;
using gs::FrameBuffers;
using gs::FrameParams;

namespace {

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct gs_ctx {
  gs_config cfg;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;  // second raster branch, forked per frame
  hipEvent_t fork = nullptr, join = nullptr;
  std::string error;
  // Scene (GaussianData contract).
  uint32_t count = 0, active_count = 0, sh_degree = 0;
  bool has_data = false, has_active = false;
  DeviceBuffer means, scales_opacity, rotations, sh, active, objects;
  uint32_t object_count = 0;
  // Per-Gaussian frame arrays, sized for slot_capacity Gaussians.
  uint32_t slot_capacity = 0;
  DeviceBuffer pmean, pconic, pcolor, tile_counts, visible_offsets, depth_a, depth_b,
      ordered_counts, intersection_offsets;
  // Intersection records, sized for `capacity`.
  uint32_t capacity = 0;
  DeviceBuffer records_a, records_b;
  // Radix histograms and scan partials, sized for max(slots, capacity).
  DeviceBuffer hist, partials;
  // Tile ranges, sized for tile_capacity + 1.
  uint32_t tile_capacity = 0;
  DeviceBuffer tile_offsets, chunk_first;
  // Raster chunk tasks and partials (RasterChunkWorkspace.ts:12-59).
  DeviceBuffer tasks, partials_px, partial_depth;
  // Output images.
  uint32_t image_w = 0, image_h = 0;
  DeviceBuffer color, depth;
  DeviceBuffer state;
  // Object table: host copy, uploaded through a ring of pinned slots.
  std::vector<float> objects_host;
  bool objects_dirty = false;
  static constexpr int kRing = 8;
  float* ring[kRing] = {};
  hipEvent_t ring_done[kRing] = {};
  size_t ring_bytes = 0;
  int ring_next = 0;
  // Captured frame graph and the launch description it was captured from.
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  FrameBuffers graph_key = {};
  bool graph_valid = false;
  // Stage profiling.
  std::vector<hipEvent_t> marks;
  uint32_t profile_capacity = 0, profile_frames = 0;
  // Last frame.
  FrameBuffers last = {};
  bool rendered = false;
  uint32_t width = 0, height = 0, tiles_x = 0, tiles_y = 0;
};

namespace {

int fail(gs_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return code;
}

int hip_fail(gs_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, GS_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Grow-only device allocation (AttributePool.ts:41-67 semantics).
int ensure(gs_ctx* ctx, DeviceBuffer& b, size_t bytes, const char* what) {
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return GS_OK;
  if (b.ptr) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
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

int ensure_slots(gs_ctx* ctx, uint32_t n) {
  if (n <= ctx->slot_capacity && ctx->pmean.ptr) return GS_OK;
  uint32_t cap = ctx->slot_capacity ? ctx->slot_capacity : 1024u;
  while (cap < n) cap = cap > 0x7fffffffu ? n : cap * 2u;
  int rc;
  size_t c = cap;
  if ((rc = ensure(ctx, ctx->pmean, c * 16, "projected-mean")) ||
      (rc = ensure(ctx, ctx->pconic, c * 16, "projected-conic")) ||
      (rc = ensure(ctx, ctx->pcolor, c * 16, "projected-color")) ||
      (rc = ensure(ctx, ctx->tile_counts, c * 4, "tile-counts")) ||
      (rc = ensure(ctx, ctx->visible_offsets, c * 4, "visible-offsets")) ||
      (rc = ensure(ctx, ctx->depth_a, c * 8, "depth-records-a")) ||
      (rc = ensure(ctx, ctx->depth_b, c * 8, "depth-records-b")) ||
      (rc = ensure(ctx, ctx->ordered_counts, c * 4, "ordered-tile-counts")) ||
      (rc = ensure(ctx, ctx->intersection_offsets, c * 4, "intersection-offsets")))
    return rc;
  ctx->slot_capacity = cap;
  return GS_OK;
}

int ensure_sort_workspace(gs_ctx* ctx) {
  uint32_t m = ctx->slot_capacity > ctx->capacity ? ctx->slot_capacity : ctx->capacity;
  uint32_t parts = div_up(m, gs::kPartition);
  if (parts == 0) parts = 1;
  int rc;
  if ((rc = ensure(ctx, ctx->hist, (size_t)parts * gs::kRadix * 4, "radix-histograms")) ||
      (rc = ensure(ctx, ctx->partials, ((size_t)parts + 64) * 4, "scan-partials")))
    return rc;
  return GS_OK;
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
}

// GaussianPass.ts:157-260 option validation.
static int validate_config(const gs_config* c, std::string& msg) {
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
  if (c->max_intersection_capacity > 0x7fffffffu) {
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
  return GS_OK;
}

extern "C" int gs_create(const gs_config* config, gs_ctx** out) {
  if (!out) return GS_E_RANGE;
  *out = nullptr;
  gs_config c;
  if (config) c = *config;
  else gs_config_default(&c);
  std::string msg;
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
  // Default stream priority only.
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming) != hipSuccess) {
    gs_destroy(ctx);
    return GS_E_DEVICE;
  }
  ctx->capacity = c.intersection_capacity;
  if ((rc = ensure(ctx, ctx->state, sizeof(DeviceState), "state")) ||
      (rc = ensure(ctx, ctx->records_a, (size_t)ctx->capacity * 8, "records-a")) ||
      (rc = ensure(ctx, ctx->records_b, (size_t)ctx->capacity * 8, "records-b"))) {
    gs_destroy(ctx);
    return rc;
  }
  if (hipMemsetAsync(ctx->state.ptr, 0, sizeof(DeviceState), ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    gs_destroy(ctx);
    return GS_E_DEVICE;
  }
  *out = ctx;
  return GS_OK;
}

extern "C" void gs_destroy(gs_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->cfg.device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DeviceBuffer* all[] = {&ctx->means,        &ctx->scales_opacity,  &ctx->rotations,
                         &ctx->sh,           &ctx->active,          &ctx->objects,
                         &ctx->pmean,        &ctx->pconic,          &ctx->pcolor,
                         &ctx->tile_counts,  &ctx->visible_offsets, &ctx->depth_a,
                         &ctx->depth_b,      &ctx->ordered_counts,  &ctx->intersection_offsets,
                         &ctx->records_a,    &ctx->records_b,       &ctx->hist,
                         &ctx->partials,     &ctx->tile_offsets,    &ctx->color,
                         &ctx->depth,        &ctx->state,           &ctx->chunk_first,
                         &ctx->tasks,        &ctx->partials_px,     &ctx->partial_depth};
  for (DeviceBuffer* b : all) release(*b);
  for (int i = 0; i < gs_ctx::kRing; i++) {
    if (ctx->ring[i]) (void)hipHostFree(ctx->ring[i]);
    if (ctx->ring_done[i]) (void)hipEventDestroy(ctx->ring_done[i]);
  }
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  for (hipEvent_t e : ctx->marks) (void)hipEventDestroy(e);
  if (ctx->fork) (void)hipEventDestroy(ctx->fork);
  if (ctx->join) (void)hipEventDestroy(ctx->join);
  if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
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
      (rc = copy_in(ctx, ctx->sh, d->sh, n * ncoef * 4, d->memory, "shCoefficients")))
    return rc;
  if (d->active_slots && d->memory == GS_MEMORY_HOST)
    for (uint32_t i = 0; i < d->active_count; i++)
      if (d->active_slots[i] >= d->count)
        return fail(ctx, GS_E_RANGE, "active slot out of range");
  ctx->has_active = d->active_slots != nullptr;
  if (ctx->has_active &&
      (rc = copy_in(ctx, ctx->active, d->active_slots, (size_t)d->active_count * 4, d->memory,
                    "activeSlots")))
    return rc;
  if ((rc = ensure_slots(ctx, d->active_count)) || (rc = ensure_sort_workspace(ctx))) return rc;
  ctx->count = d->count;
  ctx->active_count = d->active_count;
  ctx->sh_degree = d->sh_degree;
  ctx->has_data = true;
  ctx->rendered = false;
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
  if ((rc = ensure_slots(ctx, new_active_count)) || (rc = ensure_sort_workspace(ctx))) return rc;
  ctx->active_count = new_active_count;
  return GS_OK;
}

extern "C" int gs_set_objects(gs_ctx* ctx, uint32_t object_count, const float* frames) {
  if (!ctx) return GS_E_RANGE;
  if (object_count > 0 && !frames) return fail(ctx, GS_E_RANGE, "missing object frames");
  ctx->objects_host.assign(frames, frames + (size_t)object_count * 40);
  ctx->object_count = object_count;
  ctx->objects_dirty = true;
  return GS_OK;
}

// Enqueue the object table upload (ObjectFrameState.update) without
// blocking: the host copy goes to the next pinned ring slot once the copy
// that last used that slot has completed.
static int upload_objects(gs_ctx* ctx) {
  if (!ctx->objects_dirty) return GS_OK;
  size_t bytes = ctx->objects_host.size() * sizeof(float);
  int rc = ensure(ctx, ctx->objects, bytes, "object frames");
  if (rc) return rc;
  if (bytes == 0) {
    ctx->objects_dirty = false;
    return GS_OK;
  }
  if (bytes > ctx->ring_bytes) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    for (int i = 0; i < gs_ctx::kRing; i++) {
      if (ctx->ring[i]) (void)hipHostFree(ctx->ring[i]);
      ctx->ring[i] = nullptr;
      if (hipHostMalloc((void**)&ctx->ring[i], bytes, hipHostMallocDefault) != hipSuccess)
        return fail(ctx, GS_E_OOM, "allocation-failed: object staging");
      if (!ctx->ring_done[i] &&
          hipEventCreateWithFlags(&ctx->ring_done[i], hipEventDisableTiming) != hipSuccess)
        return fail(ctx, GS_E_DEVICE, "hipEventCreate");
    }
    ctx->ring_bytes = bytes;
  }
  int slot = ctx->ring_next;
  ctx->ring_next = (slot + 1) % gs_ctx::kRing;
  hipError_t e = hipEventSynchronize(ctx->ring_done[slot]);
  if (e != hipSuccess) return hip_fail(ctx, e, "object staging wait");
  if (bytes) memcpy(ctx->ring[slot], ctx->objects_host.data(), bytes);
  if (bytes) e = hipMemcpyAsync(ctx->objects.ptr, ctx->ring[slot], bytes, hipMemcpyHostToDevice,
                                ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ctx->ring_done[slot], ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "object upload");
  ctx->objects_dirty = false;
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
  int rc;
  if ((rc = ensure(ctx, ctx->records_a, (size_t)capacity * 8, "records-a")) ||
      (rc = ensure(ctx, ctx->records_b, (size_t)capacity * 8, "records-b")))
    return rc;
  ctx->capacity = capacity;
  return ensure_sort_workspace(ctx);
}

static int prepare_frame(gs_ctx* ctx, const gs_frame* f, FrameBuffers& b, void* color_dst,
                         void* depth_dst) {
  if (!ctx->has_data) return fail(ctx, GS_E_STATE, "render before data");
  if (f->width < 1 || f->height < 1) return fail(ctx, GS_E_RANGE, "empty render size");
  uint32_t tiles_x = div_up(f->width, gs::kTile), tiles_y = div_up(f->height, gs::kTile);
  // TiledGaussianPipeline.ts:294-296
  if (tiles_x > 65535 || tiles_y > 65535)
    return fail(ctx, GS_E_RANGE, "Render size exceeds the tile dispatch limit");
  if (!(f->near_plane > 0.0f && f->far_plane > f->near_plane && finite(f->far_plane)))
    return fail(ctx, GS_E_RANGE, "depthRange must be positive and increasing");
  if (!(f->filter_variance >= 0.0f && finite(f->filter_variance)))
    return fail(ctx, GS_E_RANGE, "filterVariance must be finite and nonnegative");
  if (f->boundary_axis > GS_BOUNDARY_REPEAT_Y)
    return fail(ctx, GS_E_RANGE, "unknown boundaryPolicy");
  for (int i = 0; i < 16; i++)
    if (!finite(f->projection[i])) return fail(ctx, GS_E_RANGE, "projection must be finite");
  int rc;
  uint32_t tiles = tiles_x * tiles_y;
  if (tiles > ctx->tile_capacity || !ctx->tile_offsets.ptr) {
    uint32_t cap = next_pow2(tiles);
    if ((rc = ensure(ctx, ctx->tile_offsets, ((size_t)cap + 1) * 4, "tile-offsets")) ||
        (rc = ensure(ctx, ctx->chunk_first, ((size_t)cap + 1) * 4, "raster-chunk-offsets")))
      return rc;
    ctx->tile_capacity = cap;
  }
  // rasterChunks.ts:73-78 (maxRasterChunkTasks): every chunked tile holds
  // more than chunk_size records, so ceil(2K / chunk_size) tasks suffice.
  uint32_t task_capacity = 1;
  if (ctx->cfg.raster_chunk_size) {
    task_capacity = div_up(2ull * ctx->capacity, ctx->cfg.raster_chunk_size);
    if (task_capacity < 1) task_capacity = 1;
  }
  // Heavy raster units: the chunk tasks plus tiles above 1024 records.
  uint32_t unit_capacity = task_capacity + div_up(ctx->capacity, 1024) + 1;
  if ((rc = ensure(ctx, ctx->tasks, (size_t)unit_capacity * 8, "raster-heavy-units")) ||
      (rc = ensure(ctx, ctx->partials_px, (size_t)task_capacity * gs::kBlock * 16,
                   "raster-chunk-partials")) ||
      (rc = ensure(ctx, ctx->partial_depth, (size_t)task_capacity * gs::kBlock * 4,
                   "raster-chunk-depth")))
    return rc;
  size_t pixels = (size_t)f->width * f->height;
  if ((rc = ensure(ctx, ctx->color, pixels * 8, "color-target"))) return rc;
  if (ctx->cfg.output_depth && (rc = ensure(ctx, ctx->depth, pixels * 4, "depth-target")))
    return rc;
  FrameParams& p = b.params;
  memset(&p, 0, sizeof(p));
  memcpy(p.projection, f->projection, sizeof(p.projection));
  p.viewport[0] = (float)f->width;
  p.viewport[1] = (float)f->height;
  p.viewport[2] = f->near_plane;
  p.viewport[3] = f->far_plane;
  p.filter_variance = f->filter_variance;
  p.transmittance_threshold = ctx->cfg.transmittance_threshold;
  memcpy(p.background, ctx->cfg.background, sizeof(p.background));
  p.depth_alpha_threshold = ctx->cfg.depth_alpha_threshold;
  p.invalid_depth_value = ctx->cfg.invalid_depth_value;
  p.boundary_axis = f->boundary_axis;
  p.active_count = ctx->active_count;
  p.tiles_x = tiles_x;
  p.tiles_y = tiles_y;
  p.tile_count = tiles;
  p.width = f->width;
  p.height = f->height;
  p.capacity = ctx->capacity;
  p.attempt = f->attempt;
  p.reset_valid = f->reset_valid;
  p.object_count = ctx->object_count;
  p.sh_degree = ctx->sh_degree;
  p.depth_mode = ctx->cfg.depth_sort_mode;
  p.antialias_compensated = ctx->cfg.antialias_mode == GS_AA_COMPENSATED ? 1u : 0u;
  p.subpixel_culling = ctx->cfg.subpixel_sample_culling ? 1u : 0u;
  p.chunk_size = ctx->cfg.raster_chunk_size;
  p.has_active_slots = ctx->has_active ? 1u : 0u;
  p.output_depth = ctx->cfg.output_depth ? 1u : 0u;
  p.raster_stats = ctx->cfg.raster_stats ? 1u : 0u;
  b.state = (DeviceState*)ctx->state.ptr;
  b.scene.means = (const float4*)ctx->means.ptr;
  b.scene.scales_opacity = (const uint2*)ctx->scales_opacity.ptr;
  b.scene.rotations = (const uint32_t*)ctx->rotations.ptr;
  b.scene.sh = (const uint32_t*)ctx->sh.ptr;
  b.scene.active_slots = (const uint32_t*)ctx->active.ptr;
  b.scene.objects = (const float*)ctx->objects.ptr;
  b.proj.mean = (float4*)ctx->pmean.ptr;
  b.proj.conic = (float4*)ctx->pconic.ptr;
  b.proj.color = (float4*)ctx->pcolor.ptr;
  b.proj.tile_counts = (uint32_t*)ctx->tile_counts.ptr;
  b.visible_offsets = (uint32_t*)ctx->visible_offsets.ptr;
  b.depth_a = (uint2*)ctx->depth_a.ptr;
  b.depth_b = (uint2*)ctx->depth_b.ptr;
  b.ordered_counts = (uint32_t*)ctx->ordered_counts.ptr;
  b.intersection_offsets = (uint32_t*)ctx->intersection_offsets.ptr;
  b.records_a = (uint2*)ctx->records_a.ptr;
  b.records_b = (uint2*)ctx->records_b.ptr;
  b.tile_offsets = (uint32_t*)ctx->tile_offsets.ptr;
  b.chunk_first = (uint32_t*)ctx->chunk_first.ptr;
  b.tasks = (uint2*)ctx->tasks.ptr;
  b.chunk_partials = (float4*)ctx->partials_px.ptr;
  b.partial_depth = (float*)ctx->partial_depth.ptr;
  b.task_capacity = task_capacity;
  b.unit_capacity = unit_capacity;
  b.hist = (uint32_t*)ctx->hist.ptr;
  b.partials = (uint32_t*)ctx->partials.ptr;
  b.color_out = (uint16_t*)(color_dst ? color_dst : ctx->color.ptr);
  b.depth_out = ctx->cfg.output_depth ? (float*)(depth_dst ? depth_dst : ctx->depth.ptr) : nullptr;
  b.max_n = ctx->active_count;
  b.max_k = ctx->capacity;
  b.max_tiles = tiles;
  b.depth_passes = ctx->cfg.depth_sort_mode == GS_DEPTH_FLOAT32 ? 4u : 2u;
  b.tile_passes = div_up(bits_for(tiles), gs::kRadixBits);
  if (b.tile_passes == 0) b.tile_passes = 1;
  return GS_OK;
}

static int launch(gs_ctx* ctx, const FrameBuffers& b) {
  hipStream_t s = ctx->stream;
  if (ctx->profile_capacity) {
    if (ctx->profile_frames >= ctx->profile_capacity)
      return fail(ctx, GS_E_RANGE, "profiling capacity exhausted");
    hipEvent_t* marks = &ctx->marks[(size_t)ctx->profile_frames * (gs::kStageCount + 1)];
    hipError_t e = gs::launch_frame(b, s, ctx->aux, ctx->fork, ctx->join, marks);
    if (e != hipSuccess) return hip_fail(ctx, e, "frame launch");
    ctx->profile_frames++;
    return GS_OK;
  }
  if (!ctx->cfg.capture_graph) {
    hipError_t e = gs::launch_frame(b, s, ctx->aux, ctx->fork, ctx->join, nullptr);
    if (e != hipSuccess) return hip_fail(ctx, e, "frame launch");
    return GS_OK;
  }
  if (!ctx->graph_valid || memcmp(&ctx->graph_key, &b, sizeof(b)) != 0) {
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamBeginCapture");
    hipError_t le = gs::launch_frame(b, s, ctx->aux, ctx->fork, ctx->join, nullptr);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (le != hipSuccess) e = le;
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      ctx->graph_valid = false;
      return hip_fail(ctx, e, "frame capture");
    }
    bool updated = false;
    if (ctx->graph_exec) {
      hipGraphExecUpdateResult result;
      hipGraphNode_t error_node = nullptr;
      updated = hipGraphExecUpdate(ctx->graph_exec, g, &error_node, &result) == hipSuccess &&
                result == hipGraphExecUpdateSuccess;
      if (!updated) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(ctx->graph_exec);
        ctx->graph_exec = nullptr;
      }
    }
    if (!updated) {
      e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        ctx->graph_valid = false;
        return hip_fail(ctx, e, "hipGraphInstantiate");
      }
    }
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph = g;
    ctx->graph_key = b;
    ctx->graph_valid = true;
  }
  hipError_t e = hipGraphLaunch(ctx->graph_exec, s);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphLaunch");
  return GS_OK;
}

extern "C" int gs_render(gs_ctx* ctx, const gs_frame* f, const gs_image* out) {
  if (!ctx || !f) return GS_E_RANGE;
  (void)hipSetDevice(ctx->cfg.device);
  void* color_dst = nullptr;
  void* depth_dst = nullptr;
  if (out && out->memory == GS_MEMORY_DEVICE) {
    color_dst = out->color;
    depth_dst = out->depth;
  }
  // The upload may grow the object buffer, so it precedes the launch
  // description that captures its address.
  int rc = upload_objects(ctx);
  if (rc) return rc;
  FrameBuffers b;
  memset(&b, 0, sizeof(b));
  if ((rc = prepare_frame(ctx, f, b, color_dst, depth_dst))) return rc;
  if ((rc = launch(ctx, b))) return rc;
  ctx->last = b;
  ctx->rendered = true;
  ctx->width = f->width;
  ctx->height = f->height;
  ctx->tiles_x = b.params.tiles_x;
  ctx->tiles_y = b.params.tiles_y;
  if (out && out->memory == GS_MEMORY_HOST)
    return gs_read_image(ctx, out->color, out->color ? (size_t)f->width * f->height * 8 : 0,
                         out->depth, out->depth ? (size_t)f->width * f->height * 4 : 0);
  return GS_OK;
}

extern "C" int gs_profile_begin(gs_ctx* ctx, uint32_t max_frames) {
  if (!ctx || max_frames == 0) return GS_E_RANGE;
  (void)hipSetDevice(ctx->cfg.device);
  size_t need = (size_t)max_frames * (gs::kStageCount + 1);
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
                              uint32_t* frames) {
  if (!ctx) return GS_E_RANGE;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
  double sums[gs::kStageCount] = {0};
  double total = 0.0;
  for (uint32_t f = 0; f < ctx->profile_frames; f++) {
    hipEvent_t* m = &ctx->marks[(size_t)f * (gs::kStageCount + 1)];
    for (int st = 0; st < gs::kStageCount; st++) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, m[st], m[st + 1]) != hipSuccess)
        return fail(ctx, GS_E_DEVICE, "hipEventElapsedTime");
      sums[st] += ms;
    }
    float whole = 0.0f;
    if (hipEventElapsedTime(&whole, m[0], m[gs::kStageCount]) != hipSuccess)
      return fail(ctx, GS_E_DEVICE, "hipEventElapsedTime");
    total += whole;
  }
  if (stage_ms)
    for (int st = 0; st < gs::kStageCount; st++) stage_ms[st] = sums[st];
  if (frame_ms) *frame_ms = total;
  if (frames) *frames = ctx->profile_frames;
  ctx->profile_capacity = 0;
  ctx->profile_frames = 0;
  return GS_OK;
}

extern "C" int gs_render_batch(gs_ctx* ctx, uint32_t n, const gs_frame* frames,
                               const gs_image* outs) {
  if (!ctx || (n > 0 && !frames)) return GS_E_RANGE;
  for (uint32_t i = 0; i < n; i++) {
    int rc = gs_render(ctx, &frames[i], outs ? &outs[i] : nullptr);
    if (rc) return rc;
  }
  return GS_OK;
}

extern "C" int gs_read_state(gs_ctx* ctx, gs_state* out) {
  if (!ctx || !out) return GS_E_RANGE;
  (void)hipSetDevice(ctx->cfg.device);
  DeviceState s;
  hipError_t e = hipMemcpyAsync(&s, ctx->state.ptr, sizeof(s), hipMemcpyDeviceToHost,
                                ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "state readback");
  memset(out, 0, sizeof(*out));
  out->attempt = s.out_attempt;
  out->version = s.out_version;
  out->requested = s.out_demand;
  out->overflow = s.out_overflow;
  out->visible = s.visible;
  out->intersections = s.count;
  out->capacity = ctx->capacity;
  out->tiles_x = ctx->tiles_x;
  out->tiles_y = ctx->tiles_y;
  out->raster_checked = s.checked;
  out->raster_blended = s.blended;
  return GS_OK;
}

extern "C" int gs_read_image(gs_ctx* ctx, void* color, size_t color_bytes, void* depth,
                             size_t depth_bytes) {
  if (!ctx) return GS_E_RANGE;
  if (!ctx->rendered) return fail(ctx, GS_E_STATE, "nothing rendered");
  (void)hipSetDevice(ctx->cfg.device);
  size_t pixels = (size_t)ctx->width * ctx->height;
  hipError_t e = hipSuccess;
  if (color) {
    if (color_bytes != pixels * 8) return fail(ctx, GS_E_RANGE, "color buffer size mismatch");
    e = hipMemcpyAsync(color, ctx->last.color_out, pixels * 8, hipMemcpyDeviceToHost,
                       ctx->stream);
  }
  if (e == hipSuccess && depth) {
    if (!ctx->last.depth_out) return fail(ctx, GS_E_STATE, "outputDepth is disabled");
    if (depth_bytes != pixels * 4) return fail(ctx, GS_E_RANGE, "depth buffer size mismatch");
    e = hipMemcpyAsync(depth, ctx->last.depth_out, pixels * 4, hipMemcpyDeviceToHost,
                       ctx->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "image readback");
  return GS_OK;
}

extern "C" int gs_read_debug(gs_ctx* ctx, const char* name, void* dst, size_t bytes,
                             size_t* size) {
  if (!ctx || !name) return GS_E_RANGE;
  if (!ctx->rendered) return fail(ctx, GS_E_STATE, "nothing rendered");
  gs_state st;
  int rc = gs_read_state(ctx, &st);
  if (rc) return rc;
  const FrameBuffers& b = ctx->last;
  std::string n(name);
  const void* src = nullptr;
  size_t len = 0;
  size_t a = b.params.active_count;
  if (n == "projected_mean") { src = b.proj.mean; len = a * 16; }
  else if (n == "projected_conic") { src = b.proj.conic; len = a * 16; }
  else if (n == "projected_color") { src = b.proj.color; len = a * 16; }
  else if (n == "tile_counts") { src = b.proj.tile_counts; len = a * 4; }
  else if (n == "visible_offsets") { src = b.visible_offsets; len = a * 4; }
  else if (n == "depth_records") { src = gs::sorted_depth_records(b); len = (size_t)st.visible * 8; }
  else if (n == "ordered_counts") { src = b.ordered_counts; len = (size_t)st.visible * 4; }
  else if (n == "intersection_offsets") { src = b.intersection_offsets; len = (size_t)st.visible * 4; }
  else if (n == "sorted_records") { src = gs::sorted_tile_records(b); len = (size_t)st.intersections * 8; }
  else if (n == "tile_offsets") { src = b.tile_offsets; len = ((size_t)b.params.tile_count + 1) * 4; }
  else return fail(ctx, GS_E_RANGE, "unknown debug buffer: " + n);
  if (size) *size = len;
  if (!dst) return GS_OK;
  if (bytes != len) return fail(ctx, GS_E_RANGE, "debug buffer size mismatch for " + n);
  if (len == 0) return GS_OK;
  hipError_t e = hipMemcpyAsync(dst, src, len, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "debug readback");
  return GS_OK;
}

extern "C" void* gs_stream(gs_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" int gs_device_image(gs_ctx* ctx, void** color, void** depth) {
  if (!ctx) return GS_E_RANGE;
  if (color) *color = ctx->color.ptr;
  if (depth) *depth = ctx->depth.ptr;
  return GS_OK;
}
