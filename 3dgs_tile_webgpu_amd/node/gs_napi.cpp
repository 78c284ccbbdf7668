// gs_napi.cpp - Node N-API binding of the C ABI (include/gs.h).
//
// The reference's host is TypeScript (GaussianPass.ts, TiledGaussianPipeline.ts);
// this addon lets that host call the MI355X library directly. It only
// marshals typed arrays and plain objects into the C structs: host arrays
// are copied during each call and no pointer is retained. Status codes
// become the reference's exception classes: GS_E_RANGE -> RangeError,
// GS_E_TYPE -> TypeError, anything else -> Error with `code` set.
#include <node_api.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/gs.h"

namespace {

napi_value undefined(napi_env env) {
  napi_value v;
  napi_get_undefined(env, &v);
  return v;
}

napi_value throw_status(napi_env env, int rc, const char* msg) {
  napi_value code, text, err;
  napi_create_int32(env, rc, &code);
  napi_create_string_utf8(env, msg && *msg ? msg : "gs error", NAPI_AUTO_LENGTH, &text);
  if (rc == GS_E_RANGE)
    napi_create_range_error(env, nullptr, text, &err);
  else if (rc == GS_E_TYPE)
    napi_create_type_error(env, nullptr, text, &err);
  else
    napi_create_error(env, nullptr, text, &err);
  napi_set_named_property(env, err, "code", code);
  napi_throw(env, err);
  return nullptr;
}

napi_value check(napi_env env, gs_ctx* ctx, int rc) {
  if (rc == GS_OK) return undefined(env);
  return throw_status(env, rc, ctx ? gs_last_error(ctx) : "");
}

bool get_prop(napi_env env, napi_value obj, const char* name, napi_value* out) {
  bool has = false;
  napi_has_named_property(env, obj, name, &has);
  if (!has) return false;
  napi_get_named_property(env, obj, name, out);
  napi_valuetype t;
  napi_typeof(env, *out, &t);
  return t != napi_undefined && t != napi_null;
}

double num(napi_env env, napi_value obj, const char* name, double fallback) {
  napi_value v;
  if (!get_prop(env, obj, name, &v)) return fallback;
  double d = fallback;
  napi_get_value_double(env, v, &d);
  return d;
}

// A typed array's data pointer, element count and element type.
bool typed(napi_env env, napi_value v, napi_typedarray_type want, void** data, size_t* length) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type type;
  napi_value buffer;
  size_t offset;
  napi_get_typedarray_info(env, v, &type, length, data, &buffer, &offset);
  return type == want;
}

bool typed_prop(napi_env env, napi_value obj, const char* name, napi_typedarray_type want,
                void** data, size_t* length) {
  napi_value v;
  if (!get_prop(env, obj, name, &v)) return false;
  return typed(env, v, want, data, length);
}

// A context handle is an External over this box; destroy() empties it so
// the device buffers go before garbage collection runs the finalizer. It
// also keeps the object count of the last setObjects: a frame's own object
// table must have exactly that many rows (the library copies that many).
struct Handle {
  gs_ctx* ctx;
  uint32_t object_count;
};

Handle* handle_arg(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    throw_status(env, GS_E_TYPE, "expected a context handle");
    return nullptr;
  }
  Handle* h = (Handle*)p;
  if (!h->ctx) {
    throw_status(env, GS_E_STATE, "the pass has been disposed");
    return nullptr;
  }
  return h;
}

gs_ctx* ctx_arg(napi_env env, napi_value v) {
  Handle* h = handle_arg(env, v);
  return h ? h->ctx : nullptr;
}

void finalize_ctx(napi_env, void* data, void*) {
  Handle* h = (Handle*)data;
  if (h->ctx) gs_destroy(h->ctx);
  delete h;
}

napi_value args(napi_env env, napi_callback_info info, size_t n, napi_value* argv) {
  size_t argc = n;
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  for (size_t i = argc; i < n; i++) argv[i] = undefined(env);
  return nullptr;
}

void set_num(napi_env env, napi_value obj, const char* name, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, name, x);
}

// gs_config from an options object of plain numbers (the facade maps the
// reference's option names to these).
void read_config(napi_env env, napi_value o, gs_config* c) {
  gs_config_default(c);
  napi_valuetype t;
  napi_typeof(env, o, &t);
  if (t != napi_object) return;
  c->device = (int32_t)num(env, o, "device", c->device);
  c->depth_sort_mode = (uint32_t)num(env, o, "depthSortMode", c->depth_sort_mode);
  c->antialias_mode = (uint32_t)num(env, o, "antialiasMode", c->antialias_mode);
  c->subpixel_sample_culling = (uint32_t)num(env, o, "subpixelSampleCulling", c->subpixel_sample_culling);
  c->raster_chunk_size = (uint32_t)num(env, o, "rasterChunkSize", c->raster_chunk_size);
  c->intersection_capacity = (uint32_t)num(env, o, "intersectionCapacity", c->intersection_capacity);
  c->max_intersection_capacity =
      (uint32_t)num(env, o, "maxIntersectionCapacity", c->max_intersection_capacity);
  c->output_depth = (uint32_t)num(env, o, "outputDepth", c->output_depth);
  c->transmittance_threshold = (float)num(env, o, "transmittanceThreshold", c->transmittance_threshold);
  c->depth_alpha_threshold = (float)num(env, o, "depthAlphaThreshold", c->depth_alpha_threshold);
  c->invalid_depth_value = (float)num(env, o, "invalidDepthValue", c->invalid_depth_value);
  c->raster_stats = (uint32_t)num(env, o, "rasterStats", c->raster_stats);
  c->capture_graph = (uint32_t)num(env, o, "captureGraph", c->capture_graph);
  c->tile_size = (uint32_t)num(env, o, "tileSize", c->tile_size);
  c->radix_bits = (uint32_t)num(env, o, "radixBits", c->radix_bits);
  c->raster_batch = (uint32_t)num(env, o, "rasterBatch", c->raster_batch);
  c->heavy_split_units = (uint32_t)num(env, o, "heavySplitUnits", c->heavy_split_units);
  c->heavy_fine_split_units =
      (uint32_t)num(env, o, "heavyFineSplitUnits", c->heavy_fine_split_units);
  c->view_group = (uint32_t)num(env, o, "viewGroup", c->view_group);
  c->heavy_pair_units = (uint32_t)num(env, o, "heavyPairUnits", c->heavy_pair_units);
  napi_value bg;
  if (get_prop(env, o, "background", &bg)) {
    for (uint32_t i = 0; i < 4; i++) {
      napi_value e;
      double d = 0.0;
      if (napi_get_element(env, bg, i, &e) == napi_ok) napi_get_value_double(env, e, &d);
      c->background[i] = (float)d;
    }
  }
}

napi_value config_default(napi_env env, napi_callback_info) {
  gs_config c;
  gs_config_default(&c);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "depthSortMode", c.depth_sort_mode);
  set_num(env, o, "antialiasMode", c.antialias_mode);
  set_num(env, o, "subpixelSampleCulling", c.subpixel_sample_culling);
  set_num(env, o, "rasterChunkSize", c.raster_chunk_size);
  set_num(env, o, "intersectionCapacity", c.intersection_capacity);
  set_num(env, o, "maxIntersectionCapacity", c.max_intersection_capacity);
  set_num(env, o, "transmittanceThreshold", c.transmittance_threshold);
  set_num(env, o, "depthAlphaThreshold", c.depth_alpha_threshold);
  set_num(env, o, "invalidDepthValue", c.invalid_depth_value);
  set_num(env, o, "tileSize", c.tile_size);
  set_num(env, o, "radixBits", c.radix_bits);
  set_num(env, o, "rasterBatch", c.raster_batch);
  return o;
}

napi_value create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_config c;
  read_config(env, argv[0], &c);
  gs_ctx* ctx = nullptr;
  int rc = gs_create(&c, &ctx);
  if (rc) {
    const char* msg = rc == GS_E_DEVICE ? "no usable HIP device" : "invalid configuration";
    return throw_status(env, rc, msg);
  }
  napi_value ext;
  napi_create_external(env, new Handle{ctx, 0u}, finalize_ctx, nullptr, &ext);
  return ext;
}

napi_value set_data(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  gs_data_desc d;
  memset(&d, 0, sizeof(d));
  d.count = (uint32_t)num(env, argv[1], "count", 0);
  d.active_count = (uint32_t)num(env, argv[1], "activeCount", d.count);
  d.sh_degree = (uint32_t)num(env, argv[1], "shDegree", 3);
  d.memory = GS_MEMORY_HOST;
  void* p;
  size_t n;
  size_t ncoef = (size_t)(d.sh_degree + 1) * (d.sh_degree + 1);
  if (!typed_prop(env, argv[1], "means", napi_float32_array, &p, &n) || n < (size_t)d.count * 4)
    return throw_status(env, GS_E_TYPE, "means must be a Float32Array of count * 4");
  d.means = (const float*)p;
  if (!typed_prop(env, argv[1], "scalesOpacity", napi_uint32_array, &p, &n) ||
      n < (size_t)d.count * 2)
    return throw_status(env, GS_E_TYPE, "scalesOpacity must be a Uint32Array of count * 2");
  d.scales_opacity = (const uint32_t*)p;
  if (!typed_prop(env, argv[1], "rotations", napi_uint32_array, &p, &n) || n < d.count)
    return throw_status(env, GS_E_TYPE, "rotations must be a Uint32Array of count");
  d.rotations = (const uint32_t*)p;
  if (!typed_prop(env, argv[1], "shCoefficients", napi_uint32_array, &p, &n) ||
      n < d.count * ncoef)
    return throw_status(env, GS_E_TYPE, "shCoefficients must be a Uint32Array of count * (deg+1)^2");
  d.sh = (const uint32_t*)p;
  if (typed_prop(env, argv[1], "activeSlots", napi_uint32_array, &p, &n)) {
    if (n < d.active_count) return throw_status(env, GS_E_RANGE, "activeSlots is too short");
    d.active_slots = (const uint32_t*)p;
  }
  return check(env, ctx, gs_set_data(ctx, &d));
}

napi_value set_objects(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  Handle* h = handle_arg(env, argv[0]);
  if (!h) return nullptr;
  void* p;
  size_t n;
  if (!typed(env, argv[1], napi_float32_array, &p, &n) || n % 40)
    return throw_status(env, GS_E_TYPE, "object frames must be a Float32Array of 40 per object");
  int rc = gs_set_objects(h->ctx, (uint32_t)(n / 40), (const float*)p);
  if (rc == GS_OK) h->object_count = (uint32_t)(n / 40);
  return check(env, h->ctx, rc);
}

napi_value object_frame(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  void *cam, *obj;
  size_t n1, n2;
  if (!typed(env, argv[0], napi_float64_array, &cam, &n1) || n1 != 16 ||
      !typed(env, argv[1], napi_float64_array, &obj, &n2) || n2 != 16)
    return throw_status(env, GS_E_TYPE, "camera and object matrices must be Float64Array(16)");
  bool visible = true;
  napi_get_value_bool(env, argv[2], &visible);
  napi_value buffer, out;
  void* data;
  napi_create_arraybuffer(env, 40 * sizeof(float), &data, &buffer);
  gs_object_frame((const double*)cam, (const double*)obj, visible ? 1 : 0, (float*)data);
  napi_create_typedarray(env, napi_float32_array, 40, buffer, 0, &out);
  return out;
}

// A frame object; `objects` receives a copy of its own object table, which
// must hold object_count rows of 40 floats (the count of setObjects).
bool read_frame(napi_env env, napi_value o, gs_frame* f, std::vector<float>* objects,
                uint32_t object_count) {
  memset(f, 0, sizeof(*f));
  void* p;
  size_t n;
  if (!typed_prop(env, o, "projection", napi_float32_array, &p, &n) || n != 16) {
    throw_status(env, GS_E_TYPE, "projection must be a Float32Array(16)");
    return false;
  }
  memcpy(f->projection, p, sizeof(f->projection));
  f->width = (uint32_t)num(env, o, "width", 0);
  f->height = (uint32_t)num(env, o, "height", 0);
  f->near_plane = (float)num(env, o, "near", 0.1);
  f->far_plane = (float)num(env, o, "far", 1000.0);
  f->filter_variance = (float)num(env, o, "filterVariance", 0.3);
  f->boundary_axis = (uint32_t)num(env, o, "boundaryAxis", 0);
  f->attempt = (uint32_t)num(env, o, "attempt", 1);
  f->reset_valid = (uint32_t)num(env, o, "resetValid", 0);
  if (objects && typed_prop(env, o, "objectFrames", napi_float32_array, &p, &n)) {
    if (n != (size_t)object_count * 40) {
      throw_status(env, GS_E_RANGE,
                   "objectFrames must hold 40 floats for each object of setObjects");
      return false;
    }
    objects->assign((const float*)p, (const float*)p + n);
    f->object_frames = objects->data();
  }
  return true;
}

napi_value render(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  Handle* h = handle_arg(env, argv[0]);
  if (!h) return nullptr;
  gs_frame f;
  std::vector<float> objects;
  if (!read_frame(env, argv[1], &f, &objects, h->object_count)) return nullptr;
  return check(env, h->ctx, gs_render(h->ctx, &f, nullptr));
}

// The gs_frame array of a JS array of frame objects (object tables copied).
bool read_frames(napi_env env, napi_value arr, uint32_t object_count,
                 std::vector<gs_frame>* frames, std::vector<std::vector<float>>* objects) {
  uint32_t n = 0;
  napi_get_array_length(env, arr, &n);
  frames->resize(n);
  objects->resize(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    napi_get_element(env, arr, i, &e);
    if (!read_frame(env, e, &(*frames)[i], &(*objects)[i], object_count)) return false;
  }
  return true;
}

napi_value render_batch(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  Handle* h = handle_arg(env, argv[0]);
  if (!h) return nullptr;
  std::vector<gs_frame> frames;
  std::vector<std::vector<float>> objects;
  if (!read_frames(env, argv[1], h->object_count, &frames, &objects)) return nullptr;
  return check(env, h->ctx,
               gs_render_batch(h->ctx, (uint32_t)frames.size(), frames.data(), nullptr));
}

napi_value state_object(napi_env env, const gs_state& s) {
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "attempt", s.attempt);
  set_num(env, o, "version", s.version);
  set_num(env, o, "requested", s.requested);
  set_num(env, o, "overflow", s.overflow);
  set_num(env, o, "visible", s.visible);
  set_num(env, o, "intersections", s.intersections);
  set_num(env, o, "capacity", s.capacity);
  set_num(env, o, "tilesX", s.tiles_x);
  set_num(env, o, "tilesY", s.tiles_y);
  set_num(env, o, "rasterChecked", (double)s.raster_checked);
  set_num(env, o, "rasterBlended", (double)s.raster_blended);
  set_num(env, o, "rasterCheckedHeavy", (double)s.raster_checked_heavy);
  set_num(env, o, "errorFlags", s.error_flags);
  return o;
}

napi_value read_state(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  gs_state s;
  int rc = gs_read_state(ctx, &s);
  if (rc) return check(env, ctx, rc);
  return state_object(env, s);
}

napi_value read_batch_state(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  uint32_t n = 0;
  napi_get_value_uint32(env, argv[1], &n);
  std::vector<gs_state> s(n);
  int rc = gs_read_batch_state(ctx, 0, n, s.data());
  if (rc) return check(env, ctx, rc);
  napi_value arr;
  napi_create_array_with_length(env, n, &arr);
  for (uint32_t i = 0; i < n; i++) napi_set_element(env, arr, i, state_object(env, s[i]));
  return arr;
}

napi_value resize_intersections(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  double cap = 0;
  napi_get_value_double(env, argv[1], &cap);
  if (!(cap >= 1 && cap <= 4294967295.0))
    return throw_status(env, GS_E_RANGE, "capacity must be a positive integer");
  return check(env, ctx, gs_resize_intersections(ctx, (uint32_t)cap));
}

napi_value resize_batch_intersections(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  uint32_t view = 0, cap = 0;
  napi_get_value_uint32(env, argv[1], &view);
  napi_get_value_uint32(env, argv[2], &cap);
  return check(env, ctx, gs_resize_batch_intersections(ctx, view, cap));
}

// readImage(ctx, color?: Uint16Array, depth?: Float32Array, view?: number)
napi_value read_image(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  args(env, info, 4, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  void *color = nullptr, *depth = nullptr;
  size_t nc = 0, nd = 0;
  if (!typed(env, argv[1], napi_uint16_array, &color, &nc)) color = nullptr, nc = 0;
  if (!typed(env, argv[2], napi_float32_array, &depth, &nd)) depth = nullptr, nd = 0;
  napi_valuetype t;
  napi_typeof(env, argv[3], &t);
  if (t == napi_number) {
    uint32_t view = 0;
    napi_get_value_uint32(env, argv[3], &view);
    return check(env, ctx, gs_read_batch_image(ctx, view, color, nc * 2, depth, nd * 4));
  }
  return check(env, ctx, gs_read_image(ctx, color, nc * 2, depth, nd * 4));
}

napi_value read_debug(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  char name[64] = {0};
  size_t len = 0;
  napi_get_value_string_utf8(env, argv[1], name, sizeof(name), &len);
  size_t bytes = 0;
  int rc = gs_read_debug(ctx, name, nullptr, 0, &bytes);
  if (rc) return check(env, ctx, rc);
  napi_value buffer;
  void* data = nullptr;
  napi_create_arraybuffer(env, bytes, &data, &buffer);
  if (bytes && (rc = gs_read_debug(ctx, name, data, bytes, nullptr))) return check(env, ctx, rc);
  return buffer;
}

napi_value capacity_next(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  args(env, info, 5, argv);
  double v[5] = {0, 0, 0, 2.0, 1.25};
  for (int i = 0; i < 5; i++) {
    napi_valuetype t;
    napi_typeof(env, argv[i], &t);
    if (t == napi_number) napi_get_value_double(env, argv[i], &v[i]);
  }
  uint32_t next = gs_capacity_next((uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2], v[3], v[4]);
  if (next == 0) {
    napi_value n;
    napi_get_null(env, &n);
    return n;
  }
  napi_value r;
  napi_create_uint32(env, next, &r);
  return r;
}

napi_value set_tuning(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  gs_tuning t = {(uint32_t)num(env, argv[1], "tileSize", 16), (uint32_t)num(env, argv[1], "radixBits", 8),
                 (uint32_t)num(env, argv[1], "rasterBatch", 256)};
  return check(env, ctx, gs_set_tuning(ctx, &t));
}

// A new Float32Array of n elements; *data receives its storage.
napi_value new_f32(napi_env env, size_t n, float** data) {
  napi_value buffer, out;
  void* p = nullptr;
  napi_create_arraybuffer(env, n * sizeof(float), &p, &buffer);
  napi_create_typedarray(env, napi_float32_array, n, buffer, 0, &out);
  *data = (float*)p;
  return out;
}

double arg_num(napi_env env, napi_value v, double fallback) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  double d = fallback;
  if (t == napi_number) napi_get_value_double(env, v, &d);
  return d;
}

// updateActive(ctx, start, slots: Uint32Array, newActiveCount)
napi_value update_active(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  args(env, info, 4, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  void* p = nullptr;
  size_t n = 0;
  if (!typed(env, argv[2], napi_uint32_array, &p, &n))
    return throw_status(env, GS_E_TYPE, "slots must be a Uint32Array");
  double start = arg_num(env, argv[1], -1), active = arg_num(env, argv[3], -1);
  if (!(start >= 0 && start <= 4294967295.0 && active >= 0 && active <= 4294967295.0))
    return throw_status(env, GS_E_RANGE, "start and activeCount must be non-negative integers");
  return check(env, ctx, gs_update_active(ctx, (uint32_t)start, (uint32_t)n, (const uint32_t*)p,
                                          (uint32_t)active));
}

napi_value get_tuning(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  gs_tuning t;
  int rc = gs_get_tuning(ctx, &t);
  if (rc) return check(env, ctx, rc);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "tileSize", t.tile_size);
  set_num(env, o, "radixBits", t.radix_bits);
  set_num(env, o, "rasterBatch", t.raster_batch);
  return o;
}

// autotune(ctx, frames, reps) -> {best: tuning, table: [{tileSize, radixBits, rasterBatch, ms}]}
napi_value autotune(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  Handle* h = handle_arg(env, argv[0]);
  if (!h) return nullptr;
  std::vector<gs_frame> frames;
  std::vector<std::vector<float>> objects;
  if (!read_frames(env, argv[1], h->object_count, &frames, &objects)) return nullptr;
  double reps = arg_num(env, argv[2], 3);
  if (!(reps >= 1 && reps <= 1e6)) return throw_status(env, GS_E_RANGE, "reps must be positive");
  gs_tuning best;
  gs_tune_entry table[18];
  uint32_t count = 0;
  int rc = gs_autotune(h->ctx, (uint32_t)frames.size(), frames.data(), (uint32_t)reps, &best,
                       table, 18, &count);
  if (rc) return check(env, h->ctx, rc);
  napi_value o, b, arr;
  napi_create_object(env, &o);
  napi_create_object(env, &b);
  set_num(env, b, "tileSize", best.tile_size);
  set_num(env, b, "radixBits", best.radix_bits);
  set_num(env, b, "rasterBatch", best.raster_batch);
  napi_set_named_property(env, o, "best", b);
  napi_create_array_with_length(env, count, &arr);
  for (uint32_t i = 0; i < count && i < 18; i++) {
    napi_value e;
    napi_create_object(env, &e);
    set_num(env, e, "tileSize", table[i].tuning.tile_size);
    set_num(env, e, "radixBits", table[i].tuning.radix_bits);
    set_num(env, e, "rasterBatch", table[i].tuning.raster_batch);
    set_num(env, e, "ms", table[i].ms);
    napi_set_element(env, arr, i, e);
  }
  napi_set_named_property(env, o, "table", arr);
  return o;
}

// ---- learned cull (gs_cull_*) ----

napi_value cull_init(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  double seed = arg_num(env, argv[1], 0);
  if (!(seed >= 0 && seed <= 9007199254740991.0))
    return throw_status(env, GS_E_RANGE, "seed must be a non-negative safe integer");
  return check(env, ctx, gs_cull_init(ctx, (uint64_t)seed));
}

napi_value cull_set_params(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  void* p = nullptr;
  size_t n = 0;
  if (!typed(env, argv[1], napi_float32_array, &p, &n) || n != GS_CULL_PARAMS)
    return throw_status(env, GS_E_TYPE, "params must be a Float32Array(5313)");
  return check(env, ctx, gs_cull_set_params(ctx, (const float*)p));
}

napi_value cull_get_params(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  float* data = nullptr;
  napi_value out = new_f32(env, GS_CULL_PARAMS, &data);
  int rc = gs_cull_get_params(ctx, data);
  if (rc) return check(env, ctx, rc);
  return out;
}

// Training checkpoint blob (gs_cull_save / gs_cull_load) as a Uint8Array.
napi_value cull_save(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  napi_value buffer, out;
  void* p = nullptr;
  napi_create_arraybuffer(env, GS_CULL_CHECKPOINT_BYTES, &p, &buffer);
  napi_create_typedarray(env, napi_uint8_array, GS_CULL_CHECKPOINT_BYTES, buffer, 0, &out);
  int rc = gs_cull_save(ctx, p, GS_CULL_CHECKPOINT_BYTES);
  if (rc) return check(env, ctx, rc);
  return out;
}

napi_value cull_load(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  void* p = nullptr;
  size_t n = 0;
  if (!typed(env, argv[1], napi_uint8_array, &p, &n))
    return throw_status(env, GS_E_TYPE, "a checkpoint must be a Uint8Array");
  return check(env, ctx, gs_cull_load(ctx, p, n));
}

napi_value cull_set_normalization(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  void *mu = nullptr, *inv = nullptr;
  size_t n1 = 0, n2 = 0;
  if (!typed(env, argv[1], napi_float32_array, &mu, &n1) || n1 != GS_CULL_FEATURES ||
      !typed(env, argv[2], napi_float32_array, &inv, &n2) || n2 != GS_CULL_FEATURES)
    return throw_status(env, GS_E_TYPE, "mu and invSigma must be Float32Array(16)");
  return check(env, ctx, gs_cull_set_normalization(ctx, (const float*)mu, (const float*)inv));
}

napi_value cull_get_normalization(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  float *mu = nullptr, *inv = nullptr;
  napi_value a = new_f32(env, GS_CULL_FEATURES, &mu), b = new_f32(env, GS_CULL_FEATURES, &inv);
  int rc = gs_cull_get_normalization(ctx, mu, inv);
  if (rc) return check(env, ctx, rc);
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "mu", a);
  napi_set_named_property(env, o, "invSigma", b);
  return o;
}

napi_value cull_enable(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  bool on = false;
  napi_get_value_bool(env, argv[1], &on);
  return check(env, ctx, gs_cull_enable(ctx, on ? 1 : 0, (float)arg_num(env, argv[2], 1e-4)));
}

napi_value cull_reserve(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  double cap = arg_num(env, argv[1], -1);
  if (!(cap >= 0 && cap <= 4294967295.0))
    return throw_status(env, GS_E_RANGE, "capacity must be a non-negative integer");
  return check(env, ctx, gs_cull_reserve(ctx, (uint32_t)cap));
}

// cullCollect(ctx, frame, rate, seed) -> state
napi_value cull_collect(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  args(env, info, 4, argv);
  Handle* h = handle_arg(env, argv[0]);
  if (!h) return nullptr;
  gs_frame f;
  std::vector<float> objects;
  if (!read_frame(env, argv[1], &f, &objects, h->object_count)) return nullptr;
  gs_state st;
  int rc = gs_cull_collect(h->ctx, &f, (float)arg_num(env, argv[2], 0.05),
                           (uint32_t)arg_num(env, argv[3], 0), &st);
  if (rc) return check(env, h->ctx, rc);
  return state_object(env, st);
}

napi_value cull_dataset_size(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  uint32_t n = 0;
  int rc = gs_cull_dataset_size(ctx, &n);
  if (rc) return check(env, ctx, rc);
  napi_value r;
  napi_create_uint32(env, n, &r);
  return r;
}

// cullReadDataset(ctx, first, n) -> {features: Float32Array(n * 16), targets: Float32Array(n)}
napi_value cull_read_dataset(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  args(env, info, 3, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  double first = arg_num(env, argv[1], -1), n = arg_num(env, argv[2], -1);
  if (!(first >= 0 && n >= 0 && first + n <= 4294967295.0))
    return throw_status(env, GS_E_RANGE, "first and n must be non-negative integers");
  float *f = nullptr, *t = nullptr;
  napi_value a = new_f32(env, (size_t)n * GS_CULL_FEATURES, &f), b = new_f32(env, (size_t)n, &t);
  int rc = gs_cull_read_dataset(ctx, (uint32_t)first, (uint32_t)n, f, t);
  if (rc) return check(env, ctx, rc);
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "features", a);
  napi_set_named_property(env, o, "targets", b);
  return o;
}

napi_value cull_fit_normalization(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  return check(env, ctx, gs_cull_fit_normalization(ctx));
}

// cullTrain(ctx, steps, batch, lr, underWeight, seed) -> last loss
napi_value cull_train(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  args(env, info, 6, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  double steps = arg_num(env, argv[1], -1), batch = arg_num(env, argv[2], 65536);
  double seed = arg_num(env, argv[5], 0);
  if (!(steps >= 0 && steps <= 4294967295.0 && batch >= 1 && batch <= 4294967295.0 &&
        seed >= 0 && seed <= 9007199254740991.0))
    return throw_status(env, GS_E_RANGE, "steps, batch and seed must be non-negative integers");
  float loss = 0.0f;
  int rc = gs_cull_train(ctx, (uint32_t)steps, (uint32_t)batch, (float)arg_num(env, argv[3], 2e-3),
                         (float)arg_num(env, argv[4], 4.0), (uint64_t)seed, &loss);
  if (rc) return check(env, ctx, rc);
  napi_value r;
  napi_create_double(env, loss, &r);
  return r;
}

// cullEval(ctx, features: Float32Array(n * 16)) -> Float32Array(n)
napi_value cull_eval(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  void* p = nullptr;
  size_t n = 0;
  if (!typed(env, argv[1], napi_float32_array, &p, &n) || n % GS_CULL_FEATURES)
    return throw_status(env, GS_E_TYPE, "features must be a Float32Array of 16 per sample");
  float* out = nullptr;
  napi_value r = new_f32(env, n / GS_CULL_FEATURES, &out);
  if (n == 0) return r;
  int rc = gs_cull_eval(ctx, (uint32_t)(n / GS_CULL_FEATURES), (const float*)p, out);
  if (rc) return check(env, ctx, rc);
  return r;
}

// cullGradient(ctx, features, targets, underWeight) -> {grad: Float32Array(5313), loss}
napi_value cull_gradient(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  args(env, info, 4, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  void *f = nullptr, *t = nullptr;
  size_t nf = 0, nt = 0;
  if (!typed(env, argv[1], napi_float32_array, &f, &nf) || nf % GS_CULL_FEATURES ||
      !typed(env, argv[2], napi_float32_array, &t, &nt) || nt != nf / GS_CULL_FEATURES)
    return throw_status(env, GS_E_TYPE,
                        "features (16 per sample) and targets (1 per sample) must be Float32Arrays");
  float* g = nullptr;
  napi_value grad = new_f32(env, GS_CULL_PARAMS, &g);
  float loss = 0.0f;
  int rc = gs_cull_gradient(ctx, (uint32_t)nt, (const float*)f, (const float*)t,
                            (float)arg_num(env, argv[3], 4.0), g, &loss);
  if (rc) return check(env, ctx, rc);
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "grad", grad);
  set_num(env, o, "loss", loss);
  return o;
}

napi_value last_error(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  napi_value s;
  napi_create_string_utf8(env, gs_last_error(ctx), NAPI_AUTO_LENGTH, &s);
  return s;
}

napi_value destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  args(env, info, 1, argv);
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) != napi_ok || !p)
    return throw_status(env, GS_E_TYPE, "expected a context handle");
  Handle* h = (Handle*)p;
  if (h->ctx) gs_destroy(h->ctx);
  h->ctx = nullptr;
  return undefined(env);
}

}  // namespace

NAPI_MODULE_INIT() {
  napi_property_descriptor p[] = {
      {"configDefault", nullptr, config_default, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"create", nullptr, create, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"destroy", nullptr, destroy, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setData", nullptr, set_data, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setObjects", nullptr, set_objects, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"objectFrame", nullptr, object_frame, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"render", nullptr, render, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"renderBatch", nullptr, render_batch, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readState", nullptr, read_state, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readBatchState", nullptr, read_batch_state, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"resizeIntersections", nullptr, resize_intersections, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"resizeBatchIntersections", nullptr, resize_batch_intersections, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readImage", nullptr, read_image, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readDebug", nullptr, read_debug, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"capacityNext", nullptr, capacity_next, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setTuning", nullptr, set_tuning, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"lastError", nullptr, last_error, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"updateActive", nullptr, update_active, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"getTuning", nullptr, get_tuning, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"autotune", nullptr, autotune, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullInit", nullptr, cull_init, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullSetParams", nullptr, cull_set_params, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullGetParams", nullptr, cull_get_params, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullSave", nullptr, cull_save, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullLoad", nullptr, cull_load, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullSetNormalization", nullptr, cull_set_normalization, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullGetNormalization", nullptr, cull_get_normalization, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullEnable", nullptr, cull_enable, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullReserve", nullptr, cull_reserve, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullCollect", nullptr, cull_collect, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullDatasetSize", nullptr, cull_dataset_size, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullReadDataset", nullptr, cull_read_dataset, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullFitNormalization", nullptr, cull_fit_normalization, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullTrain", nullptr, cull_train, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullEval", nullptr, cull_eval, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cullGradient", nullptr, cull_gradient, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
  };
  napi_define_properties(env, exports, sizeof p / sizeof *p, p);
  return exports;
}
