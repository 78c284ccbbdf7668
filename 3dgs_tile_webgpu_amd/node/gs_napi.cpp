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
// the device buffers go before garbage collection runs the finalizer.
struct Handle {
  gs_ctx* ctx;
};

gs_ctx* ctx_arg(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    throw_status(env, GS_E_TYPE, "expected a context handle");
    return nullptr;
  }
  gs_ctx* ctx = ((Handle*)p)->ctx;
  if (!ctx) throw_status(env, GS_E_STATE, "the pass has been disposed");
  return ctx;
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
  napi_create_external(env, new Handle{ctx}, finalize_ctx, nullptr, &ext);
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
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  void* p;
  size_t n;
  if (!typed(env, argv[1], napi_float32_array, &p, &n) || n % 40)
    return throw_status(env, GS_E_TYPE, "object frames must be a Float32Array of 40 per object");
  return check(env, ctx, gs_set_objects(ctx, (uint32_t)(n / 40), (const float*)p));
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

bool read_frame(napi_env env, napi_value o, gs_frame* f, std::vector<float>* objects) {
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
    objects->assign((const float*)p, (const float*)p + n);
    f->object_frames = objects->data();
  }
  return true;
}

napi_value render(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  gs_frame f;
  std::vector<float> objects;
  if (!read_frame(env, argv[1], &f, &objects)) return nullptr;
  return check(env, ctx, gs_render(ctx, &f, nullptr));
}

napi_value render_batch(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  args(env, info, 2, argv);
  gs_ctx* ctx = ctx_arg(env, argv[0]);
  if (!ctx) return nullptr;
  uint32_t n = 0;
  napi_get_array_length(env, argv[1], &n);
  std::vector<gs_frame> frames(n);
  std::vector<std::vector<float>> objects(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    napi_get_element(env, argv[1], i, &e);
    if (!read_frame(env, e, &frames[i], &objects[i])) return nullptr;
  }
  return check(env, ctx, gs_render_batch(ctx, n, frames.data(), nullptr));
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
  };
  napi_define_properties(env, exports, sizeof p / sizeof *p, p);
  return exports;
}
