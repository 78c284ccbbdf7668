// gs_producer.cpp - compact/rgb8e8 encoder and the seeded scene generator.
//
// Host code (no device work). The encoder follows the reference's Rust
// producer operation by operation in binary32; see include/gs_producer.h.
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gs_producer.h"
#include "gs_math.h"

extern "C" uint16_t gs_f32_to_f16(float v) { return (uint16_t)gs_float_to_half(v); }

// splat_encode.rs:234-252 (encode_quat_oct101012); Rust f32 semantics:
// round() is half away from zero, clamp before round.
extern "C" uint32_t gs_encode_quat_oct101012(const float q_in[4]) {
  float q[4] = {q_in[0], q_in[1], q_in[2], q_in[3]};
  if (q[3] < 0.0f)
    for (float& v : q) v = -v;
  float w = q[3] < 0.0f ? 0.0f : (q[3] > 1.0f ? 1.0f : q[3]);
  float theta = 2.0f * acosf(w);
  float s = sinf(theta * 0.5f);
  float axis[3];
  if (fabsf(s) < 1e-6f) {
    axis[0] = 1.0f;
    axis[1] = 0.0f;
    axis[2] = 0.0f;
  } else {
    for (int i = 0; i < 3; i++) axis[i] = q[i] / s;
  }
  float sum = fabsf(axis[0]) + fabsf(axis[1]) + fabsf(axis[2]);
  float p[2] = {axis[0] / sum, axis[1] / sum};
  if (axis[2] < 0.0f) {
    float p0 = (1.0f - fabsf(p[1])) * (p[0] >= 0.0f ? 1.0f : -1.0f);
    float p1 = (1.0f - fabsf(p[0])) * (p[1] >= 0.0f ? 1.0f : -1.0f);
    p[0] = p0;
    p[1] = p1;
  }
  uint32_t uv[2];
  for (int i = 0; i < 2; i++) {
    float x = (p[i] * 0.5f + 0.5f) * 1023.0f;
    x = x < 0.0f ? 0.0f : (x > 1023.0f ? 1023.0f : x);
    uv[i] = (uint32_t)roundf(x);
  }
  float r = theta / 3.14159265358979323846f * 4095.0f;
  r = r < 0.0f ? 0.0f : (r > 4095.0f ? 4095.0f : r);
  uint32_t ru = (uint32_t)roundf(r);
  return (ru << 20) | (uv[1] << 10) | uv[0];
}

// engine.rs:1801-1832 with model.rs:225-228 and splat_encode.rs:127-136.
extern "C" void gs_encode_compact(size_t n, const float* scales, const float* opacity,
                                  const float* quats, uint32_t* out_shape,
                                  uint32_t* out_rotation) {
  for (size_t i = 0; i < n; i++) {
    float a = opacity[i];
    // render_scale_opacity: scales * cbrt(max(opacity, 1)), opacity.min(1).
    float inflate = cbrtf(a > 1.0f ? a : 1.0f);
    float o = a < 1.0f ? a : 1.0f;
    float s[3];
    for (int k = 0; k < 3; k++) s[k] = scales[i * 3 + k] * inflate;
    // Quat::normalize (glam): q * (1 / length).
    const float* q = quats + i * 4;
    float len = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    float inv = 1.0f / len;
    float qn[4] = {q[0] * inv, q[1] * inv, q[2] * inv, q[3] * inv};
    uint32_t ln0 = gs_float_to_half(logf(s[0]));
    uint32_t ln1 = gs_float_to_half(logf(s[1]));
    uint32_t ln2 = gs_float_to_half(logf(s[2]));
    uint32_t oh = gs_float_to_half(o);
    out_shape[i * 2] = ln0 | (ln1 << 16);
    out_shape[i * 2 + 1] = ln2 | (oh << 16);
    out_rotation[i] = gs_encode_quat_oct101012(qn);
  }
}

// engine.rs:1940-1949 (pack_sh).
extern "C" uint32_t gs_pack_sh_rgb8e8(const float rgb[3]) {
  float m = 0.0f;
  for (int i = 0; i < 3; i++) m = std::max(m, fabsf(rgb[i]));  // f32::max fold
  if (m == 0.0f) return 0u;
  float e = ceilf(log2f(m));
  e = e < -126.0f ? -126.0f : (e > 127.0f ? 127.0f : e);
  int exponent = (int)e;
  float multiplier = 127.0f / gs_exp2i(exponent);
  uint32_t b[3];
  for (int i = 0; i < 3; i++) {
    float v = floorf(rgb[i] * multiplier + 0.5f);
    v = v < -127.0f ? -127.0f : (v > 127.0f ? 127.0f : v);
    b[i] = (uint32_t)(int32_t)v & 255u;
  }
  return b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)(exponent + 127) << 24);
}

extern "C" void gs_encode_sh_rgb8e8(size_t n, const float* rgb, uint32_t* out) {
  for (size_t i = 0; i < n; i++) out[i] = gs_pack_sh_rgb8e8(rgb + i * 3);
}

// ------------------------------------------------------------------ PLY ----
//
// Canonical 3DGS PLY files (CanonicalGaussianPlyLoader.ts:84-350): an ascii,
// binary_little_endian or binary_big_endian header whose first non-empty
// element is `vertex`, with scalar properties x, y, z, scale_0..2,
// rot_0..3 (w, x, y, z), opacity (logit), f_dc_0..2 and f_rest_0..(3K - 1)
// laid out channel-major (K higher-band coefficients per channel).

namespace {

struct PlyProp {
  std::string name;
  int type;  // byte size, negative for signed integers, 40 / 80 for float / double
  size_t offset;
};

thread_local char g_ply_error[160];

int ply_fail(const char* msg) {
  strncpy(g_ply_error, msg, sizeof(g_ply_error) - 1);
  g_ply_error[sizeof(g_ply_error) - 1] = 0;
  return -1;
}

// Scalar type name -> (size, kind): kind 0 unsigned, 1 signed, 2 float.
bool ply_type(const std::string& t, int& size, int& kind) {
  static const struct { const char* name; int size, kind; } types[] = {
      {"char", 1, 1},   {"int8", 1, 1},    {"uchar", 1, 0},   {"uint8", 1, 0},
      {"short", 2, 1},  {"int16", 2, 1},   {"ushort", 2, 0},  {"uint16", 2, 0},
      {"int", 4, 1},    {"int32", 4, 1},   {"uint", 4, 0},    {"uint32", 4, 0},
      {"float", 4, 2},  {"float32", 4, 2}, {"double", 8, 2},  {"float64", 8, 2}};
  for (const auto& e : types)
    if (t == e.name) {
      size = e.size;
      kind = e.kind;
      return true;
    }
  return false;
}

struct PlyHeader {
  int format = -1;  // 0 ascii, 1 little endian, 2 big endian
  uint64_t count = 0;
  size_t stride = 0, data = 0;
  std::vector<std::string> names;
  std::vector<int> sizes, kinds;
  std::vector<size_t> offsets;
};

int ply_header(const unsigned char* b, size_t n, PlyHeader& h) {
  static const char kEnd[] = "end_header";
  size_t mark = n;
  for (size_t i = 0; i + 10 <= n; i++)
    if (memcmp(b + i, kEnd, 10) == 0) {
      mark = i;
      break;
    }
  if (mark == n) return ply_fail("Invalid PLY: end_header is missing");
  size_t data = mark + 10;
  if (data < n && b[data] == 13) data++;
  if (data >= n || b[data] != 10) return ply_fail("Invalid PLY: end_header must terminate a line");
  data++;
  std::string text((const char*)b, data);
  std::vector<std::string> lines;
  size_t at = 0;
  while (at < text.size()) {
    size_t e = text.find('\n', at);
    if (e == std::string::npos) e = text.size();
    std::string line = text.substr(at, e - at);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    lines.push_back(line);
    at = e + 1;
  }
  auto tokens = [](const std::string& l) {
    std::vector<std::string> t;
    size_t i = 0;
    while (i < l.size()) {
      while (i < l.size() && isspace((unsigned char)l[i])) i++;
      size_t j = i;
      while (j < l.size() && !isspace((unsigned char)l[j])) j++;
      if (j > i) t.push_back(l.substr(i, j - i));
      i = j;
    }
    return t;
  };
  if (lines.empty() || tokens(lines[0]) != std::vector<std::string>{"ply"})
    return ply_fail("Invalid PLY signature");
  std::string element;
  bool vertex_seen = false, first_nonempty_checked = false;
  int64_t vertex_count = -1;
  for (const std::string& line : lines) {
    std::vector<std::string> t = tokens(line);
    if (t.empty()) continue;
    if (t[0] == "format") {
      if (t.size() < 2) return ply_fail("Unsupported PLY format: unknown");
      if (t[1] == "ascii") h.format = 0;
      else if (t[1] == "binary_little_endian") h.format = 1;
      else if (t[1] == "binary_big_endian") h.format = 2;
      else return ply_fail("Unsupported PLY format");
    } else if (t[0] == "element") {
      element = t.size() > 1 ? t[1] : "";
      char* end = nullptr;
      long long c = t.size() > 2 ? strtoll(t[2].c_str(), &end, 10) : -1;
      if (t.size() < 3 || c < 0 || !end || *end) return ply_fail("Invalid element count");
      if (element == "vertex") {
        vertex_count = c;
        vertex_seen = true;
      }
      if (c > 0 && !first_nonempty_checked) {
        first_nonempty_checked = true;
        if (element != "vertex")
          return ply_fail("The canonical 3DGS vertex element must be first");
      }
    } else if (t[0] == "property" && element == "vertex") {
      if (t.size() > 1 && t[1] == "list")
        return ply_fail("List properties are not supported in the vertex element");
      int size = 0, kind = 0;
      if (t.size() < 3 || !ply_type(t[1], size, kind))
        return ply_fail("Unsupported vertex property");
      h.names.push_back(t[2]);
      h.sizes.push_back(size);
      h.kinds.push_back(kind);
      h.offsets.push_back(h.stride);
      h.stride += (size_t)size;
    }
  }
  if (h.format < 0) return ply_fail("Invalid PLY: format is missing");
  if (!vertex_seen || vertex_count <= 0) return ply_fail("PLY must contain at least one vertex");
  if (vertex_count > 0xffffffffll) return ply_fail("PLY vertex count exceeds 2^32 - 1");
  h.count = (uint64_t)vertex_count;
  h.data = data;
  return 0;
}

double ply_scalar(const unsigned char* p, int size, int kind, bool little) {
  unsigned char tmp[8];
  for (int i = 0; i < size; i++) tmp[i] = little ? p[i] : p[size - 1 - i];
  if (kind == 2) {
    if (size == 4) {
      float f;
      memcpy(&f, tmp, 4);
      return f;
    }
    double d;
    memcpy(&d, tmp, 8);
    return d;
  }
  uint64_t u = 0;
  memcpy(&u, tmp, (size_t)size);
  if (kind == 1) {
    int64_t v = (int64_t)(u << (64 - 8 * size)) >> (64 - 8 * size);  // sign-extend
    return (double)v;
  }
  return (double)u;
}

}  // namespace

extern "C" const char* gs_ply_last_error(void) { return g_ply_error; }

extern "C" int gs_ply_parse(const void* data, size_t size, gs_ply_info* info, float* means,
                            float* scales, float* opacity, float* quats, float* sh) {
  g_ply_error[0] = 0;
  if (!data || !info) return ply_fail("missing PLY data");
  const unsigned char* b = (const unsigned char*)data;
  PlyHeader h;
  if (ply_header(b, size, h)) return -1;
  auto index = [&](const std::string& name) -> int {
    for (size_t i = 0; i < h.names.size(); i++)
      if (h.names[i] == name) return (int)i;
    return -1;
  };
  static const char* kRequired[] = {"x",     "y",     "z",     "scale_0", "scale_1",
                                    "scale_2", "rot_0", "rot_1", "rot_2",   "rot_3",
                                    "opacity", "f_dc_0", "f_dc_1", "f_dc_2"};
  int req[14];
  for (int i = 0; i < 14; i++)
    if ((req[i] = index(kRequired[i])) < 0) {
      snprintf(g_ply_error, sizeof(g_ply_error),
               "Not a canonical 3DGS PLY: missing property %s", kRequired[i]);
      return -1;
    }
  // f_rest_k must be exactly 0..R-1 with R divisible by three.
  std::vector<int> rest;
  {
    std::vector<std::pair<long, int>> found;
    for (size_t i = 0; i < h.names.size(); i++) {
      const std::string& n = h.names[i];
      if (n.compare(0, 7, "f_rest_") != 0 || n.size() == 7) continue;
      bool digits = true;
      for (size_t k = 7; k < n.size(); k++) digits = digits && isdigit((unsigned char)n[k]);
      if (!digits) continue;
      found.push_back({strtol(n.c_str() + 7, nullptr, 10), (int)i});
    }
    std::sort(found.begin(), found.end());
    for (size_t k = 0; k < found.size(); k++) {
      if (found[k].first != (long)k)
        return ply_fail("f_rest_* properties must be contiguous from f_rest_0");
      rest.push_back(found[k].second);
    }
  }
  if (rest.size() % 3 != 0) return ply_fail("f_rest_* property count must be divisible by three");
  size_t per_channel = rest.size() / 3, coefs = per_channel + 1;
  uint32_t degree = coefs == 1 ? 0u : coefs == 4 ? 1u : coefs == 9 ? 2u : coefs == 16 ? 3u : 9u;
  if (degree > 3)
    return ply_fail("PLY must contain one, four, nine, or sixteen SH coefficients per channel");
  info->count = (uint32_t)h.count;
  info->sh_degree = degree;
  info->format = (uint32_t)h.format;
  if (!means && !scales && !opacity && !quats && !sh) return 0;  // header only
  if (!means || !scales || !opacity || !quats || !sh) return ply_fail("missing output arrays");
  const size_t nprop = h.names.size();
  std::vector<double> ascii;
  if (h.format == 0) {
    // Whitespace-separated numbers, count * properties of them.
    ascii.resize(h.count * nprop);
    size_t at = h.data;
    for (size_t k = 0; k < ascii.size(); k++) {
      while (at < size && isspace(b[at])) at++;
      size_t start = at;
      while (at < size && !isspace(b[at])) at++;
      std::string tok((const char*)b + start, at - start);
      char* end = nullptr;
      double v = strtod(tok.c_str(), &end);
      if (tok.empty() || !end || *end || !std::isfinite(v)) {
        snprintf(g_ply_error, sizeof(g_ply_error), "Invalid ASCII PLY value at scalar %zu", k);
        return -1;
      }
      ascii[k] = v;
    }
  } else if (h.data + h.count * h.stride > size) {
    return ply_fail("Binary PLY ends before the vertex data is complete");
  }
  const bool little = h.format == 1;
  auto read = [&](size_t v, int prop) -> double {
    if (h.format == 0) return ascii[v * nprop + (size_t)prop];
    return ply_scalar(b + h.data + v * h.stride + h.offsets[(size_t)prop], h.sizes[(size_t)prop],
                      h.kinds[(size_t)prop], little);
  };
  // CanonicalGaussianPlyLoader.ts:130-175, evaluated in double like the
  // JavaScript and stored as float32.
  for (size_t g = 0; g < h.count; g++) {
    for (int k = 0; k < 3; k++) means[g * 3 + k] = (float)read(g, req[k]);
    for (int k = 0; k < 3; k++) scales[g * 3 + k] = (float)std::max(exp(read(g, req[3 + k])), 1e-6);
    opacity[g] = (float)(1.0 / (1.0 + exp(-read(g, req[10]))));
    double w = read(g, req[6]), x = read(g, req[7]), y = read(g, req[8]), z = read(g, req[9]);
    double len = sqrt(x * x + y * y + z * z + w * w);  // Math.hypot
    if (len > 1e-12) {
      quats[g * 4 + 0] = (float)(x / len);
      quats[g * 4 + 1] = (float)(y / len);
      quats[g * 4 + 2] = (float)(z / len);
      quats[g * 4 + 3] = (float)(w / len);
    } else {
      quats[g * 4 + 0] = quats[g * 4 + 1] = quats[g * 4 + 2] = 0.0f;
      quats[g * 4 + 3] = 1.0f;
    }
    float* c = sh + g * coefs * 3;
    for (int k = 0; k < 3; k++) c[k] = (float)read(g, req[11 + k]);
    for (size_t j = 1; j < coefs; j++)
      for (size_t ch = 0; ch < 3; ch++)
        c[j * 3 + ch] = (float)read(g, rest[ch * per_channel + (j - 1)]);
  }
  return 0;
}

// The store's slot buffers for one cloud (engine.rs:1801-1949 with the
// hole rows of :1917-1923): slots [0, n) hold the Gaussians in order with
// the object id in means.w; slots [n, capacity) are holes (means.w = -1,
// every other word 0), which projection skips.
extern "C" int gs_pack_scene(size_t n, size_t capacity, uint32_t object_id, uint32_t sh_degree,
                             const float* means3, const float* scales, const float* opacity,
                             const float* quats, const float* sh, float* out_means,
                             uint32_t* out_shape, uint32_t* out_rotation, uint32_t* out_sh) {
  if (n > capacity || sh_degree > 3) return -1;
  const size_t coefs = (size_t)(sh_degree + 1) * (sh_degree + 1);
  for (size_t i = 0; i < capacity; i++) {
    float* m = out_means + i * 4;
    if (i < n) {
      m[0] = means3[i * 3];
      m[1] = means3[i * 3 + 1];
      m[2] = means3[i * 3 + 2];
      m[3] = (float)object_id;
    } else {
      m[0] = m[1] = m[2] = 0.0f;
      m[3] = -1.0f;
    }
  }
  gs_encode_compact(n, scales, opacity, quats, out_shape, out_rotation);
  gs_encode_sh_rgb8e8(n * coefs, sh, out_sh);
  if (capacity > n) {
    memset(out_shape + n * 2, 0, (capacity - n) * 2 * sizeof(uint32_t));
    memset(out_rotation + n, 0, (capacity - n) * sizeof(uint32_t));
    memset(out_sh + n * coefs, 0, (capacity - n) * coefs * sizeof(uint32_t));
  }
  return 0;
}

namespace {

// xoshiro256** seeded through splitmix64 (Blackman & Vigna, 2018).
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  // Box-Muller (one value per call; the pair's second half is dropped so
  // every draw consumes exactly two uniforms).
  double normal() {
    double u1 = 1.0 - uniform();  // (0, 1]
    double u2 = uniform();
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
};

const double kShC0 = 0.2820948;  // engine.rs:26 (SH_C0), for reference only

}  // namespace

extern "C" void gs_scene_spec_default(gs_scene_spec* s) {
  memset(s, 0, sizeof(*s));
  s->seed = 1;
  s->count = 100000;
  s->clusters = 32;
  s->extent = 5.0f;
  s->cluster_sigma = 0.5f;
  s->log_scale_mean = -4.5f;
  s->log_scale_sigma = 0.6f;
  s->log_scale_min = -9.0f;
  s->log_scale_max = 0.0f;
  s->major_mean = -2.5f;
  s->major_sigma = 0.5f;
  s->minor_mean = -6.0f;
  s->minor_sigma = 0.5f;
  s->opacity_sigma = 1.5f;
  s->sh_degree = 3;
  s->dc_mean = 0.5f;
  s->dc_sigma = 0.2f;
  s->band_sigma = 0.05f;
  s->near_fraction = 0.0f;
  s->near_clusters = 8;
  s->near_depth_min = 0.5f;
  s->near_depth_max = 3.0f;
  s->near_sigma = 0.1f;
  s->camera_position[2] = 15.0;
  s->camera_fov_deg = 60.0;
  s->camera_aspect = 16.0 / 9.0;
  (void)kShC0;
}

extern "C" int gs_generate_scene(const gs_scene_spec* sp, float* means, uint32_t* shape,
                                 uint32_t* rotation, uint32_t* sh, float* scales_out,
                                 float* opacity_out, float* quats_out) {
  if (!sp || !means || !shape || !rotation || !sh || sp->sh_degree > 3 || sp->clusters == 0)
    return -1;
  const gs_scene_spec s = *sp;
  Rng root(s.seed);
  std::vector<double> centers((size_t)s.clusters * 3);
  for (auto& c : centers) c = (root.uniform() * 2.0 - 1.0) * s.extent;
  // Near-plane cluster centres inside the frustum at view depth d.
  uint32_t near_n = s.near_fraction > 0.0f ? std::max(1u, s.near_clusters) : 0u;
  std::vector<double> near_centers((size_t)near_n * 3);
  if (near_n) {
    double f[3], r[3], u[3];
    for (int i = 0; i < 3; i++) f[i] = s.camera_target[i] - s.camera_position[i];
    double fl = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= fl;
    double up[3] = {0.0, 1.0, 0.0};
    r[0] = f[1] * up[2] - f[2] * up[1];
    r[1] = f[2] * up[0] - f[0] * up[2];
    r[2] = f[0] * up[1] - f[1] * up[0];
    double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& v : r) v /= rl;
    u[0] = r[1] * f[2] - r[2] * f[1];
    u[1] = r[2] * f[0] - r[0] * f[2];
    u[2] = r[0] * f[1] - r[1] * f[0];
    double th = tan(s.camera_fov_deg * 0.5 * 3.141592653589793 / 180.0);
    for (uint32_t c = 0; c < near_n; c++) {
      double d = s.near_depth_min + root.uniform() * (s.near_depth_max - s.near_depth_min);
      double x = (root.uniform() * 2.0 - 1.0) * 0.8 * d * th * s.camera_aspect;
      double y = (root.uniform() * 2.0 - 1.0) * 0.8 * d * th;
      for (int i = 0; i < 3; i++)
        near_centers[c * 3 + i] = s.camera_position[i] + f[i] * d + r[i] * x + u[i] * y;
    }
  }
  const uint32_t ncoef = (s.sh_degree + 1) * (s.sh_degree + 1);
  const size_t chunk = 65536;
  size_t chunks = (s.count + chunk - 1) / chunk;
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  unsigned threads = (unsigned)std::min<size_t>(std::min(hw, 32u), std::max<size_t>(1, chunks));
  auto work = [&](unsigned t) {
    std::vector<float> sc(chunk * 3), op(chunk), qt(chunk * 4), rgb((size_t)ncoef * 3);
    for (size_t c = t; c < chunks; c += threads) {
      Rng rng(s.seed ^ (0xa0761d6478bd642full * (c + 1)));
      size_t b = c * chunk, e = std::min<size_t>(s.count, b + chunk);
      for (size_t i = b; i < e; i++) {
        size_t li = i - b;
        double p[3];
        bool near = near_n && rng.uniform() < s.near_fraction;
        if (near) {
          uint32_t k = (uint32_t)(rng.uniform() * near_n) % near_n;
          for (int a = 0; a < 3; a++)
            p[a] = near_centers[k * 3 + a] + rng.normal() * s.near_sigma;
        } else {
          uint32_t k = (uint32_t)(rng.uniform() * s.clusters) % s.clusters;
          for (int a = 0; a < 3; a++) p[a] = centers[k * 3 + a] + rng.normal() * s.cluster_sigma;
        }
        means[i * 4] = (float)p[0];
        means[i * 4 + 1] = (float)p[1];
        means[i * 4 + 2] = (float)p[2];
        means[i * 4 + 3] = 0.0f;  // object id 0
        double ls[3];
        if (s.anisotropic) {
          uint32_t major = (uint32_t)(rng.uniform() * 3.0) % 3u;
          for (int a = 0; a < 3; a++)
            ls[a] = (uint32_t)a == major ? s.major_mean + rng.normal() * s.major_sigma
                                         : s.minor_mean + rng.normal() * s.minor_sigma;
        } else {
          for (int a = 0; a < 3; a++) ls[a] = s.log_scale_mean + rng.normal() * s.log_scale_sigma;
        }
        for (int a = 0; a < 3; a++) {
          double v = std::min<double>(std::max<double>(ls[a], s.log_scale_min), s.log_scale_max);
          sc[li * 3 + a] = (float)exp(v);
        }
        op[li] = (float)(1.0 / (1.0 + exp(-rng.normal() * s.opacity_sigma)));
        // Shoemake's uniform rotation.
        double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        double a1 = sqrt(1.0 - u1), a2 = sqrt(u1);
        qt[li * 4] = (float)(a1 * sin(6.283185307179586 * u2));
        qt[li * 4 + 1] = (float)(a1 * cos(6.283185307179586 * u2));
        qt[li * 4 + 2] = (float)(a2 * sin(6.283185307179586 * u3));
        qt[li * 4 + 3] = (float)(a2 * cos(6.283185307179586 * u3));
        for (uint32_t k = 0; k < ncoef; k++)
          for (int ch = 0; ch < 3; ch++)
            rgb[k * 3 + ch] = k == 0 ? (float)(s.dc_mean + rng.normal() * s.dc_sigma)
                                     : (float)(rng.normal() * s.band_sigma);
        gs_encode_sh_rgb8e8(ncoef, rgb.data(), sh + i * ncoef);
      }
      size_t m = e - b;
      gs_encode_compact(m, sc.data(), op.data(), qt.data(), shape + b * 2, rotation + b);
      if (scales_out) memcpy(scales_out + b * 3, sc.data(), m * 3 * sizeof(float));
      if (opacity_out) memcpy(opacity_out + b, op.data(), m * sizeof(float));
      if (quats_out) memcpy(quats_out + b * 4, qt.data(), m * 4 * sizeof(float));
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; t++) pool.emplace_back(work, t);
  for (auto& th : pool) th.join();
  return 0;
}
