// gs_oracle.cpp - CPU oracle for the tiled Gaussian-splatting frame path.
//
// TEST INFRASTRUCTURE ONLY (see gs_oracle.h). Each function cites the
// reference file:line it restates. Paths are relative to the reference
// repository root (vinneyto/3dgs_tile_webgpu).
//
// Parity status per part (what pins it, besides the GPU comparison):
//  - compact decode, rgb8e8 decode: closed forms + the producer round-trip
//    bounds of rust/gaussian-backend/tests/backend.rs:1141-1156 and
//    test/GaussianSh.test.ts:19-40 (tests/test_formats.py).
//  - projection helpers: the known-answer tests of
//    test/ProjectionBoundary.test.ts:30-72 and the brute-force tile and
//    periodic-delta fixtures of :74-175 (tests/test_oracle_known_answers.py).
//  - scans, compaction, ordered counts, offsets: closed forms (exclusive
//    prefix with saturation) checked against numpy in tests.
//  - sorts: equal to a lexicographic (tile, depth key, ordinal) sort, checked
//    against numpy lexsort in tests.
//  - raster: brute-force FP64 blend on the same lists (gs_oracle_blend_f64),
//    the chunk-composite identity of test/rasterizationWGSL.test.ts:63-101.
//  - The bit-level float outputs of the reference's TSL bodies on a real
//    WebGPU driver are not reproducible here (WGSL builtins are
//    implementation-defined): parity unpinned at that level only.

#include "gs_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "gs_math.h"

namespace {

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };
struct M3 { V3 c[3]; };   // columns
struct M4 { V4 c[4]; };   // columns

const float kMaxF32 = 3.402823e+38f;  // projectionHelpers.ts:44,100

// Tile edge of a frame: the reference's 16 unless the scene asks for 8 or
// 32 (launch shapes of the batched renderer, checked at those sizes).
inline uint32_t tile_edge(const gs_oracle_scene& s) { return s.tile_size ? s.tile_size : 16u; }

inline float comp(const V3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline float comp(const V4& v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
inline float row(const M3& m, int r, int c) { return comp(m.c[c], r); }

// WGSL matrix products: each element sums k = 0..2 left to right.
M3 mul(const M3& a, const M3& b) {
  M3 o;
  for (int j = 0; j < 3; j++) {
    float v[3];
    for (int i = 0; i < 3; i++) {
      float s = row(a, i, 0) * row(b, 0, j);
      s = s + row(a, i, 1) * row(b, 1, j);
      s = s + row(a, i, 2) * row(b, 2, j);
      v[i] = s;
    }
    o.c[j] = {v[0], v[1], v[2]};
  }
  return o;
}
M3 transpose(const M3& a) {
  M3 o;
  for (int j = 0; j < 3; j++) o.c[j] = {row(a, j, 0), row(a, j, 1), row(a, j, 2)};
  return o;
}
V3 mul(const M3& m, const V3& v) {
  float r[3];
  for (int i = 0; i < 3; i++) {
    float s = row(m, i, 0) * v.x;
    s = s + row(m, i, 1) * v.y;
    s = s + row(m, i, 2) * v.z;
    r[i] = s;
  }
  return {r[0], r[1], r[2]};
}
V4 mul(const M4& m, const V4& v) {
  float r[4];
  for (int i = 0; i < 4; i++) {
    float s = comp(m.c[0], i) * v.x;
    s = s + comp(m.c[1], i) * v.y;
    s = s + comp(m.c[2], i) * v.z;
    s = s + comp(m.c[3], i) * v.w;
    r[i] = s;
  }
  return {r[0], r[1], r[2], r[3]};
}
float dot(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V3 normalize(const V3& v) {
  float len = gs_sqrtf(dot(v, v));
  return {v.x / len, v.y / len, v.z / len};
}
V4 normalize(const V4& v) {
  float d = ((v.x * v.x + v.y * v.y) + v.z * v.z) + v.w * v.w;
  float len = gs_sqrtf(d);
  return {v.x / len, v.y / len, v.z / len, v.w / len};
}
M4 load_m4(const float* p) {
  M4 m;
  for (int c = 0; c < 4; c++) m.c[c] = {p[c * 4], p[c * 4 + 1], p[c * 4 + 2], p[c * 4 + 3]};
  return m;
}

// ---------------------------------------------------------------- decode ----

// kernels/compactGaussians.ts:2-7 (decodeCompactScaleOpacity)
V4 decode_scale_opacity(uint32_t x, uint32_t y) {
  float sx = gs_half_to_float(x & 0xffffu), sy = gs_half_to_float(x >> 16);
  float sz = gs_half_to_float(y & 0xffffu), a = gs_half_to_float(y >> 16);
  return {gs_expf(sx), gs_expf(sy), gs_expf(sz), a};
}

// kernels/compactGaussians.ts:10-19 (decodeCompactRotation)
V4 decode_rotation(uint32_t packed) {
  float px = (float)(packed & 1023u) / 1023.0f * 2.0f - 1.0f;
  float py = (float)((packed >> 10) & 1023u) / 1023.0f * 2.0f - 1.0f;
  float z = 1.0f - gs_fabsf(px) - gs_fabsf(py);
  float t = gs_maxf(-z, 0.0f);
  float x = px >= 0.0f ? px + -t : px + t;
  float y = py >= 0.0f ? py + -t : py + t;
  V3 axis = normalize(V3{x, y, z});
  float half_angle = (float)(packed >> 20) / 4095.0f * 1.5707963267948966f;
  float s = gs_sinf(half_angle);
  return {axis.x * s, axis.y * s, axis.z * s, gs_cosf(half_angle)};
}

// kernels/projectionHelpers.ts:123-127 (decode_sh_rgb8e8):
// unpack4x8snorm = max(byte / 127, -1), times exp2(exponent byte - 127).
V3 decode_sh(uint32_t packed) {
  float m[3];
  for (int i = 0; i < 3; i++) {
    int8_t b = (int8_t)((packed >> (8 * i)) & 0xffu);
    m[i] = gs_maxf((float)b / 127.0f, -1.0f);
  }
  float scale = gs_exp2i((int)((packed >> 24) & 255u) - 127);
  return {m[0] * scale, m[1] * scale, m[2] * scale};
}

// ------------------------------------------------------------ projection ----

// kernels/projectionHelpers.ts:4-35 (gaussian_view_covariance)
M3 view_covariance(const V3& scale_input, const V4& rotation_input, const M4& mv) {
  V3 s = {gs_maxf(scale_input.x, 1e-7f), gs_maxf(scale_input.y, 1e-7f),
          gs_maxf(scale_input.z, 1e-7f)};
  V4 q = normalize(rotation_input);
  float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
  float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
  float xw = q.x * q.w, yw = q.y * q.w, zw = q.z * q.w;
  M3 r;
  r.c[0] = {1.0f - 2.0f * (yy + zz), 2.0f * (xy + zw), 2.0f * (xz - yw)};
  r.c[1] = {2.0f * (xy - zw), 1.0f - 2.0f * (xx + zz), 2.0f * (yz + xw)};
  r.c[2] = {2.0f * (xz + yw), 2.0f * (yz - xw), 1.0f - 2.0f * (xx + yy)};
  M3 d;
  d.c[0] = {s.x * s.x, 0.0f, 0.0f};
  d.c[1] = {0.0f, s.y * s.y, 0.0f};
  d.c[2] = {0.0f, 0.0f, s.z * s.z};
  M3 local = mul(mul(r, d), transpose(r));
  M3 l;
  l.c[0] = {mv.c[0].x, mv.c[0].y, mv.c[0].z};
  l.c[1] = {mv.c[1].x, mv.c[1].y, mv.c[1].z};
  l.c[2] = {mv.c[2].x, mv.c[2].y, mv.c[2].z};
  return mul(mul(l, local), transpose(l));
}

// kernels/projectionHelpers.ts:38-78 (gaussian_projection_result)
M4 projection_result(V2 center, float depth, V3 cov, float fv) {
  M4 zero = {};
  float original_determinant = gs_maxf(cov.x * cov.z - cov.y * cov.y, 0.0f);
  float s00 = cov.x + fv, s01 = cov.y, s11 = cov.z + fv;
  bool finite = s00 == s00 && gs_fabsf(s00) <= kMaxF32 && s01 == s01 &&
                gs_fabsf(s01) <= kMaxF32 && s11 == s11 && gs_fabsf(s11) <= kMaxF32;
  if (!finite) return zero;
  float mid = 0.5f * (s00 + s11);
  float radius = gs_sqrtf(0.25f * (s00 - s11) * (s00 - s11) + s01 * s01);
  float lmin = gs_clampf(mid - radius, 1e-6f, 1e4f);
  float lmax = gs_clampf(mid + radius, 1e-6f, 1e4f);
  float theta = 0.5f * gs_atan2f(2.0f * s01, s00 - s11);
  float cs = gs_cosf(theta), sn = gs_sinf(theta);
  s00 = lmin * sn * sn + lmax * cs * cs;
  s01 = (lmax - lmin) * cs * sn;
  s11 = lmin * cs * cs + lmax * sn * sn;
  float det = s00 * s11 - s01 * s01;
  if (det <= 1e-8f) return zero;
  float inv = 1.0f / det;
  M4 o;
  o.c[0] = {center.x, center.y, depth, 1.0f};
  o.c[1] = {s11 * inv, -s01 * inv, s00 * inv, det};
  o.c[2] = {s00, s01, s11, original_determinant};
  o.c[3] = {0.0f, 0.0f, 0.0f, 0.0f};
  return o;
}

// kernels/projectionHelpers.ts:80-96 (perspective_gaussian_projection)
M4 perspective_projection(V3 view, const M3& cov_view, const M4& proj, V4 viewport,
                          float fv) {
  M4 zero = {};
  float depth = -view.z;
  if (!(depth > viewport.z && depth < viewport.w)) return zero;
  V4 clip = mul(proj, V4{view.x, view.y, view.z, 1.0f});
  if (clip.w <= 0.0f) return zero;
  float ndcx = clip.x / clip.w, ndcy = clip.y / clip.w;
  V2 center = {(ndcx * 0.5f + 0.5f) * viewport.x, (0.5f - ndcy * 0.5f) * viewport.y};
  float fx = 0.5f * viewport.x * proj.c[0].x;
  float fy = 0.5f * viewport.y * proj.c[1].y;
  float inv = 1.0f / depth;
  V3 j0 = {fx * inv, 0.0f, fx * view.x * inv * inv};
  V3 j1 = {0.0f, -fy * inv, -fy * view.y * inv * inv};
  V3 cov = {dot(j0, mul(cov_view, j0)), dot(j0, mul(cov_view, j1)),
            dot(j1, mul(cov_view, j1))};
  return projection_result(center, depth, cov, fv);
}

// kernels/projectionHelpers.ts:98-116 (gaussian_projection_is_valid)
bool projection_is_valid(const M4& p) {
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 4; k++) {
      float v = comp(p.c[i], k);
      if (!(v == v) || !(gs_fabsf(v) <= kMaxF32)) return false;
    }
  V3 cov = {p.c[2].x, p.c[2].y, p.c[2].z};
  float det = cov.x * cov.z - cov.y * cov.y;
  V3 conic = {p.c[1].x, p.c[1].y, p.c[1].z};
  if (!(p.c[0].w > 0.0f && p.c[0].z > 0.0f && det > 1e-8f && cov.x > 0.0f &&
        cov.z > 0.0f && cov.x <= 1e4f && cov.z <= 1e4f && conic.x > 0.0f &&
        conic.z > 0.0f && p.c[2].w >= 0.0f))
    return false;
  V3 e = {cov.z / det, -cov.y / det, cov.x / det};
  if (!(gs_fabsf(p.c[1].w - det) <= gs_maxf(1e-6f, det * 1e-3f))) return false;
  return gs_fabsf(conic.x - e.x) <= gs_maxf(1e-6f, gs_fabsf(e.x) * 1e-3f) &&
         gs_fabsf(conic.y - e.y) <= gs_maxf(1e-6f, gs_fabsf(e.y) * 1e-3f) &&
         gs_fabsf(conic.z - e.z) <= gs_maxf(1e-6f, gs_fabsf(e.z) * 1e-3f);
}

// The largest rgb8e8 exponent of the higher SH bands minus the bias (a
// learned-cull feature; no reference counterpart): -127 without them.
float sh_top_exponent(const uint32_t* sh, uint32_t slot, uint32_t degree) {
  uint32_t n = (degree + 1) * (degree + 1);
  const uint32_t* c = sh + (size_t)slot * n;
  uint32_t top = 0u;
  for (uint32_t i = 1; i < n; i++) top = std::max(top, c[i] >> 24);
  return (float)(int)top - 127.0f;
}

// kernels/projectionHelpers.ts:118-176 (evaluate_gaussian_sh_rgb8e8)
V3 evaluate_sh(const uint32_t* sh, uint32_t slot, uint32_t degree, V3 dir) {
  float x = dir.x, y = dir.y, z = dir.z;
  uint32_t n = (degree + 1) * (degree + 1);
  const uint32_t* c = sh + (size_t)slot * n;
  V3 color;
  {
    V3 k = decode_sh(c[0]);
    float w = 0.28209479177387814f;
    color = {w * k.x, w * k.y, w * k.z};
  }
  auto add = [&](float w, uint32_t i) {
    V3 k = decode_sh(c[i]);
    color.x = color.x + w * k.x;
    color.y = color.y + w * k.y;
    color.z = color.z + w * k.z;
  };
  if (degree >= 1) {
    add(-0.4886025119029199f * y, 1);
    add(0.4886025119029199f * z, 2);
    add(-0.4886025119029199f * x, 3);
  }
  if (degree >= 2) {
    float xx = x * x, yy = y * y, zz = z * z;
    add(1.0925484305920792f * x * y, 4);
    add(-1.0925484305920792f * y * z, 5);
    add(0.31539156525252005f * (2.0f * zz - xx - yy), 6);
    add(-1.0925484305920792f * x * z, 7);
    add(0.5462742152960396f * (xx - yy), 8);
  }
  if (degree >= 3) {
    float xx = x * x, yy = y * y, zz = z * z;
    add(-0.5900435899266435f * y * (3.0f * xx - yy), 9);
    add(2.890611442640554f * x * y * z, 10);
    add(-0.4570457994644658f * y * (4.0f * zz - xx - yy), 11);
    add(0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy), 12);
    add(-0.4570457994644658f * x * (4.0f * zz - xx - yy), 13);
    add(1.445305721320277f * z * (xx - yy), 14);
    add(-0.5900435899266435f * x * (xx - 3.0f * yy), 15);
  }
  return {gs_clampf(color.x + 0.5f, 0.0f, 1.0f), gs_clampf(color.y + 0.5f, 0.0f, 1.0f),
          gs_clampf(color.z + 0.5f, 0.0f, 1.0f)};
}

// kernels/projectionHelpers.ts:185-221 (subpixel_has_sample)
bool subpixel_has_sample(V2 center, V3 conic, float thr, V2 extent, uint32_t vx,
                         uint32_t vy, uint32_t axis) {
  if (extent.x * 2.0f > 1.0f || extent.y * 2.0f > 1.0f) return true;
  V2 cc = center;
  if (axis == 1) cc.x = cc.x - gs_floorf(center.x / (float)vx) * (float)vx;
  if (axis == 2) cc.y = cc.y - gs_floorf(center.y / (float)vy) * (float)vy;
  int minx = gs_f2i(gs_ceilf(cc.x - extent.x - 0.5f));
  int miny = gs_f2i(gs_ceilf(cc.y - extent.y - 0.5f));
  int maxx = gs_f2i(gs_floorf(cc.x + extent.x - 0.5f));
  int maxy = gs_f2i(gs_floorf(cc.y + extent.y - 0.5f));
  int pminx = axis == 1 ? minx : std::max(minx, 0);
  int pminy = axis == 2 ? miny : std::max(miny, 0);
  int pmaxx = axis == 1 ? maxx : std::min(maxx, (int)vx - 1);
  int pmaxy = axis == 2 ? maxy : std::min(maxy, (int)vy - 1);
  for (int py = pminy; py <= pmaxy; py++)
    for (int px = pminx; px <= pmaxx; px++) {
      float dx = ((float)px + 0.5f) - cc.x;
      float dy = ((float)py + 0.5f) - cc.y;
      float sigma =
          0.5f * (conic.x * dx * dx + 2.0f * conic.y * dx * dy + conic.z * dy * dy);
      if (sigma <= thr) return true;
    }
  return false;
}

// kernels/tileContribution.ts:16-59 (tileContributionWGSL: the tile/ellipse
// test) inside kernels/tileContribution.ts:84-114 (gaussian_boundary_tile_hit)
bool tile_hit(V2 center, V3 conic, float thr, V2 radius, int tx, int ty, V2 viewport,
              uint32_t axis, float ts = 16.0f) {
  V2 cc = center;
  int first = 0, last = 0;
  if (axis != 0) {
    float period = axis == 1 ? viewport.x : viewport.y;
    float c = axis == 1 ? center.x : center.y;
    float r = axis == 1 ? radius.x : radius.y;
    float canon = c - gs_floorf(c / period) * period;
    if (axis == 1) cc.x = canon; else cc.y = canon;
    float tile_start = (float)(axis == 1 ? tx : ty) * ts;
    first = gs_f2i(gs_ceilf((tile_start - r - canon) / period));
    last = gs_f2i(gs_floorf((tile_start + ts + r - canon) / period));
  }
  for (int image = first; image <= last; image++) {
    V2 ic = cc;
    if (axis == 1) ic.x = ic.x + (float)image * viewport.x;
    if (axis == 2) ic.y = ic.y + (float)image * viewport.y;
    float rminx = (float)tx * ts, rminy = (float)ty * ts;
    float rmaxx = rminx + ts, rmaxy = rminy + ts;
    bool x_left = ic.x < rminx, x_right = ic.x > rmaxx;
    bool in_x = !(x_left || x_right);
    bool y_above = ic.y < rminy, y_below = ic.y > rmaxy;
    bool in_y = !(y_above || y_below);
    bool contributes = in_x && in_y;
    if (!contributes) {
      float cx = x_left ? rminx : rmaxx, cy = y_above ? rminy : rmaxy;
      float ex = x_left ? ts : -ts, ey = y_above ? ts : -ts;
      float dx = ic.x - cx, dy = ic.y - cy;
      float tx_raw = (ex * conic.x * dx + ex * conic.y * dy) / (ex * conic.x * ex);
      float ty_raw = (ey * conic.y * dx + ey * conic.z * dy) / (ey * conic.z * ey);
      float tt_x = in_y ? 0.0f : gs_clampf(tx_raw, 0.0f, 1.0f);
      float tt_y = in_x ? 0.0f : gs_clampf(ty_raw, 0.0f, 1.0f);
      float clx = cx + tt_x * ex, cly = cy + tt_y * ey;
      float ddx = clx - ic.x, ddy = cly - ic.y;
      float sigma = 0.5f * (conic.x * ddx * ddx + conic.z * ddy * ddy) + conic.y * ddx * ddy;
      contributes = sigma <= thr;
    }
    if (contributes) return true;
  }
  return false;
}

// kernels/projectionHelpers.ts:223-236 (count_contributing_tiles)
uint32_t count_tiles(V2 center, V3 conic, float thr, V2 radius, int minx, int miny,
                     int maxx, int maxy, V2 viewport, uint32_t axis, float ts = 16.0f) {
  uint32_t n = 0;
  for (int ty = miny; ty <= maxy; ty++)
    for (int tx = minx; tx <= maxx; tx++)
      if (tile_hit(center, conic, thr, radius, tx, ty, viewport, axis, ts)) n++;
  return n;
}

// kernels/tileContribution.ts:66-82 (gaussian_boundary_delta)
V2 boundary_delta(V2 pixel, V2 center, V3 conic, V2 viewport, uint32_t axis) {
  V2 cc = center;
  if (axis == 1) cc.x = cc.x - gs_floorf(center.x / viewport.x) * viewport.x;
  if (axis == 2) cc.y = cc.y - gs_floorf(center.y / viewport.y) * viewport.y;
  V2 d = {pixel.x - cc.x, pixel.y - cc.y};
  if (axis == 1) {
    float optimum = -conic.y * d.y / conic.x;
    d.x = d.x - gs_floorf((d.x - optimum) / viewport.x + 0.5f) * viewport.x;
  }
  if (axis == 2) {
    float optimum = -conic.y * d.x / conic.z;
    d.y = d.y - gs_floorf((d.y - optimum) / viewport.y + 0.5f) * viewport.y;
  }
  return d;
}

// ----------------------------------------------------------- threading ----

template <class F>
void parallel_for(uint32_t threads, size_t n, F f) {
  if (threads <= 1 || n < 4096) {
    f((size_t)0, n, 0u);
    return;
  }
  std::vector<std::thread> pool;
  size_t chunk = (n + threads - 1) / threads;
  for (uint32_t t = 0; t < threads; t++) {
    size_t b = (size_t)t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([=]() { f(b, e, t); });
  }
  for (auto& th : pool) th.join();
}

// Stable sort on the pair key, parallel: sorted runs merged pairwise in
// order, which keeps equal keys in input order.
void parallel_stable_sort(std::vector<uint64_t>& v, uint32_t threads,
                          bool (*less)(uint64_t, uint64_t)) {
  size_t n = v.size();
  if (threads <= 1 || n < 65536) {
    std::stable_sort(v.begin(), v.end(), less);
    return;
  }
  size_t runs = threads;
  std::vector<size_t> bounds(runs + 1);
  for (size_t i = 0; i <= runs; i++) bounds[i] = n * i / runs;
  {
    std::vector<std::thread> pool;
    for (size_t r = 0; r < runs; r++)
      pool.emplace_back([&, r]() {
        std::stable_sort(v.begin() + bounds[r], v.begin() + bounds[r + 1], less);
      });
    for (auto& th : pool) th.join();
  }
  std::vector<uint64_t> tmp(n);
  while (bounds.size() > 2) {
    std::vector<size_t> next;
    std::vector<std::thread> pool;
    for (size_t r = 0; r + 2 < bounds.size(); r += 2) {
      size_t a = bounds[r], m = bounds[r + 1], e = bounds[r + 2];
      pool.emplace_back([&, a, m, e]() {
        std::merge(v.begin() + a, v.begin() + m, v.begin() + m, v.begin() + e,
                   tmp.begin() + a, less);
      });
      next.push_back(a);
    }
    for (auto& th : pool) th.join();
    if ((bounds.size() - 1) % 2 == 1) {  // odd run count: copy the last run
      size_t a = bounds[bounds.size() - 2], e = bounds.back();
      std::copy(v.begin() + a, v.begin() + e, tmp.begin() + a);
      next.push_back(a);
    }
    next.push_back(n);
    v.swap(tmp);
    bounds.swap(next);
  }
}

// Records are packed as (first << 32 | second) with first the sort key.
bool key_less(uint64_t a, uint64_t b) { return (a >> 32) < (b >> 32); }

inline uint32_t sat_add(uint32_t a, uint32_t b) {
  return a + std::min(b, 0xffffffffu - a);  // kernels/scan.ts:35,57,95
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

struct gs_oracle_result {
  gs_oracle_scene scene;
  gs_oracle_counts counts;
  uint32_t width, height;
  std::vector<float> projected_mean, projected_conic, projected_color;
  std::vector<uint32_t> tile_counts, visible_offsets, ordered_counts,
      intersection_offsets, tile_offsets;
  std::vector<uint64_t> depth_records, emitted, sorted;
  std::vector<float> image, depth_image, final_t, depth_target;
  std::vector<uint16_t> image_half;
  // Learned-cull training data (scene.cull_outputs).
  std::vector<float> cull_features;
  std::unique_ptr<std::atomic<uint32_t>[]> max_contribution;
};

namespace {

// pipeline/ProjectionStage.ts:385-649 (projection kernel, default node slots)
void project_one(const gs_oracle_scene& s, uint32_t tiles_x, uint32_t tiles_y,
                 uint32_t gid, gs_oracle_result& r) {
  float* pm = &r.projected_mean[(size_t)gid * 4];
  float* pc = &r.projected_conic[(size_t)gid * 4];
  float* pk = &r.projected_color[(size_t)gid * 4];
  r.tile_counts[gid] = 0;
  pm[0] = pm[1] = pm[2] = pm[3] = 0.0f;
  uint32_t slot = s.active_slots ? s.active_slots[gid] : gid;
  const float* mean = s.means + (size_t)slot * 4;
  if (mean[3] < 0.0f) return;
  uint32_t object = gs_f2u(mean[3]);
  V4 shape = decode_scale_opacity(s.scales_opacity[(size_t)slot * 2],
                                  s.scales_opacity[(size_t)slot * 2 + 1]);
  V4 rot = decode_rotation(s.rotations[slot]);
  // Object ids beyond the published table read zero rows in the reference's
  // object tail (ObjectFrameState.ts:283): treated as invisible.
  if (object >= s.object_count) return;
  const float* frame = s.objects + (size_t)object * 40;
  M4 model_view = load_m4(frame + 16);
  V3 camera_local = {frame[32], frame[33], frame[34]};
  if (!(frame[36] > 0.0f)) return;
  V3 local = {mean[0], mean[1], mean[2]};
  V3 view_dir = normalize(V3{local.x - camera_local.x, local.y - camera_local.y,
                             local.z - camera_local.z});
  V4 view = mul(model_view, V4{local.x, local.y, local.z, 1.0f});
  M3 cov_view = view_covariance(V3{shape.x, shape.y, shape.z}, rot, model_view);
  M4 proj = load_m4(s.projection);
  V4 viewport = {s.width, s.height, s.near_plane, s.far_plane};
  M4 p = perspective_projection(V3{view.x, view.y, view.z}, cov_view, proj, viewport,
                                s.filter_variance);
  if (!projection_is_valid(p)) return;
  V2 center = {p.c[0].x, p.c[0].y};
  float depth = p.c[0].z;
  V3 conic = {p.c[1].x, p.c[1].y, p.c[1].z};
  float det = p.c[1].w;
  V3 cov = {p.c[2].x, p.c[2].y, p.c[2].z};
  float original_det = p.c[2].w;
  float opacity = gs_clampf(shape.w, 0.0f, 1.0f);
  if (s.antialias_compensated)
    opacity = opacity * gs_sqrtf(gs_clampf(original_det / det, 0.0f, 1.0f));
  if (opacity < (float)(1.0 / 255.0) || opacity != opacity) return;
  float thr = gs_logf(opacity * 255.0f);
  float extent_x = gs_sqrtf(thr * 2.0f * gs_clampf(cov.x, 1e-12f, 1e4f));
  float extent_y = gs_sqrtf(thr * 2.0f * gs_clampf(cov.z, 1e-12f, 1e4f));
  float radius_x = gs_ceilf(extent_x), radius_y = gs_ceilf(extent_y);
  if (radius_x <= 0.0f || radius_y <= 0.0f) return;
  V2 bmin = {center.x - radius_x, center.y - radius_y};
  V2 bmax = {center.x + radius_x, center.y + radius_y};
  if ((s.boundary_axis != 1 && (bmax.x < 0.0f || bmin.x >= s.width)) ||
      (s.boundary_axis != 2 && (bmax.y < 0.0f || bmin.y >= s.height)))
    return;
  if (s.subpixel_culling &&
      !subpixel_has_sample(center, conic, thr, V2{extent_x, extent_y}, gs_f2u(s.width),
                           gs_f2u(s.height), s.boundary_axis)) {
    pm[0] = center.x;
    pm[1] = center.y;
    pm[2] = depth;
    pm[3] = -1.0f;
    return;
  }
  float max_tx = (float)((int)tiles_x - 1), max_ty = (float)((int)tiles_y - 1);
  float ts = (float)tile_edge(s);
  int tmin_x = gs_f2i(gs_clampf(gs_floorf(bmin.x / ts), 0.0f, max_tx));
  int tmin_y = gs_f2i(gs_clampf(gs_floorf(bmin.y / ts), 0.0f, max_ty));
  int tmax_x = gs_f2i(gs_clampf(gs_floorf(bmax.x / ts), 0.0f, max_tx));
  int tmax_y = gs_f2i(gs_clampf(gs_floorf(bmax.y / ts), 0.0f, max_ty));
  if (s.boundary_axis == 1) { tmin_x = 0; tmax_x = (int)tiles_x - 1; }
  if (s.boundary_axis == 2) { tmin_y = 0; tmax_y = (int)tiles_y - 1; }
  V3 color = evaluate_sh(s.sh, slot, s.sh_degree, view_dir);
  uint32_t count = count_tiles(center, conic, thr, V2{radius_x, radius_y}, tmin_x,
                               tmin_y, tmax_x, tmax_y, V2{s.width, s.height},
                               s.boundary_axis, ts);
  if (count == 0) return;
  if (s.cull_outputs) {
    // Learned-cull features: log pixel radius, log conic diagonal, conic
    // correlation, opacity, position . view direction, log depth, view
    // direction, object-space position, top higher-band SH exponent, normalised screen
    // position.
    float* f = &r.cull_features[(size_t)gid * 16];
    f[0] = gs_logf_pos(radius_x > radius_y ? radius_x : radius_y);
    f[1] = gs_logf_pos(conic.x);
    f[2] = conic.y / gs_sqrtf(conic.x * conic.z);
    f[3] = gs_logf_pos(conic.z);
    f[4] = opacity;
    f[5] = local.x * view_dir.x + local.y * view_dir.y + local.z * view_dir.z;
    f[6] = gs_logf_pos(depth);
    f[7] = view_dir.x;
    f[8] = view_dir.y;
    f[9] = view_dir.z;
    f[10] = local.x;
    f[11] = local.y;
    f[12] = local.z;
    f[13] = sh_top_exponent(s.sh, slot, s.sh_degree);
    f[14] = center.x * (1.0f / (float)s.width);
    f[15] = center.y * (1.0f / (float)s.height);
  }
  pm[0] = center.x;
  pm[1] = center.y;
  pm[2] = depth;
  pm[3] = opacity;
  pc[0] = conic.x;
  pc[1] = conic.y;
  pc[2] = conic.z;
  pc[3] = radius_x + radius_y * 512.0f;
  pk[0] = gs_clampf(color.x, 0.0f, 1.0f);
  pk[1] = gs_clampf(color.y, 0.0f, 1.0f);
  pk[2] = gs_clampf(color.z, 0.0f, 1.0f);
  pk[3] = gs_b2f(slot ^ 0x3f800000u);  // pipeline/activeSlots.ts:4-8
  r.tile_counts[gid] = count;
}

// kernels/intersections.ts:43-111 (emit_intersections), one depth rank.
uint32_t emit_one(const gs_oracle_result& r, uint32_t tiles_x, uint32_t tiles_y,
                  uint32_t rank, uint32_t capacity, std::vector<uint64_t>& out) {
  const gs_oracle_scene& s = r.scene;
  uint32_t gid = (uint32_t)(r.depth_records[rank] & 0xffffffffu);
  const float* mean = &r.projected_mean[(size_t)gid * 4];
  const float* conic = &r.projected_conic[(size_t)gid * 4];
  uint32_t packed = gs_f2u(conic[3]);
  V2 radius = {(float)(packed % 512u), (float)(packed / 512u)};
  V2 center = {mean[0], mean[1]};
  float thr = gs_logf(mean[3] * 255.0f);
  int max_tx = (int)tiles_x - 1, max_ty = (int)tiles_y - 1;
  float ts = (float)tile_edge(s);
  int tmin_x = std::clamp(gs_f2i(gs_floorf((center.x - radius.x) / ts)), 0, max_tx);
  int tmin_y = std::clamp(gs_f2i(gs_floorf((center.y - radius.y) / ts)), 0, max_ty);
  int tmax_x = std::clamp(gs_f2i(gs_floorf((center.x + radius.x) / ts)), 0, max_tx);
  int tmax_y = std::clamp(gs_f2i(gs_floorf((center.y + radius.y) / ts)), 0, max_ty);
  if (s.boundary_axis == 1) { tmin_x = 0; tmax_x = max_tx; }
  if (s.boundary_axis == 2) { tmin_y = 0; tmax_y = max_ty; }
  uint32_t reserved = r.ordered_counts[rank];
  uint32_t base = r.intersection_offsets[rank];
  uint32_t local = 0;
  V3 c = {conic[0], conic[1], conic[2]};
  for (int ty = tmin_y; ty <= tmax_y; ty++)
    for (int tx = tmin_x; tx <= tmax_x; tx++)
      if (tile_hit(center, c, thr, radius, tx, ty, V2{s.width, s.height},
                   s.boundary_axis, ts)) {
        if (local < reserved && base + local < capacity)
          out[base + local] = ((uint64_t)((uint32_t)ty * tiles_x + (uint32_t)tx) << 32) | gid;
        local++;
      }
  uint32_t sentinels = 0;
  for (uint32_t pad = local; pad < reserved; pad++) {
    if (base + pad < capacity) out[base + pad] = ((uint64_t)(tiles_x * tiles_y) << 32) | gid;
    sentinels++;
  }
  return sentinels;
}

struct PixelAccum {
  float c[3];
  float t;
  float d;
  uint32_t checked, blended;
};

// pipeline/TileRasterizer.ts:547-709: front-to-back blend of records
// [begin, end) of one tile for one pixel, starting from T = 1.
PixelAccum blend_range(const gs_oracle_result& r, uint32_t begin, uint32_t end,
                       float px, float py) {
  const gs_oracle_scene& s = r.scene;
  PixelAccum a = {{0.0f, 0.0f, 0.0f}, 1.0f, 0.0f, 0, 0};
  for (uint32_t i = begin; i < end; i++) {
    a.checked++;
    uint32_t gid = (uint32_t)(r.sorted[i] & 0xffffffffu);
    const float* mean = &r.projected_mean[(size_t)gid * 4];
    const float* conic = &r.projected_conic[(size_t)gid * 4];
    const float* color = &r.projected_color[(size_t)gid * 4];
    float thr = gs_logf(mean[3] * 255.0f);  // TileRasterizer.ts:579-581
    V2 d = boundary_delta(V2{px, py}, V2{mean[0], mean[1]},
                          V3{conic[0], conic[1], conic[2]}, V2{s.width, s.height},
                          s.boundary_axis);
    float power = (conic[0] * (d.x * d.x) + conic[1] * 2.0f * d.x * d.y +
                   conic[2] * (d.y * d.y)) *
                  -0.5f;
    if (power > 0.0f || power < -thr) continue;
    // default rasterAlphaNode: opacity * exp(power) (GaussianContextNodes.ts:143)
    float alpha = gs_clampf(mean[3] * gs_expf(power), 0.0f, 0.99f);
    if (alpha < (float)(1.0 / 255.0)) continue;
    float contribution = a.t * alpha;
    if (s.cull_outputs) {
      std::atomic<uint32_t>& m = r.max_contribution[gid];
      uint32_t bits = gs_f2b(contribution), old = m.load(std::memory_order_relaxed);
      while (old < bits && !m.compare_exchange_weak(old, bits, std::memory_order_relaxed)) {
      }
    }
    a.c[0] = a.c[0] + color[0] * contribution;
    a.c[1] = a.c[1] + color[1] * contribution;
    a.c[2] = a.c[2] + color[2] * contribution;
    a.d = a.d + mean[2] * contribution;
    a.blended++;
    a.t = a.t * (1.0f - alpha);
    if (a.t < s.transmittance_threshold) break;
  }
  return a;
}

// pipeline/finalizeGaussianDepth.ts:18-68 for one pixel, with the default
// depth node of nodes/GaussianContextNodes.ts:125-131 (encoding 0); the
// other encodings are the custom nodes test/ScalarOutput.test.ts:61-96 sets.
float finalize_depth(float weighted, float t, float alpha_threshold, float invalid,
                     float near_plane, float far_plane, uint32_t encoding) {
  float alpha = 1.0f - t;
  float depth = 0.0f;
  if (alpha > 0.0f) depth = weighted / alpha;  // never 0 / 0
  auto finite = [](float v) { return v == v && gs_fabsf(v) <= kMaxF32; };
  bool scalar_valid = alpha > 0.0f && finite(weighted) && finite(depth) && depth > 0.0f;
  bool valid = scalar_valid && alpha > alpha_threshold;  // rasterOutputValid
  float out = invalid;
  if (valid) {
    float encoded;
    if (encoding == 0) {
      float view_z = -depth;
      encoded = gs_clampf((near_plane + view_z) * far_plane / ((far_plane - near_plane) * view_z),
                          0.0f, 1.0f);
    } else if (encoding == 1) {
      encoded = depth;
    } else if (encoding == 2) {
      volatile float zero = 0.0f;
      encoded = depth / zero;
    } else {
      encoded = alpha;
    }
    if (finite(encoded)) out = encoded;
  }
  return out;
}

// pipeline/TileRasterizer.ts:477-805 (direct and chunk targets) and :807-902
// (ordered chunk composite), :930-964 (storeFinalPixel) for one tile.
// `checked` and `blended` follow the reference's counters (:617, :750-768):
// each chunk task of a chunked tile blends its whole chunk from T = 1, so
// every chunk is counted in full, even those the composite's early stop
// never reads. Tiles above 1024 records or above the chunk size also add
// their checked pairs to `checked_heavy`.
void raster_tile(gs_oracle_result& r, uint32_t tile, uint64_t& checked, uint64_t& blended,
                 uint64_t& checked_heavy, uint32_t& chunk_tasks) {
  const gs_oracle_scene& s = r.scene;
  uint32_t tiles_x = r.counts.tiles_x;
  uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
  uint32_t begin = r.tile_offsets[tile];
  uint32_t count = r.tile_offsets[tile + 1] - begin;
  uint32_t cs = s.raster_chunk_size;
  bool chunked = cs != 0 && count > cs;
  bool heavy = chunked || count > 1024u;
  uint32_t chunks = chunked ? (count + cs - 1) / cs : 0;
  chunk_tasks += chunks;
  float bg_a = gs_clampf(s.background[3], 0.0f, 1.0f);
  uint32_t edge = tile_edge(s);
  uint64_t tile_checked = 0;
  for (uint32_t ly = 0; ly < edge; ly++)
    for (uint32_t lx = 0; lx < edge; lx++) {
      uint32_t x = tx * edge + lx, y = ty * edge + ly;
      if (x >= r.width || y >= r.height) continue;
      float px = (float)x + 0.5f, py = (float)y + 0.5f;
      float c[3] = {0.0f, 0.0f, 0.0f};
      float t = 1.0f, d = 0.0f;
      if (!chunked) {
        PixelAccum a = blend_range(r, begin, begin + count, px, py);
        c[0] = a.c[0]; c[1] = a.c[1]; c[2] = a.c[2];
        t = a.t; d = a.d;
        tile_checked += a.checked;
        blended += a.blended;
      } else {
        bool stopped = false;
        for (uint32_t k = 0; k < chunks; k++) {
          uint32_t b = k * cs, e = std::min(b + cs, count);
          PixelAccum a = blend_range(r, begin + b, begin + e, px, py);
          tile_checked += a.checked;
          blended += a.blended;
          if (stopped) continue;  // the composite has stopped reading partials
          c[0] = c[0] + a.c[0] * t;
          c[1] = c[1] + a.c[1] * t;
          c[2] = c[2] + a.c[2] * t;
          d = d + a.d * t;
          t = t * a.t;
          if (t < s.transmittance_threshold) stopped = true;
        }
      }
      size_t o = ((size_t)y * r.width + x) * 4;
      float out[4];
      for (int ch = 0; ch < 3; ch++) out[ch] = c[ch] + s.background[ch] * t * bg_a;
      out[3] = 1.0f - t * (1.0f - bg_a);
      for (int ch = 0; ch < 4; ch++) {
        r.image[o + ch] = out[ch];
        r.image_half[o + ch] = (uint16_t)gs_float_to_half(out[ch]);
      }
      r.depth_image[(size_t)y * r.width + x] = d;
      r.final_t[(size_t)y * r.width + x] = t;
      if (s.output_depth)
        r.depth_target[(size_t)y * r.width + x] =
            finalize_depth(d, t, s.depth_alpha_threshold, s.invalid_depth_value, s.near_plane,
                           s.far_plane, 0u);
    }
  checked += tile_checked;
  if (heavy) checked_heavy += tile_checked;
}

}  // namespace

extern "C" gs_oracle_result* gs_oracle_run(const gs_oracle_scene* sp) {
  if (!sp || sp->width < 1.0f || sp->height < 1.0f || sp->sh_degree > 3 ||
      sp->active_count > sp->count)
    return nullptr;
  gs_oracle_result* res = new gs_oracle_result();
  gs_oracle_result& r = *res;
  r.scene = *sp;
  const gs_oracle_scene& s = r.scene;
  uint32_t threads = s.threads ? s.threads : std::max(1u, std::thread::hardware_concurrency());
  std::memset(&r.counts, 0, sizeof(r.counts));
  r.width = gs_f2u(s.width);
  r.height = gs_f2u(s.height);
  // pipeline/TiledGaussianPipeline.ts:292-293
  uint32_t edge = tile_edge(s);
  if (edge != 8 && edge != 16 && edge != 32) {
    delete res;
    return nullptr;
  }
  uint32_t tiles_x = (r.width + edge - 1) / edge, tiles_y = (r.height + edge - 1) / edge;
  uint32_t tiles = tiles_x * tiles_y;
  r.counts.tiles_x = tiles_x;
  r.counts.tiles_y = tiles_y;
  uint32_t n = s.active_count;

  double t0 = now_s();
  r.projected_mean.assign((size_t)n * 4, 0.0f);
  r.projected_conic.assign((size_t)n * 4, 0.0f);
  r.projected_color.assign((size_t)n * 4, 0.0f);
  r.tile_counts.assign(n, 0);
  if (s.cull_outputs) {
    r.cull_features.assign((size_t)n * 16, 0.0f);
    r.max_contribution.reset(new std::atomic<uint32_t>[n ? n : 1]);
    for (uint32_t i = 0; i < n; i++) r.max_contribution[i].store(0u);
  }
  parallel_for(threads, n, [&](size_t b, size_t e, uint32_t) {
    for (size_t g = b; g < e; g++) project_one(s, tiles_x, tiles_y, (uint32_t)g, r);
  });
  double t1 = now_s();

  // kernels/scan.ts:78-83 + visibility.ts:9-61: visible flags, exclusive scan,
  // compaction in ascending ordinal order.
  r.visible_offsets.assign(n, 0);
  uint32_t v = 0;
  for (uint32_t g = 0; g < n; g++) {
    r.visible_offsets[g] = v;
    if (r.projected_mean[(size_t)g * 4 + 3] > 0.0f) v = sat_add(v, 1);
  }
  r.counts.visible = v;
  r.depth_records.assign(v, 0);
  for (uint32_t g = 0; g < n; g++) {
    if (!(r.projected_mean[(size_t)g * 4 + 3] > 0.0f)) continue;
    float depth = r.projected_mean[(size_t)g * 4 + 2];
    uint32_t key;
    if (s.depth_mode == 0) {
      key = gs_f2b(depth);
    } else {  // visibility.ts:40-44
      float q = gs_clampf((depth - s.near_plane) / (s.far_plane - s.near_plane), 0.0f, 1.0f);
      key = gs_f2u(__builtin_rintf(q * 65535.0f));
    }
    r.depth_records[r.visible_offsets[g]] = ((uint64_t)key << 32) | g;
  }
  double t2 = now_s();

  // RadixSorter "depth": stable LSD sort on the depth key (RadixSorter.ts:215-223).
  parallel_stable_sort(r.depth_records, threads, key_less);
  double t3 = now_s();

  // visibility.ts:63-76 gather + the intersections ExclusiveScanStage.
  r.ordered_counts.assign(v, 0);
  r.intersection_offsets.assign(v, 0);
  uint32_t total = 0;
  for (uint32_t k = 0; k < v; k++) {
    uint32_t gid = (uint32_t)(r.depth_records[k] & 0xffffffffu);
    r.ordered_counts[k] = r.tile_counts[gid];
    r.intersection_offsets[k] = total;
    total = sat_add(total, r.ordered_counts[k]);
  }
  // intersections.ts:10-41 (prepare_dispatch): overflow publishes nothing.
  r.counts.total = total;
  r.counts.overflow = total > s.capacity ? 1u : 0u;
  uint32_t count = r.counts.overflow ? 0u : total;
  r.counts.count = count;
  double t4 = now_s();

  r.emitted.assign(count, 0);
  if (count > 0) {
    std::vector<uint32_t> sentinel_counts(threads, 0);
    parallel_for(threads, v, [&](size_t b, size_t e, uint32_t th) {
      uint32_t local = 0;
      for (size_t k = b; k < e; k++)
        local += emit_one(r, tiles_x, tiles_y, (uint32_t)k, count, r.emitted);
      sentinel_counts[th] += local;
    });
    for (uint32_t c : sentinel_counts) r.counts.sentinels += c;
  }
  double t5 = now_s();

  // RadixSorter "tile": stable sort on the tile id.
  r.sorted = r.emitted;
  parallel_stable_sort(r.sorted, threads, key_less);
  double t6 = now_s();

  // kernels/tileOffsets.ts:4-128: first index per tile, sentinel tile
  // included, empty tiles filled by the suffix minimum.
  r.tile_offsets.assign((size_t)tiles + 1, 0xffffffffu);
  r.tile_offsets[tiles] = count;
  for (uint32_t i = 0; i < count; i++) {
    uint32_t tile = (uint32_t)(r.sorted[i] >> 32);
    if (tile <= tiles && (i == 0 || (uint32_t)(r.sorted[i - 1] >> 32) != tile))
      r.tile_offsets[tile] = i;
  }
  for (int64_t t = (int64_t)tiles - 1; t >= 0; t--)
    r.tile_offsets[t] = std::min(r.tile_offsets[t], r.tile_offsets[t + 1]);
  double t7 = now_s();

  // Raster. On overflow the reference dispatches no raster workgroups and
  // keeps the previous image (intersections.ts:114-134); the oracle then
  // reports an all-zero image and the caller checks the overflow flag.
  r.image.assign((size_t)r.width * r.height * 4, 0.0f);
  r.image_half.assign((size_t)r.width * r.height * 4, 0);
  r.depth_image.assign((size_t)r.width * r.height, 0.0f);
  r.final_t.assign((size_t)r.width * r.height, 1.0f);
  if (s.output_depth)
    r.depth_target.assign((size_t)r.width * r.height, s.invalid_depth_value);
  if (!r.counts.overflow) {
    std::vector<uint64_t> ch(threads, 0), bl(threads, 0), hv(threads, 0);
    std::vector<uint32_t> tasks(threads, 0);
    parallel_for(threads, tiles, [&](size_t b, size_t e, uint32_t th) {
      for (size_t t = b; t < e; t++)
        raster_tile(r, (uint32_t)t, ch[th], bl[th], hv[th], tasks[th]);
    });
    for (uint32_t i = 0; i < threads; i++) {
      r.counts.checked += ch[i];
      r.counts.blended += bl[i];
      r.counts.checked_heavy += hv[i];
      r.counts.chunk_tasks += tasks[i];
    }
  }
  double t8 = now_s();
  double st[8] = {t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4, t6 - t5, t7 - t6, t8 - t7};
  for (int i = 0; i < 8; i++) r.counts.stage_seconds[i] = st[i];
  return res;
}

extern "C" void gs_oracle_counts_of(const gs_oracle_result* r, gs_oracle_counts* out) {
  *out = r->counts;
}

namespace {
size_t copy_out(const void* src, size_t size, void* dst, size_t bytes) {
  if (!dst) return size;
  if (bytes != size) return 0;
  if (size) std::memcpy(dst, src, size);
  return size;
}
size_t copy_pairs(const std::vector<uint64_t>& v, void* dst, size_t bytes) {
  size_t size = v.size() * 8;
  if (!dst) return size;
  if (bytes != size) return 0;
  uint32_t* o = (uint32_t*)dst;
  for (size_t i = 0; i < v.size(); i++) {  // uvec2(first, second)
    o[2 * i] = (uint32_t)(v[i] >> 32);
    o[2 * i + 1] = (uint32_t)(v[i] & 0xffffffffu);
  }
  return size;
}
}  // namespace

extern "C" size_t gs_oracle_get(const gs_oracle_result* r, const char* name, void* dst,
                                size_t bytes) {
  std::string n(name);
#define GS_VEC(field) copy_out(r->field.data(), r->field.size() * sizeof(r->field[0]), dst, bytes)
  if (n == "projected_mean") return GS_VEC(projected_mean);
  if (n == "projected_conic") return GS_VEC(projected_conic);
  if (n == "projected_color") return GS_VEC(projected_color);
  if (n == "tile_counts") return GS_VEC(tile_counts);
  if (n == "visible_offsets") return GS_VEC(visible_offsets);
  if (n == "ordered_counts") return GS_VEC(ordered_counts);
  if (n == "intersection_offsets") return GS_VEC(intersection_offsets);
  if (n == "tile_offsets") return GS_VEC(tile_offsets);
  if (n == "image_f32") return GS_VEC(image);
  if (n == "image_f16") return GS_VEC(image_half);
  if (n == "depth_f32") return GS_VEC(depth_image);
  if (n == "final_t") return GS_VEC(final_t);
  if (n == "depth_target") return GS_VEC(depth_target);
  if (n == "cull_features") return GS_VEC(cull_features);
  if (n == "max_contribution") {
    size_t cnt = r->scene.cull_outputs ? r->scene.active_count : 0;
    std::vector<uint32_t> v(cnt);
    for (size_t i = 0; i < cnt; i++) v[i] = r->max_contribution[i].load();
    return copy_out(v.data(), v.size() * 4, dst, bytes);
  }
  if (n == "depth_records") return copy_pairs(r->depth_records, dst, bytes);
  if (n == "emitted_records") return copy_pairs(r->emitted, dst, bytes);
  if (n == "sorted_records") return copy_pairs(r->sorted, dst, bytes);
#undef GS_VEC
  return 0;
}

extern "C" void gs_oracle_free(gs_oracle_result* r) { delete r; }

// Numerators b (of the packed formats' ranges) for which the kernels'
// gs_div_small(b, d, 1/d) differs from the IEEE quotient b / d, summed over
// d = 127 (b in [-128, 127]), 1023 ([0, 1023]) and 4095 ([0, 4095]).
extern "C" uint32_t gs_oracle_div_small_mismatches() {
  struct Range {
    float d;
    int lo, hi;
  } ranges[] = {{127.0f, -128, 127}, {1023.0f, 0, 1023}, {4095.0f, 0, 4095}};
  uint32_t bad = 0;
  for (const Range& r : ranges) {
    volatile float d = r.d;  // the quotient below is evaluated at run time
    for (int b = r.lo; b <= r.hi; b++)
      if (gs_f2b(gs_div_small((float)b, r.d, 1.0f / r.d)) != gs_f2b((float)b / d)) bad++;
  }
  return bad;
}

// FP64 blend: the same per-tile lists and projected f32 values, blended in
// double precision with one uninterrupted front-to-back stream per pixel.
extern "C" void gs_oracle_blend_f64(const gs_oracle_result* rp, double* out) {
  const gs_oracle_result& r = *rp;
  const gs_oracle_scene& s = r.scene;
  uint32_t tiles_x = r.counts.tiles_x;
  uint32_t tiles = tiles_x * r.counts.tiles_y;
  double bg_a = std::min(std::max((double)s.background[3], 0.0), 1.0);
  uint32_t threads = s.threads ? s.threads : std::max(1u, std::thread::hardware_concurrency());
  parallel_for(threads, tiles, [&](size_t b, size_t e, uint32_t) {
    for (size_t tile = b; tile < e; tile++) {
      uint32_t tx = (uint32_t)tile % tiles_x, ty = (uint32_t)tile / tiles_x;
      uint32_t begin = r.tile_offsets[tile], end = r.tile_offsets[tile + 1];
      uint32_t edge = tile_edge(s);
      for (uint32_t ly = 0; ly < edge; ly++)
        for (uint32_t lx = 0; lx < edge; lx++) {
          uint32_t x = tx * edge + lx, y = ty * edge + ly;
          if (x >= r.width || y >= r.height) continue;
          double px = x + 0.5, py = y + 0.5;
          double c[3] = {0, 0, 0}, t = 1.0;
          if (!r.counts.overflow)
            for (uint32_t i = begin; i < end; i++) {
              uint32_t gid = (uint32_t)(r.sorted[i] & 0xffffffffu);
              const float* mean = &r.projected_mean[(size_t)gid * 4];
              const float* conic = &r.projected_conic[(size_t)gid * 4];
              const float* color = &r.projected_color[(size_t)gid * 4];
              double dx = px - mean[0], dy = py - mean[1];
              if (s.boundary_axis != 0) {
                float d2[2];
                float pp[2] = {(float)px, (float)py};
                gs_oracle_boundary_delta(pp, mean, conic, &s.width, s.boundary_axis, d2);
                dx = d2[0];
                dy = d2[1];
              }
              double power =
                  -0.5 * (conic[0] * dx * dx + 2.0 * conic[1] * dx * dy + conic[2] * dy * dy);
              double thr = std::log((double)mean[3] * 255.0);
              if (power > 0.0 || power < -thr) continue;
              double alpha = std::min(std::max(mean[3] * std::exp(power), 0.0), 0.99);
              if (alpha < 1.0 / 255.0) continue;
              for (int ch = 0; ch < 3; ch++) c[ch] += color[ch] * t * alpha;
              t *= 1.0 - alpha;
              if (t < s.transmittance_threshold) break;
            }
          size_t o = ((size_t)y * r.width + x) * 4;
          for (int ch = 0; ch < 3; ch++) out[o + ch] = c[ch] + s.background[ch] * t * bg_a;
          out[o + 3] = 1.0 - t * (1.0 - bg_a);
        }
    }
  });
}

// ------------------------------------------------ exported helper tests ----

extern "C" void gs_oracle_view_covariance(const float scale[3], const float rotation[4],
                                          const float model_view[16], float out[9]) {
  M3 c = view_covariance(V3{scale[0], scale[1], scale[2]},
                         V4{rotation[0], rotation[1], rotation[2], rotation[3]},
                         load_m4(model_view));
  for (int j = 0; j < 3; j++) {
    out[j * 3] = c.c[j].x;
    out[j * 3 + 1] = c.c[j].y;
    out[j * 3 + 2] = c.c[j].z;
  }
}

static void store_m4(const M4& m, float out[16]) {
  for (int c = 0; c < 4; c++) {
    out[c * 4] = m.c[c].x;
    out[c * 4 + 1] = m.c[c].y;
    out[c * 4 + 2] = m.c[c].z;
    out[c * 4 + 3] = m.c[c].w;
  }
}

extern "C" void gs_oracle_perspective_projection(const float view[3], const float cov_view[9],
                                                 const float projection[16],
                                                 const float viewport[4],
                                                 float filter_variance, float out[16]) {
  M3 c;
  for (int j = 0; j < 3; j++) c.c[j] = {cov_view[j * 3], cov_view[j * 3 + 1], cov_view[j * 3 + 2]};
  store_m4(perspective_projection(V3{view[0], view[1], view[2]}, c, load_m4(projection),
                                  V4{viewport[0], viewport[1], viewport[2], viewport[3]},
                                  filter_variance),
           out);
}

extern "C" void gs_oracle_projection_result(const float center[2], float depth,
                                            const float covariance[3], float filter_variance,
                                            float out[16]) {
  store_m4(projection_result(V2{center[0], center[1]}, depth,
                             V3{covariance[0], covariance[1], covariance[2]}, filter_variance),
           out);
}

extern "C" int gs_oracle_projection_is_valid(const float p[16]) {
  return projection_is_valid(load_m4(p)) ? 1 : 0;
}

extern "C" int gs_oracle_tile_hit(const float center[2], const float conic[3],
                                  float power_threshold, const float radius[2], int tile_x,
                                  int tile_y, const float viewport[2], uint32_t axis) {
  return tile_hit(V2{center[0], center[1]}, V3{conic[0], conic[1], conic[2]},
                  power_threshold, V2{radius[0], radius[1]}, tile_x, tile_y,
                  V2{viewport[0], viewport[1]}, axis)
             ? 1
             : 0;
}

extern "C" uint32_t gs_oracle_count_tiles(const float center[2], const float conic[3],
                                          float power_threshold, const float radius[2],
                                          int minx, int miny, int maxx, int maxy,
                                          const float viewport[2], uint32_t axis) {
  return count_tiles(V2{center[0], center[1]}, V3{conic[0], conic[1], conic[2]},
                     power_threshold, V2{radius[0], radius[1]}, minx, miny, maxx, maxy,
                     V2{viewport[0], viewport[1]}, axis);
}

extern "C" int gs_oracle_subpixel_has_sample(const float center[2], const float conic[3],
                                             float power_threshold, const float extent[2],
                                             uint32_t vx, uint32_t vy, uint32_t axis) {
  return subpixel_has_sample(V2{center[0], center[1]}, V3{conic[0], conic[1], conic[2]},
                             power_threshold, V2{extent[0], extent[1]}, vx, vy, axis)
             ? 1
             : 0;
}

extern "C" void gs_oracle_boundary_delta(const float pixel[2], const float center[2],
                                         const float conic[3], const float viewport[2],
                                         uint32_t axis, float out[2]) {
  V2 d = boundary_delta(V2{pixel[0], pixel[1]}, V2{center[0], center[1]},
                        V3{conic[0], conic[1], conic[2]}, V2{viewport[0], viewport[1]}, axis);
  out[0] = d.x;
  out[1] = d.y;
}

extern "C" void gs_oracle_decode_scale_opacity(uint32_t x, uint32_t y, float out[4]) {
  V4 v = decode_scale_opacity(x, y);
  out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
}

extern "C" void gs_oracle_decode_rotation(uint32_t packed, float out[4]) {
  V4 v = decode_rotation(packed);
  out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
}

extern "C" void gs_oracle_decode_sh_rgb8e8(uint32_t packed, float out[3]) {
  V3 v = decode_sh(packed);
  out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

extern "C" void gs_oracle_evaluate_sh(const uint32_t* words, uint32_t degree, const float dir[3],
                                      float out[3]) {
  V3 v = evaluate_sh(words, 0u, degree, V3{dir[0], dir[1], dir[2]});
  out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

extern "C" float gs_oracle_finalize_depth(float weighted_depth, float transmittance,
                                          float depth_alpha_threshold, float invalid_value,
                                          float near_plane, float far_plane, uint32_t encoding) {
  return finalize_depth(weighted_depth, transmittance, depth_alpha_threshold, invalid_value,
                        near_plane, far_plane, encoding);
}

// Elementwise evaluation of the shared elementary functions, for the
// accuracy tests against libm. fn: 0 exp, 1 log, 2 sin, 3 cos, 4 atan,
// 5 atan2 (second operand in b), 6 half->float (input bits), 7 float->half,
// 8 log of positive normals, 9 exp of non-positive arguments.
extern "C" void gs_oracle_math(int fn, const float* a, const float* b, float* out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    float x = a[i];
    switch (fn) {
      case 0: out[i] = gs_expf(x); break;
      case 1: out[i] = gs_logf(x); break;
      case 2: out[i] = gs_sinf(x); break;
      case 3: out[i] = gs_cosf(x); break;
      case 4: out[i] = gs_atanf(x); break;
      case 5: out[i] = gs_atan2f(x, b[i]); break;
      case 6: out[i] = gs_half_to_float(gs_f2b(x) & 0xffffu); break;
      case 7: out[i] = gs_b2f(gs_float_to_half(x)); break;
      case 8: out[i] = gs_logf_pos(x); break;
      case 9: out[i] = gs_expf_nonpos(x); break;
      default: out[i] = 0.0f;
    }
  }
}
