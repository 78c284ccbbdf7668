// gs_project.h - per-Gaussian projection math for the gfx950 kernels.
//
// Device functions for the projection stage, in the operation order of the
// reference WGSL so that tile counts, radii, keys and conics are bit-exact
// against the CPU restatement: every binary32 operation is spelled out,
// left-to-right sums, no contraction (compiled with -ffp-contract=off).
#ifndef GS_PROJECT_H
#define GS_PROJECT_H

#include "gs_internal.h"
#include "gs_math.h"

namespace gs {

#define GS_DEV __device__ __forceinline__

constexpr float kMaxF32 = 3.402823e+38f;  // projectionHelpers.ts:44,100

// m[col][row], as WGSL matrices.
GS_DEV void mat3_mul(const float a[3][3], const float b[3][3], float o[3][3]) {
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      float s = a[0][i] * b[j][0];
      s = s + a[1][i] * b[j][1];
      s = s + a[2][i] * b[j][2];
      o[j][i] = s;
    }
}

GS_DEV void mat3_transpose(const float a[3][3], float o[3][3]) {
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++) o[j][i] = a[i][j];
}

GS_DEV void mat3_vec(const float m[3][3], const float v[3], float o[3]) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float s = m[0][i] * v[0];
    s = s + m[1][i] * v[1];
    s = s + m[2][i] * v[2];
    o[i] = s;
  }
}

GS_DEV float dot3(const float a[3], const float b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// kernels/compactGaussians.ts:10-19
GS_DEV void decode_rotation(uint32_t packed, float q[4]) {
  float px = gs_div_small((float)(packed & 1023u), 1023.0f, 1.0f / 1023.0f) * 2.0f - 1.0f;
  float py = gs_div_small((float)((packed >> 10) & 1023u), 1023.0f, 1.0f / 1023.0f) * 2.0f - 1.0f;
  float z = 1.0f - gs_fabsf(px) - gs_fabsf(py);
  float t = gs_maxf(-z, 0.0f);
  float x = px >= 0.0f ? px + -t : px + t;
  float y = py >= 0.0f ? py + -t : py + t;
  float len = gs_sqrtf((x * x + y * y) + z * z);
  float ax = x / len, ay = y / len, az = z / len;
  float half_angle = gs_div_small((float)(packed >> 20), 4095.0f, 1.0f / 4095.0f) * 1.5707963267948966f;
  float s = gs_sinf(half_angle);
  q[0] = ax * s;
  q[1] = ay * s;
  q[2] = az * s;
  q[3] = gs_cosf(half_angle);
}

// kernels/projectionHelpers.ts:123-127
GS_DEV void decode_sh(uint32_t packed, float out[3]) {
  float scale = gs_exp2i((int)((packed >> 24) & 255u) - 127);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    int8_t b = (int8_t)((packed >> (8 * i)) & 0xffu);
    out[i] = gs_maxf(gs_div_small((float)b, 127.0f, 1.0f / 127.0f), -1.0f) * scale;
  }
}

// kernels/projectionHelpers.ts:4-35
GS_DEV void view_covariance(const float scale_in[3], const float rot[4],
                            const float mv[4][4], float out[3][3]) {
  float s0 = gs_maxf(scale_in[0], 1e-7f);
  float s1 = gs_maxf(scale_in[1], 1e-7f);
  float s2 = gs_maxf(scale_in[2], 1e-7f);
  float qd = ((rot[0] * rot[0] + rot[1] * rot[1]) + rot[2] * rot[2]) + rot[3] * rot[3];
  float ql = gs_sqrtf(qd);
  float qx = rot[0] / ql, qy = rot[1] / ql, qz = rot[2] / ql, qw = rot[3] / ql;
  float xx = qx * qx, yy = qy * qy, zz = qz * qz;
  float xy = qx * qy, xz = qx * qz, yz = qy * qz;
  float xw = qx * qw, yw = qy * qw, zw = qz * qw;
  float r[3][3] = {
      {1.0f - 2.0f * (yy + zz), 2.0f * (xy + zw), 2.0f * (xz - yw)},
      {2.0f * (xy - zw), 1.0f - 2.0f * (xx + zz), 2.0f * (yz + xw)},
      {2.0f * (xz + yw), 2.0f * (yz - xw), 1.0f - 2.0f * (xx + yy)}};
  float d[3][3] = {{s0 * s0, 0.0f, 0.0f}, {0.0f, s1 * s1, 0.0f}, {0.0f, 0.0f, s2 * s2}};
  float rd[3][3], rt[3][3], local[3][3];
  mat3_mul(r, d, rd);
  mat3_transpose(r, rt);
  mat3_mul(rd, rt, local);
  float l[3][3] = {{mv[0][0], mv[0][1], mv[0][2]},
                   {mv[1][0], mv[1][1], mv[1][2]},
                   {mv[2][0], mv[2][1], mv[2][2]}};
  float lc[3][3], lt[3][3];
  mat3_mul(l, local, lc);
  mat3_transpose(l, lt);
  mat3_mul(lc, lt, out);
}

// Result of the default projection root (projectionHelpers.ts:38-96).
struct Projected {
  bool valid;
  float cx, cy, depth;
  float conic[3];
  float det;
  float cov[3];
  float original_det;
};

// kernels/projectionHelpers.ts:38-78 followed by the validity test :98-116.
GS_DEV void projection_result(float cx, float cy, float depth, float c0, float c1,
                              float c2, float fv, Projected& p) {
  p.valid = false;
  float original_det = gs_maxf(c0 * c2 - c1 * c1, 0.0f);
  float s00 = c0 + fv, s01 = c1, s11 = c2 + fv;
  bool finite = s00 == s00 && gs_fabsf(s00) <= kMaxF32 && s01 == s01 &&
                gs_fabsf(s01) <= kMaxF32 && s11 == s11 && gs_fabsf(s11) <= kMaxF32;
  if (!finite) return;
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
  if (det <= 1e-8f) return;
  float inv = 1.0f / det;
  float k0 = s11 * inv, k1 = -s01 * inv, k2 = s00 * inv;
  // gaussian_projection_is_valid: rows 0..2 finite, then the range tests.
  float rows[12] = {cx, cy, depth, 1.0f, k0, k1, k2, det, s00, s01, s11, original_det};
#pragma unroll
  for (int i = 0; i < 12; i++)
    if (!(rows[i] == rows[i]) || !(gs_fabsf(rows[i]) <= kMaxF32)) return;
  float vdet = s00 * s11 - s01 * s01;
  if (!(depth > 0.0f && vdet > 1e-8f && s00 > 0.0f && s11 > 0.0f && s00 <= 1e4f &&
        s11 <= 1e4f && k0 > 0.0f && k2 > 0.0f && original_det >= 0.0f))
    return;
  float e0 = s11 / vdet, e1 = -s01 / vdet, e2 = s00 / vdet;
  if (!(gs_fabsf(det - vdet) <= gs_maxf(1e-6f, vdet * 1e-3f))) return;
  if (!(gs_fabsf(k0 - e0) <= gs_maxf(1e-6f, gs_fabsf(e0) * 1e-3f) &&
        gs_fabsf(k1 - e1) <= gs_maxf(1e-6f, gs_fabsf(e1) * 1e-3f) &&
        gs_fabsf(k2 - e2) <= gs_maxf(1e-6f, gs_fabsf(e2) * 1e-3f)))
    return;
  p.valid = true;
  p.cx = cx;
  p.cy = cy;
  p.depth = depth;
  p.conic[0] = k0;
  p.conic[1] = k1;
  p.conic[2] = k2;
  p.det = det;
  p.cov[0] = s00;
  p.cov[1] = s01;
  p.cov[2] = s11;
  p.original_det = original_det;
}

// kernels/projectionHelpers.ts:80-96
GS_DEV void perspective_projection(const float view[3], const float cv[3][3],
                                   const FrameParams& fp, const FrameCamera& cam,
                                   Projected& p) {
  p.valid = false;
  float depth = -view[2];
  if (!(depth > cam.near_plane && depth < cam.far_plane)) return;
  const float* P = cam.projection;
  float clip[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float s = P[0 * 4 + i] * view[0];
    s = s + P[1 * 4 + i] * view[1];
    s = s + P[2 * 4 + i] * view[2];
    s = s + P[3 * 4 + i] * 1.0f;
    clip[i] = s;
  }
  if (clip[3] <= 0.0f) return;
  float ndcx = clip[0] / clip[3], ndcy = clip[1] / clip[3];
  float cx = (ndcx * 0.5f + 0.5f) * fp.viewport[0];
  float cy = (0.5f - ndcy * 0.5f) * fp.viewport[1];
  float fx = 0.5f * fp.viewport[0] * P[0];
  float fy = 0.5f * fp.viewport[1] * P[5];
  float inv = 1.0f / depth;
  float j0[3] = {fx * inv, 0.0f, fx * view[0] * inv * inv};
  float j1[3] = {0.0f, -fy * inv, -fy * view[1] * inv * inv};
  float a[3], b[3];
  mat3_vec(cv, j0, a);
  mat3_vec(cv, j1, b);
  projection_result(cx, cy, depth, dot3(j0, a), dot3(j0, b), dot3(j1, b),
                    fp.filter_variance, p);
}

// kernels/projectionHelpers.ts:185-221
GS_DEV bool subpixel_has_sample(float cx, float cy, const float conic[3], float thr,
                                float ex, float ey, uint32_t vx, uint32_t vy,
                                uint32_t axis) {
  if (ex * 2.0f > 1.0f || ey * 2.0f > 1.0f) return true;
  float ccx = cx, ccy = cy;
  if (axis == 1) ccx = ccx - gs_floorf(cx / (float)vx) * (float)vx;
  if (axis == 2) ccy = ccy - gs_floorf(cy / (float)vy) * (float)vy;
  int minx = gs_f2i(gs_ceilf(ccx - ex - 0.5f));
  int miny = gs_f2i(gs_ceilf(ccy - ey - 0.5f));
  int maxx = gs_f2i(gs_floorf(ccx + ex - 0.5f));
  int maxy = gs_f2i(gs_floorf(ccy + ey - 0.5f));
  int pminx = axis == 1 ? minx : max(minx, 0);
  int pminy = axis == 2 ? miny : max(miny, 0);
  int pmaxx = axis == 1 ? maxx : min(maxx, (int)vx - 1);
  int pmaxy = axis == 2 ? maxy : min(maxy, (int)vy - 1);
  for (int py = pminy; py <= pmaxy; py++)
    for (int px = pminx; px <= pmaxx; px++) {
      float dx = ((float)px + 0.5f) - ccx;
      float dy = ((float)py + 0.5f) - ccy;
      float sigma =
          0.5f * (conic[0] * dx * dx + 2.0f * conic[1] * dx * dy + conic[2] * dy * dy);
      if (sigma <= thr) return true;
    }
  return false;
}

// kernels/tileContribution.ts:271-313 inside :339-367. One tile, one image;
// ts is the tile edge in pixels (a power of two, so the products are exact).
GS_DEV bool tile_rect_hit(float icx, float icy, const float conic[3], float thr, int tx,
                          int ty, float ts) {
  float rminx = (float)tx * ts, rminy = (float)ty * ts;
  float rmaxx = rminx + ts, rmaxy = rminy + ts;
  bool x_left = icx < rminx, x_right = icx > rmaxx;
  bool in_x = !(x_left || x_right);
  bool y_above = icy < rminy, y_below = icy > rmaxy;
  bool in_y = !(y_above || y_below);
  if (in_x && in_y) return true;
  float cx = x_left ? rminx : rmaxx, cy = y_above ? rminy : rmaxy;
  float ex = x_left ? ts : -ts, ey = y_above ? ts : -ts;
  float dx = icx - cx, dy = icy - cy;
  float tx_raw = (ex * conic[0] * dx + ex * conic[1] * dy) / (ex * conic[0] * ex);
  float ty_raw = (ey * conic[1] * dx + ey * conic[2] * dy) / (ey * conic[2] * ey);
  float tt_x = in_y ? 0.0f : gs_clampf(tx_raw, 0.0f, 1.0f);
  float tt_y = in_x ? 0.0f : gs_clampf(ty_raw, 0.0f, 1.0f);
  float clx = cx + tt_x * ex, cly = cy + tt_y * ey;
  float ddx = clx - icx, ddy = cly - icy;
  float sigma =
      0.5f * (conic[0] * ddx * ddx + conic[2] * ddy * ddy) + conic[1] * ddx * ddy;
  return sigma <= thr;
}

GS_DEV bool tile_hit(float cx, float cy, const float conic[3], float thr, float rx,
                     float ry, int tx, int ty, float vw, float vh, uint32_t axis, float ts) {
  if (axis == 0) return tile_rect_hit(cx, cy, conic, thr, tx, ty, ts);
  float period = axis == 1 ? vw : vh;
  float c = axis == 1 ? cx : cy;
  float r = axis == 1 ? rx : ry;
  float canon = c - gs_floorf(c / period) * period;
  float tile_start = (float)(axis == 1 ? tx : ty) * ts;
  int first = gs_f2i(gs_ceilf((tile_start - r - canon) / period));
  int last = gs_f2i(gs_floorf((tile_start + ts + r - canon) / period));
  for (int image = first; image <= last; image++) {
    float icx = axis == 1 ? canon + (float)image * vw : cx;
    float icy = axis == 2 ? canon + (float)image * vh : cy;
    if (tile_rect_hit(icx, icy, conic, thr, tx, ty, ts)) return true;
  }
  return false;
}

// kernels/tileContribution.ts:321-337
GS_DEV void boundary_delta(float px, float py, float cx, float cy, float c0, float c1,
                           float c2, float vw, float vh, uint32_t axis, float& dx,
                           float& dy) {
  float ccx = cx, ccy = cy;
  if (axis == 1) ccx = ccx - gs_floorf(cx / vw) * vw;
  if (axis == 2) ccy = ccy - gs_floorf(cy / vh) * vh;
  dx = px - ccx;
  dy = py - ccy;
  if (axis == 1) {
    float optimum = -c1 * dy / c0;
    dx = dx - gs_floorf((dx - optimum) / vw + 0.5f) * vw;
  }
  if (axis == 2) {
    float optimum = -c1 * dx / c2;
    dy = dy - gs_floorf((dy - optimum) / vh + 0.5f) * vh;
  }
}

}  // namespace gs

#endif
