// gs_kernels.hip - gfx950 kernels of the tiled Gaussian-splatting frame.
//
// Stage map (reference file:line -> kernel here):
//   projection         ProjectionStage.ts:385-649         -> k_project
//   visible scan+compact scan.ts:78-83, visibility.ts:37-61 -> k_scan_reduce /
//                      k_scan_partials / k_scan_downsweep (fused compaction)
//   depth radix sort   RadixSorter.ts:215-313, radix.ts    -> k_radix_hist /
//                      plain scan / k_radix_scatter (8-bit digits)
//   ordered counts + intersection scan visibility.ts:63-76, scan.ts
//                      -> the same scan kernels with a gathering load
//   prepare dispatch / output state intersections.ts:10-41, 114-134
//                      -> folded into k_scan_partials (kScanIntersections)
//   emission           intersections.ts:43-111             -> k_emit
//   tile radix sort    TiledGaussianPipeline.ts:442-446    -> radix kernels
//   tile offsets       tileOffsets.ts:4-128                -> k_tile_offsets
//   raster + chunks + composite TileRasterizer.ts:477-964  -> k_raster
//
// Every size that the reference reads through indirect dispatch (V, K) is
// read from DeviceState on the device, so the frame can be replayed as a
// captured graph; grids are sized by capacity and idle blocks exit at once.
#include <algorithm>
#include <utility>

#include "gs_cull.h"
#include "gs_internal.h"
#include "gs_project.h"

namespace gs {

// ------------------------------------------------------------ utilities ----

GS_DEV uint32_t sat_add(uint32_t a, uint32_t b) {
  uint32_t s = a + b;
  return s < a ? 0xffffffffu : s;
}

GS_DEV uint32_t lane_id() { return __lane_id(); }

GS_DEV uint32_t eval_len(const LenSpec& l) {
  if (!l.ptr) return l.constant;
  uint32_t v = *l.ptr;
  if (l.log2_div) v = (uint32_t)(((uint64_t)v + (1u << l.log2_div) - 1u) >> l.log2_div);
  return v * l.mul;
}

// Inclusive saturating scan across one wave64.
GS_DEV uint32_t wave_inclusive_scan(uint32_t v) {
  uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v = sat_add(v, o);
  }
  return v;
}

// Exclusive saturating scan across the 256-thread block; *total gets the sum.
GS_DEV uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* lds4) {
  uint32_t incl = wave_inclusive_scan(v);
  uint32_t prev = __shfl_up(incl, 1, 64);
  uint32_t excl = lane_id() == 0 ? 0u : prev;
  uint32_t wave = threadIdx.x >> 6;
  if (lane_id() == 63) lds4[wave] = incl;
  __syncthreads();
  uint32_t w0 = lds4[0], w1 = lds4[1], w2 = lds4[2], w3 = lds4[3];
  uint32_t before = 0;
  if (wave > 0) before = w0;
  if (wave > 1) before = sat_add(before, w1);
  if (wave > 2) before = sat_add(before, w2);
  *total = sat_add(sat_add(w0, w1), sat_add(w2, w3));
  __syncthreads();
  return sat_add(before, excl);
}

// -------------------------------------------------------------- project ----

// Outcome of projecting one Gaussian: nothing to draw, culled by the
// subpixel sample test (its centre and depth are still published), or
// visible with `count` tiles.
enum ProjKind : int { kProjNone = 0, kProjSubpixel = 1, kProjVisible = 2 };
struct ProjOut {
  int kind;
  float4 mean, conic, color;
  float thr;  // ln(255 opacity)
  float3 support;  // the raster's support bound terms (support_terms)
  uint32_t count;
  uint2 hits;  // tile-hit record (kHitMasked)
};

// A visible Gaussian's tile hits, 8 bytes, written by the projection and
// gathered into depth order by the tile-count scan, so that emission reads
// it in sequence. A box of at most 32 tiles (in a grid of at most 2^26
// tiles): x = the row-major hit mask over the box, y = kHitMasked |
// (width - 1) << 26 | the index of the box's first tile; the hit count is
// the mask's popcount. Otherwise x = the hit count and y = 0, and emission
// tests the tiles again from the projected mean and conic.
constexpr uint32_t kHitMasked = 1u << 31;
constexpr uint32_t kHitOriginBits = 26;
GS_DEV uint32_t hit_count(uint2 h) {
  return (h.y & kHitMasked) ? (uint32_t)__popc(h.x) : h.x;
}

// Projection modes: the plain path, the learned cull, and feature
// collection for its training set.
enum ProjMode : int { kModePlain = 0, kModeCull = 1, kModeCollect = 2 };

// A Gaussian projected up to its tile footprint (ProjectionStage.ts:385-
// 600): kind is kProjNone, kProjSubpixel (mean published with w = -1) or
// kProjVisible for a candidate whose exact tile count is still to be taken.
struct ProjGeom {
  int kind;
  uint32_t slot;
  float cx, cy, depth, opacity;
  float conic[3];
  float thr, rx, ry;
  float dir[3];    // normalised object-space offset from the camera (SH)
  float local[3];  // object-space mean (learned-cull features)
};

GS_DEV ProjGeom project_geometry(const FrameParams& fp, const SceneView& sc, uint32_t gid) {
  ProjGeom o;
  o.kind = kProjNone;
  uint32_t slot = fp.has_active_slots ? sc.active_slots[gid] : gid;
  o.slot = slot;
  float4 mean = sc.means[slot];
  if (mean.w < 0.0f) return o;
  uint32_t object = gs_f2u(mean.w);
  if (object >= fp.object_count) return o;
  const float* frame = sc.objects + (size_t)object * 40;
  if (!(frame[36] > 0.0f)) return o;
  uint2 so = sc.scales_opacity[slot];
  float scale[3] = {gs_expf(gs_half_to_float(so.x & 0xffffu)),
                    gs_expf(gs_half_to_float(so.x >> 16)),
                    gs_expf(gs_half_to_float(so.y & 0xffffu))};
  float base_opacity = gs_half_to_float(so.y >> 16);
  float rot[4];
  decode_rotation(sc.rotations[slot], rot);
  float mv[4][4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) mv[c][r] = frame[16 + c * 4 + r];
  float cam[3] = {frame[32], frame[33], frame[34]};
  float local[3] = {mean.x, mean.y, mean.z};
  float dir[3] = {local[0] - cam[0], local[1] - cam[1], local[2] - cam[2]};
  {
    float len = gs_sqrtf(dot3(dir, dir));
    dir[0] = dir[0] / len;
    dir[1] = dir[1] / len;
    dir[2] = dir[2] / len;
  }
  float view[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float s = mv[0][i] * local[0];
    s = s + mv[1][i] * local[1];
    s = s + mv[2][i] * local[2];
    s = s + mv[3][i] * 1.0f;
    view[i] = s;
  }
  float cv[3][3];
  view_covariance(scale, rot, mv, cv);
  Projected p;
  perspective_projection(view, cv, fp, *fp.camera, p);
  if (!p.valid) return o;
  float opacity = gs_clampf(base_opacity, 0.0f, 1.0f);
  if (fp.antialias_compensated)
    opacity = opacity * gs_sqrtf(gs_clampf(p.original_det / p.det, 0.0f, 1.0f));
  if (opacity < (float)(1.0 / 255.0) || opacity != opacity) return o;
  float thr = gs_logf(opacity * 255.0f);
  float ext_x = gs_sqrtf(thr * 2.0f * gs_clampf(p.cov[0], 1e-12f, 1e4f));
  float ext_y = gs_sqrtf(thr * 2.0f * gs_clampf(p.cov[2], 1e-12f, 1e4f));
  float rx = gs_ceilf(ext_x), ry = gs_ceilf(ext_y);
  if (rx <= 0.0f || ry <= 0.0f) return o;
  float bminx = p.cx - rx, bminy = p.cy - ry, bmaxx = p.cx + rx, bmaxy = p.cy + ry;
  float vw = fp.viewport[0], vh = fp.viewport[1];
  uint32_t axis = fp.boundary_axis;
  if ((axis != 1 && (bmaxx < 0.0f || bminx >= vw)) ||
      (axis != 2 && (bmaxy < 0.0f || bminy >= vh)))
    return o;
  o.cx = p.cx;
  o.cy = p.cy;
  o.depth = p.depth;
  if (fp.subpixel_culling &&
      !subpixel_has_sample(p.cx, p.cy, p.conic, thr, ext_x, ext_y, gs_f2u(vw), gs_f2u(vh),
                           axis)) {
    o.kind = kProjSubpixel;
    return o;
  }
  o.kind = kProjVisible;
  o.opacity = opacity;
  o.conic[0] = p.conic[0];
  o.conic[1] = p.conic[1];
  o.conic[2] = p.conic[2];
  o.thr = thr;
  o.rx = rx;
  o.ry = ry;
  o.dir[0] = dir[0];
  o.dir[1] = dir[1];
  o.dir[2] = dir[2];
  o.local[0] = local[0];
  o.local[1] = local[1];
  o.local[2] = local[2];
  return o;
}

// The exact tile count, hit record and SH colour of a candidate
// (ProjectionStage.ts:600-649): visible if it hits a tile.
// The per-Gaussian terms of the raster's support bound (record_rect): with
// det = c0 c2 - c1^2, two_t = 2 thr (1 + 1e-6), g = 3e-7 and, for the rows,
// n = det - g (c2 + |c1|) c0, the half-height h of the rows a block may
// need obeys h^2 = two_t c0 / n + g (c0 + |c1|) c0 / n * O^2, O bounding
// |dx| over the block's columns; columns likewise with c0 and c2 exchanged.
// Returned: (rows' constant term, columns' constant term, the larger of the
// two O^2 coefficients), each rounded once from binary64; an axis without a
// bound (not positive definite, infinite cutoff, n <= 0) gets +inf, which
// record_rect reads as "no narrowing".
GS_DEV float3 support_terms(const float conic[3], float thr) {
  const double g = 3.0e-7;
  const double c0 = conic[0], c1a = fabs((double)conic[1]), c2 = conic[2];
  const double det = c0 * c2 - (double)conic[1] * (double)conic[1];
  const double T = thr;
  const float inf = gs_b2f(0x7f800000u);
  float3 t = make_float3(inf, inf, 0.0f);
  if (!(det > 0.0 && c0 > 0.0 && c2 > 0.0 && T >= 0.0 && T < 1e30)) return t;
  const double two_t = 2.0 * T * (1.0 + 1e-6);
  const double n_rows = det - g * (c2 + c1a) * c0, n_cols = det - g * (c0 + c1a) * c2;
  double b = 0.0;
  if (n_rows > 0.0) {
    t.x = (float)(two_t * c0 / n_rows);
    b = fmax(b, g * (c0 + c1a) * c0 / n_rows);
  }
  if (n_cols > 0.0) {
    t.y = (float)(two_t * c2 / n_cols);
    b = fmax(b, g * (c2 + c1a) * c2 / n_cols);
  }
  // Rounded up so that the binary32 coefficient is never below the bound.
  float bf = (float)b;
  if ((double)bf < b) bf = gs_b2f(gs_f2b(bf) + 1u);
  t.z = bf;
  return t;
}

// A visible Gaussian's bounding box in tiles (ProjectionStage.ts:590-618).
struct TileBox {
  int tmin_x, tmin_y, tmax_x, tmax_y;
  GS_DEV uint32_t width() const { return (uint32_t)(tmax_x - tmin_x + 1); }
  GS_DEV uint32_t height() const { return (uint32_t)(tmax_y - tmin_y + 1); }
};
constexpr uint32_t kMaskTiles = 32;  // boxes up to this many tiles keep a hit mask

GS_DEV TileBox tile_box(const FrameParams& fp, const ProjGeom& g) {
  float bminx = g.cx - g.rx, bminy = g.cy - g.ry, bmaxx = g.cx + g.rx, bmaxy = g.cy + g.ry;
  float max_tx = (float)((int)fp.tiles_x - 1), max_ty = (float)((int)fp.tiles_y - 1);
  // x / ts as x * (1 / ts): the tile edge is a power of two, so both are the
  // exact quotient rounded once, the same binary32 value.
  float inv_ts = 1.0f / (float)fp.tile_size;
  TileBox b;
  b.tmin_x = gs_f2i(gs_clampf(gs_floorf(bminx * inv_ts), 0.0f, max_tx));
  b.tmin_y = gs_f2i(gs_clampf(gs_floorf(bminy * inv_ts), 0.0f, max_ty));
  b.tmax_x = gs_f2i(gs_clampf(gs_floorf(bmaxx * inv_ts), 0.0f, max_tx));
  b.tmax_y = gs_f2i(gs_clampf(gs_floorf(bmaxy * inv_ts), 0.0f, max_ty));
  if (fp.boundary_axis == 1) {
    b.tmin_x = 0;
    b.tmax_x = (int)fp.tiles_x - 1;
  }
  if (fp.boundary_axis == 2) {
    b.tmin_y = 0;
    b.tmax_y = (int)fp.tiles_y - 1;
  }
  return b;
}

// The box's tiles passing the exact tile test, in one thread; `mask` gets
// the row-major hit bits of a box of at most kMaskTiles tiles.
GS_DEV uint32_t count_box(const FrameParams& fp, const ProjGeom& g, const TileBox& b,
                          uint32_t& mask) {
  const float conic[3] = {g.conic[0], g.conic[1], g.conic[2]};
  const float vw = fp.viewport[0], vh = fp.viewport[1], ts = (float)fp.tile_size;
  const bool small = b.width() * b.height() <= kMaskTiles;
  uint32_t count = 0u, bitpos = 0u;
  mask = 0u;
  for (int ty = b.tmin_y; ty <= b.tmax_y; ty++)
    for (int tx = b.tmin_x; tx <= b.tmax_x; tx++, bitpos++)
      if (tile_hit(g.cx, g.cy, conic, g.thr, g.rx, g.ry, tx, ty, vw, vh, fp.boundary_axis,
                   ts)) {
        count++;
        if (small) mask |= 1u << bitpos;
      }
  return count;
}

// Tile i of a box `bw` tiles wide in row-major order: row = i / bw through
// a binary32 reciprocal (exact for the quotients of tile grids), then one
// integer correction either way.
GS_DEV void box_tile(uint32_t i, uint32_t bw, float inv_bw, uint32_t& row, uint32_t& col) {
  int r = (int)(((float)i + 0.5f) * inv_bw);
  int c = (int)i - r * (int)bw;
  if (c < 0) {
    r--;
    c += (int)bw;
  } else if (c >= (int)bw) {
    r++;
    c -= (int)bw;
  }
  row = (uint32_t)r;
  col = (uint32_t)c;
}

// Tile counts of the wave's boxes above kMaskTiles tiles (flagged `big`),
// one box at a time over all 64 lanes: a thread walking a near-plane
// Gaussian's thousands of tiles alone would hold its wave for all of them.
// The count does not depend on the order the tiles are tested in. Needs
// the whole wave.
GS_DEV void count_big_boxes(const FrameParams& fp, bool big, const ProjGeom& g,
                            const TileBox& b, uint32_t& count) {
  const uint32_t lane = lane_id();
  const float vw = fp.viewport[0], vh = fp.viewport[1], ts = (float)fp.tile_size;
  for (uint64_t m = __ballot(big); m != 0ull; m &= m - 1ull) {
    const int src = __builtin_ctzll(m);
    const float cx = __shfl(g.cx, src, 64), cy = __shfl(g.cy, src, 64);
    const float conic[3] = {__shfl(g.conic[0], src, 64), __shfl(g.conic[1], src, 64),
                            __shfl(g.conic[2], src, 64)};
    const float thr = __shfl(g.thr, src, 64), rx = __shfl(g.rx, src, 64),
                ry = __shfl(g.ry, src, 64);
    const int tx0 = __shfl(b.tmin_x, src, 64), ty0 = __shfl(b.tmin_y, src, 64);
    const uint32_t bw = (uint32_t)__shfl((int)b.width(), src, 64);
    const uint32_t n = bw * (uint32_t)__shfl((int)b.height(), src, 64);
    const float inv_bw = 1.0f / (float)bw;
    uint32_t c = 0u;
    for (uint32_t i = lane; i < n; i += 64u) {
      uint32_t row, col;
      box_tile(i, bw, inv_bw, row, col);
      if (tile_hit(cx, cy, conic, thr, rx, ry, tx0 + (int)col, ty0 + (int)row, vw, vh,
                   fp.boundary_axis, ts))
        c++;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += (uint32_t)__shfl_xor((int)c, d, 64);
    if ((int)lane == src) count = c;
  }
}

// The rest of ProjectionStage.ts:385-649 once the box's tiles are counted:
// the hit record and the SH colour.
GS_DEV ProjOut project_finish_counted(const FrameParams& fp, const SceneView& sc,
                                      const ProjGeom& g, const TileBox& box, uint32_t count,
                                      uint32_t mask) {
  ProjOut o;
  o.kind = kProjNone;
  o.count = 0u;
  o.hits = make_uint2(0u, 0u);
  if (g.kind != kProjVisible) {
    o.kind = g.kind;
    if (g.kind == kProjSubpixel) o.mean = make_float4(g.cx, g.cy, g.depth, -1.0f);
    return o;
  }
  const uint32_t slot = g.slot;
  const float thr = g.thr, rx = g.rx, ry = g.ry;
  const float conic[3] = {g.conic[0], g.conic[1], g.conic[2]};
  if (count == 0) return o;
  const uint32_t bw = box.width(), bh = box.height();
  const bool masked = bw * bh <= kMaskTiles && fp.tile_count <= (1u << kHitOriginBits);
  const uint32_t origin = (uint32_t)box.tmin_y * fp.tiles_x + (uint32_t)box.tmin_x;
  o.hits = masked ? make_uint2(mask, kHitMasked | ((bw - 1u) << kHitOriginBits) | origin)
                  : make_uint2(count, 0u);
  // SH colour (projectionHelpers.ts:118-176), direction = normalized
  // object-space offset from the camera.
  uint32_t degree = fp.sh_degree;
  uint32_t ncoef = (degree + 1) * (degree + 1);
  const uint32_t* shc = sc.sh + (size_t)slot * ncoef;
  // Degree 3 (16 words, 64-byte aligned rows) is read as four 16-byte
  // vectors: one load instruction per 16 bytes instead of per word.
  uint32_t shw[16];
  if (degree == 3) {
    const uint4* sh4 = reinterpret_cast<const uint4*>(shc);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      uint4 q4 = sh4[v];
      shw[4 * v] = q4.x;
      shw[4 * v + 1] = q4.y;
      shw[4 * v + 2] = q4.z;
      shw[4 * v + 3] = q4.w;
    }
  } else {
#pragma unroll
    for (int v = 0; v < 16; v++) shw[v] = (uint32_t)v < ncoef ? shc[v] : 0u;
  }
  float x = g.dir[0], y = g.dir[1], z = g.dir[2];
  float k[3];
  decode_sh(shw[0], k);
  float w0 = 0.28209479177387814f;
  float col[3] = {w0 * k[0], w0 * k[1], w0 * k[2]};
#define GS_SH_ADD(wexpr, idx)                       \
  {                                                 \
    float w_ = (wexpr);                             \
    decode_sh(shw[idx], k);                         \
    col[0] = col[0] + w_ * k[0];                    \
    col[1] = col[1] + w_ * k[1];                    \
    col[2] = col[2] + w_ * k[2];                    \
  }
  if (degree >= 1) {
    GS_SH_ADD(-0.4886025119029199f * y, 1);
    GS_SH_ADD(0.4886025119029199f * z, 2);
    GS_SH_ADD(-0.4886025119029199f * x, 3);
  }
  if (degree >= 2) {
    float xx = x * x, yy = y * y, zz = z * z;
    GS_SH_ADD(1.0925484305920792f * x * y, 4);
    GS_SH_ADD(-1.0925484305920792f * y * z, 5);
    GS_SH_ADD(0.31539156525252005f * (2.0f * zz - xx - yy), 6);
    GS_SH_ADD(-1.0925484305920792f * x * z, 7);
    GS_SH_ADD(0.5462742152960396f * (xx - yy), 8);
  }
  if (degree >= 3) {
    float xx = x * x, yy = y * y, zz = z * z;
    GS_SH_ADD(-0.5900435899266435f * y * (3.0f * xx - yy), 9);
    GS_SH_ADD(2.890611442640554f * x * y * z, 10);
    GS_SH_ADD(-0.4570457994644658f * y * (4.0f * zz - xx - yy), 11);
    GS_SH_ADD(0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy), 12);
    GS_SH_ADD(-0.4570457994644658f * x * (4.0f * zz - xx - yy), 13);
    GS_SH_ADD(1.445305721320277f * z * (xx - yy), 14);
    GS_SH_ADD(-0.5900435899266435f * x * (xx - 3.0f * yy), 15);
  }
#undef GS_SH_ADD
  float r = gs_clampf(col[0] + 0.5f, 0.0f, 1.0f);
  float g_ = gs_clampf(col[1] + 0.5f, 0.0f, 1.0f);
  float b = gs_clampf(col[2] + 0.5f, 0.0f, 1.0f);
  o.kind = kProjVisible;
  o.mean = make_float4(g.cx, g.cy, g.depth, g.opacity);
  o.conic = make_float4(conic[0], conic[1], conic[2], rx + ry * 512.0f);
  o.color = make_float4(r, g_, b, gs_b2f(slot ^ 0x3f800000u));
  o.thr = thr;
  o.support = support_terms(conic, thr);
  o.count = count;
  return o;
}

// The largest rgb8e8 exponent of the higher SH bands, minus the bias: the
// log2 scale of the strongest view-dependent colour term (a learned-cull
// feature; -127 without higher bands).
GS_DEV uint32_t sh_top_byte(const uint32_t* __restrict__ sh, uint32_t degree, uint32_t slot) {
  uint32_t ncoef = (degree + 1) * (degree + 1);
  const uint32_t* shc = sh + (size_t)slot * ncoef;
  uint32_t top = 0u;
  if (degree == 3) {
    const uint4* sh4 = reinterpret_cast<const uint4*>(shc);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      uint4 q4 = sh4[v];
      if (v > 0) top = max(top, q4.x >> 24);
      top = max(max(top, q4.y >> 24), max(q4.z >> 24, q4.w >> 24));
    }
  } else {
    for (uint32_t v = 1; v < ncoef; v++) top = max(top, shc[v] >> 24);
  }
  return top;
}

// The scene's exponent bytes, computed once per upload: the feature then
// reads one byte per Gaussian instead of its 64-byte SH row.
__global__ __launch_bounds__(kBlock) void k_sh_top(const uint32_t* __restrict__ sh, uint32_t n,
                                                   uint32_t degree, uint8_t* __restrict__ out) {
  uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
  if (slot < n) out[slot] = (uint8_t)sh_top_byte(sh, degree, slot);
}

GS_DEV float sh_top_exponent(const FrameParams& fp, const SceneView& sc, uint32_t slot) {
  if (sc.sh_top) return (float)(int)sc.sh_top[slot] - 127.0f;
  uint32_t degree = fp.sh_degree;
  uint32_t ncoef = (degree + 1) * (degree + 1);
  const uint32_t* shc = sc.sh + (size_t)slot * ncoef;
  uint32_t top = 0u;
  if (degree == 3) {
    const uint4* sh4 = reinterpret_cast<const uint4*>(shc);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      uint4 q4 = sh4[v];
      if (v > 0) top = max(top, q4.x >> 24);
      top = max(max(top, q4.y >> 24), max(q4.z >> 24, q4.w >> 24));
    }
  } else {
    for (uint32_t v = 1; v < ncoef; v++) top = max(top, shc[v] >> 24);
  }
  return (float)(int)top - 127.0f;
}

// ------------------------------------------------------ group launches ----
//
// Every kernel of the frame takes its arguments for a group of up to
// kGroupMax views at once (by value, in the kernel argument segment, so a
// captured launch keeps them): block b works for view b / bpv as its block
// b % bpv, or, in the persistent raster kernels, walks the views in order.
// A single frame is a group of one. Batches render several views per
// group, so each stage is one launch over all of them: the short,
// latency-bound launches (sort passes, scans, the heavy-unit list) carry
// several views' work, and one view's raster tail overlaps the next one's
// units.
constexpr int kGroupMax = kViewGroupMax;
// Kernels are instantiated for groups of one (kN = 1: the view index is a
// constant, so the arguments are read at fixed offsets as for a plain
// launch) and of up to kGroupMax views.
template <class A, int kN = kGroupMax>
struct ViewArgs {
  A v[kN];
  uint32_t n;    // views in the group
  uint32_t bpv;  // blocks per view
};

// ProjectionStage.ts:385-649 for active ordinal gid (live lanes; every
// lane of the wave must call it). Boxes above kMaskTiles tiles are counted
// by the whole wave. With kMode != plain it also fills the 16 cull features
// of a visible Gaussian.
template <int kMode>
GS_DEV ProjOut project_gaussian(const FrameParams& fp, const SceneView& sc, uint32_t gid,
                                bool live, float feat[kCullIn]) {
  ProjGeom g;
  g.kind = kProjNone;
  if (live) g = project_geometry(fp, sc, gid);
  TileBox b = {0, 0, -1, -1};
  uint32_t count = 0u, mask = 0u;
  bool big = false;
  if (g.kind == kProjVisible) {
    b = tile_box(fp, g);
    big = b.width() * b.height() > kMaskTiles;
    if (!big) count = count_box(fp, g, b, mask);
  }
  count_big_boxes(fp, big, g, b, count);
  ProjOut o = project_finish_counted(fp, sc, g, b, count, mask);
  if (kMode != kModePlain && o.kind == kProjVisible)
    cull_features(feat, g.rx, g.ry, g.conic, g.opacity, g.depth, g.dir, g.local,
                  sh_top_exponent(fp, sc, g.slot), g.cx, g.cy, fp.inv_viewport[0],
                  fp.inv_viewport[1]);
  return o;
}

GS_DEV void store_projection(const ProjectedView& out, uint32_t gid, const ProjOut& o) {
  if (o.kind == kProjVisible) {
    out.mean[gid] = o.mean;
    Splat* s = &out.splat[gid];
    s->mean = o.mean;
    s->conic = o.conic;
    s->color = o.color;
    s->aux = make_float4(o.thr, o.support.x, o.support.y, o.support.z);
    out.tile_counts[gid] = o.count;
    out.hits[gid] = o.hits;
  } else {
    out.tile_counts[gid] = 0u;
    out.mean[gid] = o.kind == kProjSubpixel ? o.mean : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// Projection of every active Gaussian; the collect variant also stores the
// features and clears the per-Gaussian contribution record of the
// instrumented raster.
struct ProjectArgs {
  FrameParams fp;
  DeviceState* state;
  SceneView sc;
  ProjectedView out;
  float* feats;
  uint32_t* record;
  const CullNet* net;
};

template <int kMode, int kN>
__global__ __launch_bounds__(kBlock) void k_project(const ViewArgs<ProjectArgs, kN> grp) {
  const uint32_t view = kN == 1 ? 0u : blockIdx.x / grp.bpv;
  const ProjectArgs& a = grp.v[view];
  const FrameParams& fp = a.fp;
  DeviceState* __restrict__ state = a.state;
  const SceneView& sc = a.sc;
  const ProjectedView& out = a.out;
  float* __restrict__ feats = a.feats;
  uint32_t* __restrict__ record = a.record;
  const uint32_t block = blockIdx.x - view * grp.bpv;
  uint32_t gid = block * kBlock + threadIdx.x;
  if (gid == 0) {
    state->checked = 0ull;
    state->blended = 0ull;
    state->checked_heavy = 0ull;
    state->sort_error = 0u;
  }
  const bool live = gid < fp.active_count;
  if (__ballot(live) == 0ull) return;  // wave-uniform
  float feat[kCullIn];
  ProjOut o = project_gaussian<kMode>(fp, sc, gid, live, feat);
  if (!live) return;
  if constexpr (kMode == kModeCollect) {
    record[gid] = 0u;
    if (o.kind == kProjVisible)
      for (int k = 0; k < kCullIn; k++) feats[(size_t)gid * kCullIn + k] = feat[k];
  }
  store_projection(out, gid, o);
}

// The learned-cull projection, in two phases. Every Gaussian is projected
// up to its tile footprint; every wave with a candidate runs the MLP over
// its 64 lanes (matrix cores, weights staged in LDS per block) and drops
// those predicted below the threshold as if they touched no tile. The
// survivors (a minority once the cull is trained) are then gathered in
// LDS, packed at the front of the block, and only they take the exact tile
// count and SH colour -- in dense waves, so the culled lanes cost nothing
// there.
struct CullSurvivor {
  uint32_t gid, slot;
  float cx, cy, depth, opacity, c0, c1, c2, thr, rx, ry, d0, d1, d2;
};
// Register budget: left to itself the compiler keeps the network's operands
// in 108 VGPRs (4 waves per SIMD, for the projection phase too); asking for
// 6 waves per SIMD gives 80 VGPRs with no spills (3M views at 1080p, 8 views
// per launch: 1,849 -> 1,976 culled frames/s).
template <int kN>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void
k_project_cull(const ViewArgs<ProjectArgs, kN> grp) {
  const uint32_t view = kN == 1 ? 0u : blockIdx.x / grp.bpv;
  const ProjectArgs& a = grp.v[view];
  const FrameParams& fp = a.fp;
  DeviceState* __restrict__ state = a.state;
  const SceneView& sc = a.sc;
  const ProjectedView& out = a.out;
  const CullNet* __restrict__ net = a.net;
  const uint32_t block = blockIdx.x - view * grp.bpv;
  struct Mlp {
    CullLdsNet net;
    __align__(16) uint16_t x[kBlock / 64][64 * kCullIn];
  };
  // The survivors overwrite the network once every wave has run it.
  __shared__ union {
    Mlp mlp;
    CullSurvivor surv[kBlock];
  } lds;
  __shared__ uint32_t wave_kept[kBlock / 64];
  uint32_t gid = block * kBlock + threadIdx.x;
  if (gid == 0) {
    state->checked = 0ull;
    state->blended = 0ull;
    state->checked_heavy = 0ull;
    state->sort_error = 0u;
  }
  // The network's loads are issued first, so that their latency overlaps
  // the projection's.
  cull_stage_net(net, lds.mlp.net);
  ProjGeom g;
  g.kind = kProjNone;
  if (gid < fp.active_count) g = project_geometry(fp, sc, gid);
  bool cand = g.kind == kProjVisible;
  float feat[kCullIn];
#pragma unroll
  for (int k = 0; k < kCullIn; k++) feat[k] = 0.0f;
  if (cand)
    cull_features(feat, g.rx, g.ry, g.conic, g.opacity, g.depth, g.dir, g.local,
                  sh_top_exponent(fp, sc, g.slot), g.cx, g.cy, fp.inv_viewport[0],
                  fp.inv_viewport[1]);
  __syncthreads();
  bool keep = false;
  if (__ballot(cand) != 0ull) {  // wave-uniform
    float y = cull_infer_wave(lds.mlp.net, lds.mlp.x[threadIdx.x >> 6], feat);
    keep = cand && !(y < lds.mlp.net.log_threshold);
  }
  // Culled, subpixel and invisible Gaussians are published now.
  if (gid < fp.active_count && !keep) {
    ProjOut o;
    o.kind = g.kind == kProjSubpixel ? kProjSubpixel : kProjNone;
    o.mean = make_float4(g.cx, g.cy, g.depth, -1.0f);
    store_projection(out, gid, o);
  }
  // Pack the survivors at the front of the block, in thread order.
  const uint64_t kept = __ballot(keep);
  const uint32_t wave = threadIdx.x >> 6;
  if (lane_id() == 0) wave_kept[wave] = (uint32_t)__popcll(kept);
  __syncthreads();  // every wave is past the network: its LDS is free
  uint32_t before = 0u, total = 0u;
#pragma unroll
  for (uint32_t w = 0; w < kBlock / 64; w++) {
    uint32_t c = wave_kept[w];
    if (w < wave) before += c;
    total += c;
  }
  if (keep) {
    uint64_t lower = lane_id() == 0 ? 0ull : (kept & (~0ull >> (64 - lane_id())));
    CullSurvivor& v = lds.surv[before + (uint32_t)__popcll(lower)];
    v = {gid, g.slot, g.cx, g.cy, g.depth, g.opacity, g.conic[0], g.conic[1], g.conic[2],
         g.thr, g.rx, g.ry, g.dir[0], g.dir[1], g.dir[2]};
  }
  __syncthreads();
  // Waves holding survivors finish them, the lanes past the last survivor
  // joining in the wave's count of large boxes.
  if ((threadIdx.x & ~63u) < total) {  // wave-uniform
    const bool live = threadIdx.x < total;
    uint32_t sgid = 0u;
    ProjGeom s;
    s.kind = kProjNone;
    if (live) {
      const CullSurvivor v = lds.surv[threadIdx.x];
      sgid = v.gid;
      s.kind = kProjVisible;
      s.slot = v.slot;
      s.cx = v.cx;
      s.cy = v.cy;
      s.depth = v.depth;
      s.opacity = v.opacity;
      s.conic[0] = v.c0;
      s.conic[1] = v.c1;
      s.conic[2] = v.c2;
      s.thr = v.thr;
      s.rx = v.rx;
      s.ry = v.ry;
      s.dir[0] = v.d0;
      s.dir[1] = v.d1;
      s.dir[2] = v.d2;
    }
    TileBox box = {0, 0, -1, -1};
    uint32_t count = 0u, mask = 0u;
    bool big = false;
    if (live) {
      box = tile_box(fp, s);
      big = box.width() * box.height() > kMaskTiles;
      if (!big) count = count_box(fp, s, box, mask);
    }
    count_big_boxes(fp, big, s, box, count);
    if (live) store_projection(out, sgid, project_finish_counted(fp, sc, s, box, count, mask));
  }
}

// ----------------------------------------------------------------- scans ----
//
// Single-pass exclusive scan with decoupled look-back (one kernel per scan).
// Block b scans partition b of kScanPartition elements; wave w of the block
// owns the partition's w-th run of 64 * kScanItems elements and loads it in
// coalesced rounds of 64 (element k * 64 + lane of the run), scanning each
// round across the wave with a running carry. The block's total is
// published at once (aggregate), wave 0 then walks back over the preceding
// partitions 64 at a time until it meets an inclusive prefix, and thread 0
// publishes the partition's inclusive prefix. The last partition's prefix
// is the scan total, which feeds DeviceState by mode. All additions
// saturate; saturating addition of unsigned values is associative and
// commutative, so the result equals the reference's serial Blelloch scan
// (scan.ts:27-62) bit for bit.
// Partitions are taken in block order: workgroups are dispatched in
// increasing order on each XCD, so the lowest unfinished partition is
// always resident and every wait ends. No per-block atomic is taken (on
// this GPU same-address atomics from many blocks serialize at about 12 ns
// each). The status words start at zero: the next radix sort's clearing
// kernel zeroes them after each scan (k_sort_histograms). A look-back wait is
// bounded; exceeding it raises bit 1 of DeviceState::sort_error.

constexpr uint32_t kLookbackSpinLimit = 1u << 20;
constexpr uint64_t kScanAggregate = 1ull << 32;
constexpr uint64_t kScanPrefix = 2ull << 32;

struct ScanWorkspace {
  unsigned long long* status;  // per partition: flag << 32 | saturating sum
};

struct LoadVisible {  // scan.ts:78-83: 1 if projectedMean.w > 0
  const float4* mean;
  GS_DEV uint32_t operator()(uint32_t i) const { return mean[i].w > 0.0f ? 1u : 0u; }
};
struct LoadOrdered {  // visibility.ts:63-76: tileCounts[sorted[rank].id]
  const uint2* sorted_a;   // the depth sort's last output ...
  const uint2* sorted_b;   // ... and its last input (when that pass was skipped)
  const uint32_t* in_place;
  const uint2* hits;       // per ordinal: tile-hit record (kHitMasked)
  uint2* ordered_hits;     // the same records in depth order, for emission
  GS_DEV uint32_t operator()(uint32_t r) const {
    const uint2* sorted = *in_place ? sorted_b : sorted_a;
    uint2 h = hits[sorted[r].y];
    ordered_hits[r] = h;
    return hit_count(h);
  }
};
struct StoreCompact {  // visibility.ts:37-61 (+ visibleOffsets)
  const float4* mean;
  uint32_t depth_mode;
  const FrameCamera* camera;
  uint32_t* offsets;
  uint2* records;
  GS_DEV void operator()(uint32_t i, uint32_t prefix, uint32_t flag) const {
    offsets[i] = prefix;
    if (!flag) return;
    float depth = mean[i].z;
    uint32_t key;
    if (depth_mode == 0) {
      key = gs_f2b(depth);
    } else {
      float near = camera->near_plane, far = camera->far_plane;
      float q = gs_clampf((depth - near) / (far - near), 0.0f, 1.0f);
      key = gs_f2u(__builtin_rintf(q * 65535.0f));
    }
    records[prefix] = make_uint2(key, i);
  }
};
struct StoreOrdered {
  uint32_t* ordered;
  uint32_t* offsets;
  GS_DEV void operator()(uint32_t r, uint32_t prefix, uint32_t value) const {
    ordered[r] = value;
    offsets[r] = prefix;
  }
};

// What a scan's total feeds.
GS_DEV void scan_finish(ScanMode mode, uint32_t total, uint32_t capacity,
                        const FrameCamera* camera, DeviceState* state) {
  if (mode == kScanVisible) {
    state->visible = total;
  } else if (mode == kScanIntersections) {
    // intersections.ts:10-41 and :114-134: never publish a truncated set.
    uint32_t overflow = total > capacity ? 1u : 0u;
    state->total = total;
    state->count = overflow ? 0u : total;
    state->overflow = overflow;
    uint32_t attempt = camera->attempt;
    uint32_t last_version = camera->reset_valid ? 0u : state->out_version;
    state->out_attempt = attempt;
    state->out_version = overflow ? last_version : attempt;
    state->out_demand = total;
    state->out_overflow = overflow;
  }
}

GS_DEV uint32_t wave_sum_sat(uint32_t v) { return __shfl(wave_inclusive_scan(v), 63, 64); }

template <class Load, class Store>
struct ScanArgs {
  Load load;
  Store store;
  LenSpec len_spec;
  ScanWorkspace ws;
  ScanMode mode;
  uint32_t capacity;
  const FrameCamera* camera;
  DeviceState* state;
};

template <class Load, class Store, int kN>
__global__ __launch_bounds__(kBlock) void k_scan(const ViewArgs<ScanArgs<Load, Store>, kN> grp) {
  const uint32_t view = kN == 1 ? 0u : blockIdx.x / grp.bpv;
  const ScanArgs<Load, Store>& a = grp.v[view];
  const Load& load = a.load;
  const Store& store = a.store;
  const LenSpec& len_spec = a.len_spec;
  const ScanWorkspace& ws = a.ws;
  const ScanMode mode = a.mode;
  const uint32_t capacity = a.capacity;
  const FrameCamera* __restrict__ camera = a.camera;
  DeviceState* __restrict__ state = a.state;
  __shared__ uint32_t s_prefix;
  __shared__ uint32_t wave_tot[kBlock / 64];
  const uint32_t t = threadIdx.x, wave = t >> 6, lane = lane_id();
  const uint32_t len = eval_len(len_spec);
  const uint32_t nparts = (uint32_t)(((uint64_t)len + kScanPartition - 1) / kScanPartition);
  const uint32_t part = blockIdx.x - view * grp.bpv;
  if (part < nparts) {
    const uint32_t run = part * (uint32_t)kScanPartition + wave * 64u * kScanItems;
    uint32_t val[kScanItems], ex[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      uint32_t i = run + k * 64u + lane;
      val[k] = i < len ? load(i) : 0u;
    }
    uint32_t carry = 0u;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      uint32_t incl = wave_inclusive_scan(val[k]);
      uint32_t prev = __shfl_up(incl, 1, 64);
      ex[k] = sat_add(carry, lane == 0 ? 0u : prev);
      carry = sat_add(carry, __shfl(incl, 63, 64));
    }
    if (lane == 0) wave_tot[wave] = carry;
    __syncthreads();
    uint32_t before = 0u, block_total = 0u;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; w++) {
      if (w < wave) before = sat_add(before, wave_tot[w]);
      block_total = sat_add(block_total, wave_tot[w]);
    }
    if (t == 0)
      __hip_atomic_store(&ws.status[part],
                         (part == 0 ? kScanPrefix : kScanAggregate) | block_total,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (wave == 0) {
      uint32_t excl = 0u;
      int32_t q = (int32_t)part - 1;
      uint32_t spins = 0u;
      while (q >= 0) {
        int32_t j = q - (int32_t)lane;
        unsigned long long v =
            j >= 0 ? __hip_atomic_load(&ws.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : kScanPrefix;  // before the first partition: an empty prefix
        uint32_t flag = (uint32_t)(v >> 32);
        uint64_t prefix_lanes = __ballot(flag == 2u);
        uint32_t stop = prefix_lanes ? (uint32_t)__builtin_ctzll(prefix_lanes) : 63u;
        uint64_t needed = stop == 63u ? ~0ull : ((2ull << stop) - 1ull);
        if (__ballot(flag == 0u) & needed) {
          if (++spins >= kLookbackSpinLimit) {
            if (lane == 0) atomicOr(&state->sort_error, 2u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl = sat_add(excl, wave_sum_sat(lane <= stop ? (uint32_t)v : 0u));
        if (prefix_lanes) break;
        q -= 64;
      }
      if (t == 0) {
        s_prefix = excl;
        uint32_t incl = sat_add(excl, block_total);
        if (part > 0)
          __hip_atomic_store(&ws.status[part], kScanPrefix | incl, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        if (part == nparts - 1) scan_finish(mode, incl, capacity, camera, state);
      }
    }
    __syncthreads();
    uint32_t prefix = sat_add(s_prefix, before);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      uint32_t i = run + k * 64u + lane;
      if (i < len) store(i, sat_add(prefix, ex[k]), val[k]);
    }
  } else if (part == 0 && t == 0) {
    scan_finish(mode, 0u, capacity, camera, state);  // empty sequence
  }
}

// ----------------------------------------------------------- radix sort ----
//
// Stable LSD sort of (key, value) pairs with kBits-bit digits (4, 6 or 8),
// one kernel per digit pass ("onesweep" with decoupled look-back):
//  - k_sort_histograms counts the digits of every pass in one read of the
//    keys (global digit totals, hence each digit's start in the output),
//    and clears what the passes and the next scan need at zero;
//  - block b of k_sort_pass takes partition b of kSortThreads * kItems keys
//    (kItems = 8, or 16 for sorts sized for at least kLargeSort keys); wave
//    w holds the partition's w-th run of 64 * kItems consecutive keys in
//    registers (key r*64 + lane of the run). Per round, the wave finds each
//    key's same-digit peers with kBits + 1 ballots (the extra bit marks keys
//    past the end); the key's rank in the run is the wave's running count of
//    its digit plus the popcount of its lower peers. One thread per digit
//    then turns the waves' counts into offsets in wave order, which is
//    input order, publishes the partition's digit count, walks
//    back over the preceding partitions' published counts (kLookbackWindow
//    partitions per step) until it meets an inclusive prefix, and publishes
//    its own inclusive prefix. The keys go to their place in the
//    partition's digit order in LDS and from there to the output, each
//    digit's run written by consecutive threads. Keys keep their input
//    order within a digit.
// A pass whose digit totals put every key in one digit is a stable identity
// permutation; its blocks only copy their keys (depth keys of one frame
// often share the top byte). When the caller passes `in_place` for the last
// pass, such a pass does not even copy: it records 1 there (0 otherwise) and
// the consumers read the pass's input buffer.
// Passes alternate between two sets of status rows: each block of pass p
// clears its own row in the other set for pass p + 1 (pass p - 1, its last
// user, has finished; every pass has the same partitions). A block waits
// only on lower partitions, which in-order dispatch has already made
// resident (as for the scans); the wait is bounded anyway, and a block that
// exceeds the bound raises DeviceState::sort_error instead of waiting on.

constexpr int kMaxPasses = kMaxSortPasses;
// Partitions read per look-back step. Every step issues its loads
// unconditionally, so a wide window mostly reads status rows it does not
// need: measured on this GPU, 4 beats 1, 2, 8, 32 and 64 (1M/1080p frame:
// depth + tile sort 0.133 ms at 4, 0.153 at 32, 0.182 at 64; 79 M tile keys:
// 1.18 ms at 4, 1.57 at 32).
constexpr int kLookbackWindow = 4;
// Blocks of the digit-total kernel: a quarter of the 4096-key partitions
// the sort is sized for, between 256 and 1024 (79 M keys: 0.27 ms at 256
// blocks, 0.17 at 1024). Of these, only as many count as the keys present
// fill with kHistKeys each; the others only clear.
constexpr uint32_t kHistBlocksMin = 256, kHistBlocksMax = 1024;
// Sorts of at least this many keys (the capacity the grid is sized for) use
// partitions of 16 keys per thread instead of 8: half the partitions, half
// the look-back and status traffic (79 M tile keys: 1.07 -> 0.94 ms); below
// it the larger partitions leave compute units idle (1M frame depth sort
// 0.077 -> 0.087 ms).
constexpr uint32_t kLargeSort = 16u << 20;
constexpr int kHistItems = 16;         // keys per lane in flight (digit totals)
constexpr uint32_t kHistKeys = kBlock * kHistItems;  // keys per counting block and round
constexpr int kSortWaves = kSortThreads / 64;
// k_sort_pass gives each of the up to kMaxRadix digits its own thread (digit
// totals, offsets, look-back, status rows) and its block scan takes four
// waves; fewer threads would leave digits without an output offset.
static_assert(kSortThreads >= kMaxRadix && kSortThreads % 64 == 0,
              "a radix-pass block needs a thread per digit");

// Exclusive scan of one value per digit thread (t < 256) across a block of
// kSortThreads threads; every thread of the block must call it.
GS_DEV uint32_t digit_exclusive_scan(uint32_t v, uint32_t* lds4) {
  uint32_t incl = wave_inclusive_scan(v);
  uint32_t prev = __shfl_up(incl, 1, 64);
  uint32_t excl = lane_id() == 0 ? 0u : prev;
  uint32_t wave = threadIdx.x >> 6;
  if (wave < 4u && lane_id() == 63) lds4[wave] = incl;
  __syncthreads();
  uint32_t before = 0u;
  for (uint32_t w = 0; w < 4u && w < wave; w++) before = sat_add(before, lds4[w]);
  return sat_add(before, excl);
}
constexpr uint32_t kStatusAggregate = 1u << 30;
constexpr uint32_t kStatusPrefix = 2u << 30;
constexpr uint32_t kStatusCount = (1u << 30) - 1u;

// Views into a lane's sort workspace. The digit totals of the frame's two
// sorts sit side by side (kMaxPasses rows of kHistShards x 256 words each):
// `ghist` is this sort's rows, `ghist_other` the other sort's, which this
// sort's counting kernel zeroes for their next use (the other sort has
// finished by then, and its rows start at zero before the first frame).
// Then come two sets of partition status rows [set][parts][256]; and the
// scans' status words, which the counting kernel also zeroes.
struct SortWorkspace {
  uint32_t* ghist;
  uint32_t* ghist_other;
  uint32_t* status;
  uint32_t parts;
  unsigned long long* scan_status;
  uint32_t scan_words;
};

GS_DEV uint32_t* status_row(const SortWorkspace& ws, uint32_t set, uint32_t part) {
  return ws.status + ((size_t)set * ws.parts + part) * 256u;
}

// Total of digit t in pass `pass` (the sum of its shard copies).
GS_DEV uint32_t digit_total(const SortWorkspace& ws, uint32_t pass, uint32_t t) {
  const uint32_t* row = ws.ghist + (size_t)pass * kHistShards * 256u + t;
  uint32_t v = 0u;
#pragma unroll
  for (int sh = 0; sh < kHistShards; sh++) v += row[sh * 256];
  return v;
}

// Same-digit peers of this lane within the wave (kBits + 1 ballots).
template <int kBits>
GS_DEV uint64_t digit_peers(uint32_t d) {
  uint64_t peers = ~0ull;
#pragma unroll
  for (int bit = 0; bit <= kBits; bit++) {
    uint64_t bal = __ballot((d >> bit) & 1u);
    peers &= ((d >> bit) & 1u) ? bal : ~bal;
  }
  return peers;
}

// Digit totals of every pass in one read of the keys, into this sort's
// rows (zero on entry). Every block first clears, striding over the grid:
// the other sort's rows, the status rows of the first pass (set 0) for the
// partitions present, and the scans' status words -- their last users (the
// other sort, the previous sort's passes, the scan just before) have
// finished. Then the first ceil(count / kHistKeys) blocks (at most the
// grid) count rounds of kHistKeys keys, all of a lane's loads in flight at
// once, into LDS, and add their nonzero counts into shard blockIdx %
// kHistShards of the totals; the other blocks are done. Counting blocks
// only for the keys present keeps the global atomics per total near
// count / (kHistKeys kHistShards): a grid sized for a sort's capacity
// would add every block's counts (at about 12 ns per same-address atomic
// on this GPU, 1,024 blocks cost 12 us per total).
struct HistArgs {
  const uint2* in;
  const uint32_t* count_ptr;
  uint32_t passes;
  SortWorkspace ws;
};

template <int kBits, int kN>
__global__ __launch_bounds__(kBlock) void k_sort_histograms(const ViewArgs<HistArgs, kN> grp) {
  const uint32_t view = kN == 1 ? 0u : blockIdx.x / grp.bpv;
  const HistArgs& a = grp.v[view];
  const uint2* __restrict__ in = a.in;
  const uint32_t* __restrict__ count_ptr = a.count_ptr;
  const uint32_t passes = a.passes;
  const SortWorkspace& ws = a.ws;
  const uint32_t block = blockIdx.x - view * grp.bpv, grid = grp.bpv;
  constexpr uint32_t kRadix = 1u << kBits;
  __shared__ uint32_t h[kMaxPasses * kRadix];
  const uint32_t count = *count_ptr, lane = lane_id();
  {
    const uint32_t stride = grid * kBlock, first = block * kBlock + threadIdx.x;
    const uint32_t nparts = (count + kSortPartition - 1) / kSortPartition;
    for (uint32_t i = first; i < (uint32_t)(kMaxPasses * kHistShards * 256); i += stride)
      ws.ghist_other[i] = 0u;
    for (uint32_t i = first; i < nparts * 256u; i += stride) ws.status[i] = 0u;
    for (uint32_t i = first; i < ws.scan_words; i += stride) ws.scan_status[i] = 0ull;
  }
  const uint32_t counting = min(grid, max(1u, (count + kHistKeys - 1) / kHistKeys));
  if (block >= counting) return;  // block-uniform
  for (uint32_t i = threadIdx.x; i < passes * kRadix; i += kBlock) h[i] = 0u;
  __syncthreads();
  // Each wave takes runs of 64 * kHistItems keys, all loads in flight at once.
  const uint32_t stride = counting * kHistKeys;
  for (uint32_t base = (block * kBlock + (threadIdx.x & ~63u)) * kHistItems; base < count;
       base += stride) {
    uint32_t key[kHistItems];
#pragma unroll
    for (int k = 0; k < kHistItems; k++) {
      uint32_t i = base + k * 64u + lane;
      key[k] = i < count ? in[i].x : 0u;
    }
#pragma unroll
    for (int k = 0; k < kHistItems; k++) {
      if (base + k * 64u + lane >= count) continue;
#pragma unroll
      for (int p = 0; p < kMaxPasses; p++)
        if ((uint32_t)p < passes)
          atomicAdd(&h[p * kRadix + ((key[k] >> (p * kBits)) & (kRadix - 1))], 1u);
    }
  }
  __syncthreads();
  uint32_t* shard = ws.ghist + (block % kHistShards) * 256u;
  for (uint32_t i = threadIdx.x; i < passes * kRadix; i += kBlock)
    if (h[i]) atomicAdd(&shard[(i / kRadix) * kHistShards * 256u + i % kRadix], h[i]);
}

struct SortPassArgs {
  const uint2* in;
  uint2* out;
  const uint32_t* count_ptr;
  uint32_t pass;
  SortWorkspace ws;
  DeviceState* state;
  uint32_t* in_place;
};

template <int kBits, int kItems, int kN>
__global__ __launch_bounds__(kSortThreads) void k_sort_pass(const ViewArgs<SortPassArgs, kN> grp) {
  const uint32_t view = kN == 1 ? 0u : blockIdx.x / grp.bpv;
  const SortPassArgs& a = grp.v[view];
  const uint2* __restrict__ in = a.in;
  uint2* __restrict__ out = a.out;
  const uint32_t* __restrict__ count_ptr = a.count_ptr;
  const uint32_t pass = a.pass;
  const SortWorkspace& ws = a.ws;
  DeviceState* __restrict__ state = a.state;
  uint32_t* __restrict__ in_place = a.in_place;
  static_assert(kItems % kSortItems == 0, "partitions are whole workspace partitions");
  constexpr uint32_t kRadix = 1u << kBits;
  constexpr uint32_t kRun = 64u * kItems;  // keys of one wave, consecutive
  constexpr uint32_t kPartition = kSortThreads * kItems;
  __shared__ uint32_t wave_cnt[kSortWaves][kRadix];  // running counts, then offsets
  __shared__ uint32_t digit_base[kRadix];
  __shared__ uint32_t lds4[4], lds4b[4];
  // The partition's keys in digit order, and where each digit's keys start
  // among them.
  __shared__ uint32_t part_start[kRadix];
  __shared__ uint2 stage[kPartition];
  const uint32_t t = threadIdx.x;
  const uint32_t count = *count_ptr;
  const uint32_t nparts = (count + kPartition - 1) / kPartition;
  const uint32_t part = blockIdx.x - view * grp.bpv;
  if (in_place) {
    // The last pass: a digit shared by every key leaves the order as it is.
    const uint32_t total = t < kRadix ? digit_total(ws, pass, t) : 0u;
    bool uniform = __syncthreads_or(t < kRadix && total == count);
    if (part == 0 && t == 0) *in_place = uniform ? 1u : 0u;
    if (uniform) return;
  }
  if (part >= nparts) return;  // block-uniform
  const uint32_t set = pass & 1u;
  const uint32_t shift = pass * kBits;
  const uint32_t wave = t >> 6, lane = lane_id();
  const uint32_t run = part * kPartition + wave * kRun;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  // Every global load first: the wave's run of keys and this digit's total.
  uint2 rec[kItems];
  uint32_t dig[kItems], rank[kItems];
#pragma unroll
  for (int r = 0; r < kItems; r++) {
    uint32_t i = run + r * 64u + lane;
    rec[r] = i < count ? in[i] : make_uint2(0u, 0u);
  }
  const uint32_t my_total = t < kRadix ? digit_total(ws, pass, t) : 0u;
  for (uint32_t i = t; i < kSortWaves * kRadix; i += kSortThreads) (&wave_cnt[0][0])[i] = 0u;
  // The row this partition uses in the next pass.
  if (t < 256u) status_row(ws, set ^ 1u, part)[t] = 0u;
  // A pass whose digit is the same for every key keeps the order: copy.
  if (__syncthreads_or(t < kRadix && my_total == count)) {
#pragma unroll
    for (int r = 0; r < kItems; r++) {
      uint32_t i = run + r * 64u + lane;
      if (i < count) out[i] = rec[r];
    }
    return;
  }
  // Rank within the wave's run, in order: same-digit peers of the round
  // plus the wave's running count of the digit (read by every lane before
  // the group's highest lane adds the group size; the wave runs in step).
#pragma unroll
  for (int r = 0; r < kItems; r++) {
    dig[r] = run + r * 64u + lane < count ? ((rec[r].x >> shift) & (kRadix - 1)) : kRadix;
    uint64_t peers = digit_peers<kBits>(dig[r]);
    uint32_t before = dig[r] < kRadix ? wave_cnt[wave][dig[r]] : 0u;
    rank[r] = before + (uint32_t)__popcll(peers & lt_mask);
    if (dig[r] < kRadix && (peers >> lane) == 1ull)
      wave_cnt[wave][dig[r]] = before + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  uint32_t local = 0u;
  if (t < kRadix) {
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) {
      uint32_t c = wave_cnt[w][t];
      wave_cnt[w][t] = local;
      local += c;
    }
    __hip_atomic_store(&status_row(ws, set, part)[t],
                       (part == 0 ? kStatusPrefix : kStatusAggregate) | local, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  // Each digit's start in the output: exclusive scan of the digit totals.
  uint32_t digit_start = digit_exclusive_scan(my_total, lds4);
  // Each digit's start within the partition: exclusive scan of its counts.
  uint32_t start = digit_exclusive_scan(local, lds4b);
  if (t < kRadix) part_start[t] = start;
  if (t < kRadix) {
    uint32_t excl = 0u;
    if (part > 0) {
      // Look back kLookbackWindow partitions per step (independent loads),
      // nearest first, up to the first one not yet published.
      int32_t q = (int32_t)part - 1;
      uint32_t spins = 0;
      for (;;) {
        uint32_t v[kLookbackWindow];
#pragma unroll
        for (int j = 0; j < kLookbackWindow; j++)
          v[j] = q - j >= 0 ? __hip_atomic_load(&status_row(ws, set, q - j)[t], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
                            : kStatusPrefix;  // partition 0 is always a prefix
        int used = 0;
        bool found = false, stopped = false;
#pragma unroll
        for (int j = 0; j < kLookbackWindow; j++) {
          uint32_t flag = v[j] & ~kStatusCount;
          if (!stopped && flag != 0u) {
            excl += v[j] & kStatusCount;
            used++;
            if (flag == kStatusPrefix) found = true;
          }
          stopped = stopped || flag == 0u || flag == kStatusPrefix;
        }
        if (found) break;
        q -= used;
        if (used == 0) {
          if (++spins >= kLookbackSpinLimit) {
            atomicOr(&state->sort_error, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __hip_atomic_store(&status_row(ws, set, part)[t], kStatusPrefix | (excl + local),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    digit_base[t] = digit_start + excl;
  }
  __syncthreads();
  // Keys to their place in the partition's digit order in LDS, then out in
  // that order: consecutive threads write consecutive records of a digit's
  // run, so a wave's store covers a few lines instead of one line per key
  // (6M/4K tile sort, two passes over 79 M keys: 0.94 -> 0.84 ms).
#pragma unroll
  for (int r = 0; r < kItems; r++) {
    uint32_t d = dig[r];
    if (d < kRadix) stage[part_start[d] + wave_cnt[wave][d] + rank[r]] = rec[r];
  }
  __syncthreads();
  const uint32_t valid = min((uint32_t)kPartition, count - part * (uint32_t)kPartition);
#pragma unroll
  for (int r = 0; r < kItems; r++) {
    uint32_t i = t + (uint32_t)r * kSortThreads;
    if (i < valid) {
      uint2 k = stage[i];
      uint32_t d = (k.x >> shift) & (kRadix - 1);
      out[digit_base[d] - part_start[d] + i] = k;
    }
  }
}

// ---------------------------------------------------------------- emission ----

// intersections.ts:43-111: per depth rank, row-major bbox tiles, the same
// exact tile test as the projection; sentinel (T, id) pads reserved slots.
// Boxes without a hit mask (above kMaskTiles tiles) are emitted by the
// whole wave, one box at a time in chunks of 64 tiles: lane l tests the
// chunk's l-th tile and its key goes after the hits of the earlier tiles
// (a ballot prefix), which is the row-major order of the serial walk.
struct EmitArgs {
  FrameParams fp;
  const DeviceState* state;
  const uint2* sorted_a;
  const uint2* sorted_b;
  const Splat* splat;
  const uint2* ordered_hits;
  const uint32_t* offsets;
  uint2* records;
};

template <int kN>
__global__ __launch_bounds__(kBlock) void k_emit(const ViewArgs<EmitArgs, kN> grp) {
  const uint32_t view = kN == 1 ? 0u : blockIdx.x / grp.bpv;
  const EmitArgs& a = grp.v[view];
  const FrameParams& fp = a.fp;
  const DeviceState* __restrict__ state = a.state;
  const uint2* __restrict__ sorted_a = a.sorted_a;
  const uint2* __restrict__ sorted_b = a.sorted_b;
  const Splat* __restrict__ splat = a.splat;
  const uint2* __restrict__ ordered_hits = a.ordered_hits;
  const uint32_t* __restrict__ offsets = a.offsets;
  uint2* __restrict__ records = a.records;
  uint32_t rank = (blockIdx.x - view * grp.bpv) * kBlock + threadIdx.x;
  if (state->overflow) return;  // uniform
  const bool live = rank < state->visible;
  if (__ballot(live) == 0ull) return;  // wave-uniform
  uint32_t capacity = fp.capacity;
  // The depth sort's last output, or its last input if that pass was skipped.
  const uint2* sorted_depth = state->depth_in_place ? sorted_b : sorted_a;
  uint32_t gid = 0u, reserved = 0u, base = 0u, local = 0u;
  uint2 h = make_uint2(0u, 0u);
  if (live) {
    gid = sorted_depth[rank].y;
    h = ordered_hits[rank];
    reserved = hit_count(h);
    base = offsets[rank];
  }
  const bool masked = (h.y & kHitMasked) != 0u;
  if (live && masked) {
    // The projection's hits, in the same row-major order.
    const uint32_t origin = h.y & ((1u << kHitOriginBits) - 1u);
    const uint32_t bw = ((h.y >> kHitOriginBits) & 31u) + 1u;
    for (uint32_t m = h.x; m != 0u; m &= m - 1u, local++) {
      uint32_t bit = (uint32_t)__builtin_ctz(m);
      uint32_t tile = origin + (bit / bw) * fp.tiles_x + bit % bw;
      if (base + local < capacity) records[base + local] = make_uint2(tile, gid);
    }
  }
  // Unmasked boxes: the tile footprint again from the projected mean and
  // conic, emitted by the whole wave.
  const bool big = live && !masked && reserved != 0u;
  float cx = 0.0f, cy = 0.0f, thr = 0.0f, rx = 0.0f, ry = 0.0f;
  float conic[3] = {0.0f, 0.0f, 0.0f};
  int tmin_x = 0, tmin_y = 0, tmax_x = -1, tmax_y = -1;
  if (big) {
    float4 mean = splat[gid].mean;
    float4 conic4 = splat[gid].conic;
    uint32_t packed = gs_f2u(conic4.w);
    rx = (float)(packed % 512u);
    ry = (float)(packed / 512u);
    conic[0] = conic4.x;
    conic[1] = conic4.y;
    conic[2] = conic4.z;
    cx = mean.x;
    cy = mean.y;
    thr = gs_logf(mean.w * 255.0f);
    int max_tx = (int)fp.tiles_x - 1, max_ty = (int)fp.tiles_y - 1;
    float ts = (float)fp.tile_size, inv_ts = 1.0f / ts;  // a power of two: exact
    tmin_x = min(max(gs_f2i(gs_floorf((mean.x - rx) * inv_ts)), 0), max_tx);
    tmin_y = min(max(gs_f2i(gs_floorf((mean.y - ry) * inv_ts)), 0), max_ty);
    tmax_x = min(max(gs_f2i(gs_floorf((mean.x + rx) * inv_ts)), 0), max_tx);
    tmax_y = min(max(gs_f2i(gs_floorf((mean.y + ry) * inv_ts)), 0), max_ty);
    if (fp.boundary_axis == 1) {
      tmin_x = 0;
      tmax_x = max_tx;
    }
    if (fp.boundary_axis == 2) {
      tmin_y = 0;
      tmax_y = max_ty;
    }
  }
  const uint32_t lane = lane_id();
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const float vw = fp.viewport[0], vh = fp.viewport[1], ts = (float)fp.tile_size;
  for (uint64_t m = __ballot(big); m != 0ull; m &= m - 1ull) {
    const int src = __builtin_ctzll(m);
    const float scx = __shfl(cx, src, 64), scy = __shfl(cy, src, 64);
    const float sconic[3] = {__shfl(conic[0], src, 64), __shfl(conic[1], src, 64),
                             __shfl(conic[2], src, 64)};
    const float sthr = __shfl(thr, src, 64), srx = __shfl(rx, src, 64),
                sry = __shfl(ry, src, 64);
    const int tx0 = __shfl(tmin_x, src, 64), ty0 = __shfl(tmin_y, src, 64);
    const uint32_t bw = (uint32_t)(__shfl(tmax_x, src, 64) - tx0 + 1);
    const uint32_t n = bw * (uint32_t)(__shfl(tmax_y, src, 64) - ty0 + 1);
    const uint32_t sbase = (uint32_t)__shfl((int)base, src, 64);
    const uint32_t sres = (uint32_t)__shfl((int)reserved, src, 64);
    const uint32_t sgid = (uint32_t)__shfl((int)gid, src, 64);
    const float inv_bw = 1.0f / (float)bw;
    uint32_t run = 0u;  // hits of the box's earlier chunks
    for (uint32_t i0 = 0; i0 < n; i0 += 64u) {
      const uint32_t i = i0 + lane;
      uint32_t row = 0u, col = 0u;
      bool hit = false;
      if (i < n) {
        box_tile(i, bw, inv_bw, row, col);
        hit = tile_hit(scx, scy, sconic, sthr, srx, sry, tx0 + (int)col, ty0 + (int)row, vw, vh,
                       fp.boundary_axis, ts);
      }
      const uint64_t hb = __ballot(hit);
      const uint32_t pos = run + (uint32_t)__popcll(hb & lt_mask);
      if (hit && pos < sres && sbase + pos < capacity)
        records[sbase + pos] =
            make_uint2((uint32_t)(ty0 + (int)row) * fp.tiles_x + (uint32_t)(tx0 + (int)col), sgid);
      run += (uint32_t)__popcll(hb);
    }
    if ((int)lane == src) local = run;
  }
  if (!live) return;
  for (uint32_t pad = local; pad < reserved; pad++)
    if (base + pad < capacity) records[base + pad] = make_uint2(fp.tile_count, gid);
}

// ------------------------------------------------------------ tile ranges ----

// tileOffsets.ts:4-128 in one pass: record i starts every tile in
// (tile[i - 1], tile[i]] (tile[-1] = -1, tile[K] = T), so off[t] = i there:
// every tile gets the index of the first record of the first non-empty tile
// at or after it (the suffix minimum), and off[T] ends the last range (K, or
// the first sentinel). Runs of empty tiles can be long (a frame with few
// occupied tiles), so a lane with a run above 8 tiles hands it to its whole
// wave, which fills it 64 tiles at a time.
struct TileOffsetArgs {
  FrameParams fp;
  const DeviceState* state;
  const uint2* sorted;
  uint32_t* off;
};

template <int kN>
__global__ __launch_bounds__(kBlock) void k_tile_offsets(const ViewArgs<TileOffsetArgs, kN> grp) {
  const uint32_t view = kN == 1 ? 0u : blockIdx.x / grp.bpv;
  const TileOffsetArgs& a = grp.v[view];
  const FrameParams& fp = a.fp;
  const DeviceState* __restrict__ state = a.state;
  const uint2* __restrict__ sorted = a.sorted;
  uint32_t* __restrict__ off = a.off;
  const uint32_t block = blockIdx.x - view * grp.bpv, grid = grp.bpv;
  const uint32_t count = state->count, tiles = fp.tile_count, lane = lane_id();
  for (uint32_t base = block * kBlock + (threadIdx.x & ~63u); base <= count;
       base += grid * kBlock) {
    uint32_t i = base + lane;
    uint32_t lo = 1u, hi = 0u;  // empty
    if (i <= count) {
      lo = i == 0 ? 0u : sorted[i - 1].x + 1u;
      hi = i == count ? tiles : sorted[i].x;
      if (hi > tiles) hi = tiles;
    }
    bool wide = lo <= hi && hi - lo >= 8u;
    if (!wide)
      for (uint32_t t = lo; t <= hi; t++) off[t] = i;
    uint64_t m = __ballot(wide);
    while (m) {
      int src = __builtin_ctzll(m);
      m &= m - 1ull;
      uint32_t a = __shfl(lo, src, 64), b = __shfl(hi, src, 64), v = __shfl(i, src, 64);
      for (uint32_t t = a + lane; t <= b; t += 64u) off[t] = v;
    }
  }
}

// ---------------------------------------------------------------- raster ----
//
// One block per tile (or per heavy unit, below), one pixel per thread.
// Records are taken a batch at a time, front to back:
//  phase A: the thread owning record r stages it in LDS and tests only the
//    pixels of its support rectangle (support_rect below): every pixel
//    outside the rectangle is proven to fail the reference's power test, so
//    leaving it out changes nothing. Pixels inside run the exact test and,
//    when power lies in [-ln(255 a), 0], set bit r of their candidate mask;
//  phase B: each pixel walks its candidate bits in record order and
//    evaluates the reference's blend (alpha = clamp(a e^power, 0, 0.99),
//    skip alpha < 1/255, C += c T alpha, T *= 1 - alpha, stop at T < 1e-4).
// The blend visits exactly the records the reference blends, in the same
// order with the same binary32 operations. The block leaves once every
// pixel has stopped. Global loads of the next batch are issued before the
// current batch is processed.
//
// Tiles with more than chunk_size records are split, as in the reference,
// into consecutive chunks blended from T = 1 by independent tasks and
// composited in order (TileRasterizer.ts:331-412, 807-902).

// Support rectangle of a record: the pixels outside [x0, x1] x [y0, y1]
// (global pixel indices) are those where the binary32 quadratic form
// q = c0 dx^2 + 2 c1 dx dy + c2 dy^2 provably exceeds 2 thr, i.e. the
// reference's test `power < -thr` holds. With S = |c0| dx^2 + 2|c1 dx dy| +
// |c2| dy^2, the evaluated q obeys q_hat >= q - g S (g = 3e-7 > 4 unit
// roundoffs), q >= (det / c2) dx^2 and S <= (c0 + |c1|) dx^2 + (c2 + |c1|)
// dy^2; so over any set of rows with |dy| <= Y, q_hat > 2 thr whenever
// dx^2 (det / c2 - g (c0 + |c1|)) > 2 thr + g (c2 + |c1|) Y^2. The same
// bound with x and y exchanged limits rows. The per-Gaussian terms come
// from the projection (support_terms, binary64 rounded once); here
// h^2 = a + b O^2 and the root are taken in binary32 (a few roundings,
// < 1e-6 relative) and widened by a 1e-5 relative margin, and pixels left
// out lie at least one pixel beyond the bound (which also covers the
// rounding of the pixel bounds). Rows are bounded first (over the block's
// columns), then columns over the rows that remain, so a block covering
// part of a tile can drop a record before looking at its columns.
//
// Narrows [p0, p1] along one axis (centre pc, constant term a) given that
// the other axis (centre oc) only takes pixels o0..o1.
GS_DEV void support_axis(float a, float b, float pc, float oc, int o0, int o1, int& p0,
                         int& p1) {
  const float O = gs_maxf(fabsf(((float)o0 + 0.5f) - oc), fabsf(((float)o1 + 0.5f) - oc)) + 1.0f;
  const float h = gs_sqrtf(a + b * (O * O)) * (1.0f + 1e-5f) + 1e-6f;
  const float lo = gs_floorf(pc - h - 0.5f), hi = gs_ceilf(pc + h - 0.5f);
  if (lo > (float)p0) p0 = lo > (float)p1 ? p1 + 1 : (int)lo;
  if (hi < (float)p1) p1 = hi < (float)p0 ? p0 - 1 : (int)hi;
}

// The reference's per-pair power (TileRasterizer.ts:624-656), exactly.
GS_DEV float pair_power(float pxc, float pyc, float4 m, float4 c, float vw, float vh,
                        uint32_t axis) {
  float dx, dy;
  if (axis == 0) {
    dx = pxc - m.x;
    dy = pyc - m.y;
  } else {
    boundary_delta(pxc, pyc, m.x, m.y, c.x, c.y, c.z, vw, vh, axis, dx, dy);
  }
  return (c.x * (dx * dx) + c.y * 2.0f * dx * dy + c.z * (dy * dy)) * -0.5f;
}

struct PixelSum {
  float r, g, b, d, t;
  uint32_t checked, blended;
};

// Block shapes. A tile of edge E has E*E pixels. Light tiles (up to
// kLightMax records) use one thread per pixel, max(E*E, batch) threads, and
// stage `batch` records per round (128 or 256, tunable). Heavy units (tiles
// above kLightMax records and chunk tasks) use a group of lanes per pixel
// at edges 8 and 16 (heavy_lanes) and stage one record per thread per
// round:
//  - whole tiles: one block per unit, four lanes per pixel (256 threads at
//    edge 8, 1024 at 16);
//  - at edge 16, frames with more than FrameParams::split_units units use
//    two lanes per pixel: first whole tiles (one block of 512 threads per
//    unit, three to a CU), above pair_units row slabs (two blocks of 256
//    threads per unit, each covering 8 rows, six to a CU), above
//    fine_split_units four blocks of 128 threads covering 4 rows, twelve
//    to a CU. A slab stops as soon as its own pixels are saturated and a
//    unit's work spreads over CUs, at the price of staging every record
//    once per slab, which pays only when the units outnumber the CUs many
//    times over (or saturate early, as the 6M/4K frame's do).
// Edge 32 uses one block of 1024 threads per unit (one lane per pixel, 512
// records per round).
constexpr uint32_t kLightMax = 1024;  // records up to which a tile is light
constexpr int kHeavyThreads = 1024;  // edge 32
constexpr int log2_edge(int e) { return e == 8 ? 3 : (e == 16 ? 4 : 5); }
constexpr int light_threads(int e, int batch) { return e * e > batch ? e * e : batch; }
// Heavy shapes (DeviceState::heavy_split; all but kShapeWhole at edge 16
// only): whole tiles with four lanes per pixel, 8-row and 4-row slabs, and
// whole tiles with two lanes per pixel (kShapePair, a "slab" of all 16 rows).
constexpr int kShapeWhole = 0, kShapeSlab8 = 1, kShapeSlab4 = 2, kShapePair = 3;
constexpr int heavy_slab_rows(int e, int shape) {
  return e == 16 && shape != kShapeWhole
             ? (shape == kShapeSlab8 ? 8 : (shape == kShapeSlab4 ? 4 : 16))
             : e;
}
constexpr int heavy_slabs(int e, int shape) { return e / heavy_slab_rows(e, shape); }
constexpr int heavy_slab_pixels(int e, int shape) { return e * heavy_slab_rows(e, shape); }
// Lanes per pixel at edges 8 and 16: four for whole tiles in frames with
// about as many units as CUs, which wait on the longest unit (more lanes,
// shorter units); two for the shapes of frames with many more units than
// CUs, which are bound by total VALU work and occupancy (two lanes split a
// candidate step's blends two ways instead of four and make 512- or
// smaller-thread blocks, several to a CU).
constexpr int heavy_lanes(int e, int shape) { return e == 16 && shape != kShapeWhole ? 2 : 4; }
constexpr int heavy_threads(int e, int shape) {
  return e == 32 ? kHeavyThreads : heavy_lanes(e, shape) * heavy_slab_pixels(e, shape);
}
constexpr int heavy_batch(int e, int shape) { return e == 32 ? 512 : heavy_threads(e, shape); }

template <int kPixels, int kBatch>
struct RasterLds {
  static constexpr int kWords = kBatch / 32;
  float4 mean[kBatch];   // cx, cy, depth, opacity
  float4 conic[kBatch];  // a, b, c, ln(255 opacity)
  float4 color[kBatch];
  uint32_t gid[kBatch];  // active ordinals (instrumented raster only)
  uint32_t mask[kWords][kPixels];  // [word][pixel] candidate bits of the round
};

// Blend one candidate (record j of the round) into the pixel sum.
struct Candidate {
  float alpha, depth;
  float4 color;
};

template <class Lds>
GS_DEV Candidate eval_candidate(const Lds& L, uint32_t j, float pxc, float pyc, float vw,
                                float vh, uint32_t axis) {
  float4 gm = L.mean[j];
  float4 gc = L.conic[j];
  float power = pair_power(pxc, pyc, gm, gc, vw, vh, axis);
  Candidate c;
  c.alpha = gs_clampf(gm.w * gs_expf_nonpos(power), 0.0f, 0.99f);
  c.depth = gm.z;
  c.color = L.color[j];
  return c;
}

// A record's candidate pixels in a block: [x0, x1] x [y0, y1], empty when
// x0 > x1 or y0 > y1.
struct Rect {
  int x0, x1, y0, y1;
};
// The support rectangle (above) of the record with mean m and support
// terms sup = (ln(255 opacity), rows' term, columns' term, O^2 coefficient)
// within columns tx0..tx0+E-1 and rows ry0..ry1; under the periodic
// boundary policies every pixel there is a candidate.
template <int kEdge>
GS_DEV Rect record_rect(float4 m, float4 sup, int tx0, int ry0, int ry1, uint32_t axis) {
  Rect r = {tx0, tx0 + kEdge - 1, ry0, ry1};
  if (axis != 0) return r;
  support_axis(sup.y, sup.w, m.y, m.x, r.x0, r.x1, r.y0, r.y1);  // rows
  if (r.y0 > r.y1) return r;
  support_axis(sup.z, sup.w, m.x, m.y, r.y0, r.y1, r.x0, r.x1);  // columns
  return r;
}

// Splat fields a raster round stages: mean, (conic.xyz, ln(255 opacity)),
// colour, and the support terms.
GS_DEV void gather_splat(const Splat* __restrict__ splat, uint32_t id, float4& m, float4& c,
                         float4& k, float4& sup) {
  const Splat& g = splat[id];
  m = g.mean;
  float4 cn = g.conic;
  sup = g.aux;
  c = make_float4(cn.x, cn.y, cn.z, sup.x);
  k = g.color;
}

// The candidate rows of mark_candidates under the plain boundary policy:
// pair_power with its row terms hoisted, the same operations in the same
// order; pixel centres step by 1.0f exactly (small integers + 0.5). With
// kBranchFree every tested pixel takes an LDS atomic (an OR of 0 when it
// fails), which saves the branch around it but costs LDS work per failing
// pixel.
template <int kEdge, bool kBranchFree, class Lds>
GS_DEV void mark_rows(Lds& L, uint32_t bit, uint32_t word, const Rect& rc, int tx0, int ry0,
                      float4 m, float4 c) {
  const float cy2 = c.y * 2.0f;
  for (int py = rc.y0; py <= rc.y1; py++) {
    const float dy = ((float)py + 0.5f) - m.y;
    const float qy = c.z * (dy * dy);
    uint32_t* row = &L.mask[word][(uint32_t)((py - ry0) * kEdge)];
    float pxf = (float)rc.x0 + 0.5f;
    for (int i = rc.x0 - tx0; i <= rc.x1 - tx0; i++, pxf += 1.0f) {  // column in the tile
      const float dx = pxf - m.x;
      const float power = (c.x * (dx * dx) + cy2 * dx * dy + qy) * -0.5f;
      const bool fails = power > 0.0f || power < -c.w;
      if (kBranchFree)
        atomicOr(&row[i], fails ? 0u : bit);
      else if (!fails)
        atomicOr(&row[i], bit);
    }
  }
}

// Set the candidate bits of the record in round slot `slot` for the pixels
// of rc (absolute pixel bounds, inside the block's rows) that pass the
// reference's test (TileRasterizer.ts:624-656); pixel p of the block is
// (px - tx0) + (py - ry0) * E. Under the plain policy a wave picks the
// atomic form by its records' shape: an ellipse fills pi/4 sqrt(1 - r^2) of
// its bounding rectangle (r^2 = c1^2 / (c0 c2)), so a wave holding a thin
// tilted one (r^2 > 0.9, filling under a quarter) keeps the branch and the
// rest go branch-free. Both forms set the same bits.
template <int kEdge, class Lds>
GS_DEV void mark_candidates(Lds& L, uint32_t slot, const Rect& rc, int tx0, int ry0, float4 m,
                            float4 c, float vw, float vh, uint32_t axis) {
  uint32_t bit = 1u << (slot & 31u), word = slot >> 5;
  if (axis == 0) {
    const bool thin = c.y * c.y > 0.9f * (c.x * c.z);
    if (__builtin_amdgcn_ballot_w64(thin) != 0u)  // uniform over the wave
      mark_rows<kEdge, false>(L, bit, word, rc, tx0, ry0, m, c);
    else
      mark_rows<kEdge, true>(L, bit, word, rc, tx0, ry0, m, c);
    return;
  }
  for (int py = rc.y0; py <= rc.y1; py++) {
    float pyf = (float)py + 0.5f;
    for (int px = rc.x0; px <= rc.x1; px++) {
      float power = pair_power((float)px + 0.5f, pyf, m, c, vw, vh, axis);
      if (!(power > 0.0f || power < -c.w))
        atomicOr(&L.mask[word][(uint32_t)((py - ry0) * kEdge + (px - tx0))], bit);
    }
  }
}

// Blend records [cb, ce) (cb < ce) of `tile` into one pixel per thread (threads
// 0..E*E-1), from T = 1. All kThreads threads of the block must call it;
// threads below kBatch stage one record per round (a record whose support
// rectangle is empty sets no bits and is not staged). The instrumented variant
// (kRecord) also raises record[id] to every alpha * T that Gaussian id
// contributes (binary32 bits; non-negative, so unsigned order).
template <int kEdge, int kThreads, int kBatch, bool kRecord>
GS_DEV PixelSum blend_tile_range(const FrameParams& fp, uint32_t tile, uint32_t cb, uint32_t ce,
                                 const uint2* __restrict__ sorted,
                                 const Splat* __restrict__ splat,
                                 RasterLds<kEdge * kEdge, kBatch>& L, uint32_t& x, uint32_t& y,
                                 bool& active, uint32_t* __restrict__ record) {
  constexpr int kPixels = kEdge * kEdge;
  constexpr int kMaskWords = kBatch / 32;
  static_assert(kThreads >= kPixels && kThreads >= kBatch, "block too small");
  uint32_t t = threadIdx.x;
  bool pixel_thread = t < (uint32_t)kPixels;
  bool stager = t < (uint32_t)kBatch;
  int tx0 = (int)((tile % fp.tiles_x) * kEdge), ty0 = (int)((tile / fp.tiles_x) * kEdge);
  x = (uint32_t)tx0 + (t & (uint32_t)(kEdge - 1));
  y = (uint32_t)ty0 + ((t >> log2_edge(kEdge)) & (uint32_t)(kEdge - 1));
  active = pixel_thread && x < fp.width && y < fp.height;
  float pxc = (float)x + 0.5f, pyc = (float)y + 0.5f;
  float vw = fp.viewport[0], vh = fp.viewport[1];
  uint32_t axis = fp.boundary_axis;
  float threshold = fp.transmittance_threshold;
  PixelSum s = {0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0u, 0u};
  bool done = !active;
  // Each thread holds the splat of its slot one round ahead and the ordinal
  // two rounds ahead. Slots past the range load its last record (never
  // staged): unconditional loads land in the loop-carried registers, where
  // a conditional one makes the compiler wait for the load at the join.
  const uint32_t last = ce - 1u;
  float4 m, c, k, sp;
  uint32_t cg = sorted[min(cb + t, last)].y;
  uint32_t cg_next = sorted[min(cb + kBatch + t, last)].y;
  gather_splat(splat, cg, m, c, k, sp);
  __syncthreads();
  for (uint32_t batch = cb; batch < ce; batch += kBatch) {
    for (uint32_t i = t; i < (uint32_t)(kMaskWords * kPixels); i += kThreads)
      (&L.mask[0][0])[i] = 0u;
    __syncthreads();
    bool hit = stager && batch + t < ce;
    Rect crc = {0, -1, 0, -1};
    if (hit) {
      crc = record_rect<kEdge>(m, sp, tx0, ty0, ty0 + kEdge - 1, axis);
      hit = crc.x0 <= crc.x1 && crc.y0 <= crc.y1;
    }
    if (hit) {
      L.mean[t] = m;
      L.conic[t] = c;
      L.color[t] = k;
      if (kRecord) L.gid[t] = cg;
    }
    float4 cm = m, cc = c;
    cg = cg_next;
    gather_splat(splat, cg, m, c, k, sp);
    cg_next = sorted[min(batch + 2u * kBatch + t, last)].y;
    if (hit) mark_candidates<kEdge>(L, t, crc, tx0, ty0, cm, cc, vw, vh, axis);
    __syncthreads();
    // Ordered blend over this pixel's candidates: each lane walks its own
    // candidate bits across the round's mask words, four at a time (their
    // loads and exponentials are independent; only the blend is serial).
    if (pixel_thread && !done) {
      uint32_t n = min((uint32_t)kBatch, ce - batch);
      uint32_t stop = n;
      uint32_t w = 0;
      uint32_t bits = L.mask[0][t];
      while (true) {
        uint32_t j[4];
        bool v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          while (bits == 0u && w + 1u < (uint32_t)kMaskWords) {
            w++;
            bits = L.mask[w][t];
          }
          v[q] = bits != 0u;
          j[q] = v[q] ? w * 32u + (uint32_t)__builtin_ctz(bits) : (q == 0 ? 0u : j[0]);
          bits &= bits - 1u;
        }
        if (!v[0]) break;
        // Slots past the pixel's last candidate evaluate the step's first
        // candidate again; their results are never taken.
        Candidate cd[4];
#pragma unroll
        for (int q = 0; q < 4; q++) cd[q] = eval_candidate(L, j[q], pxc, pyc, vw, vh, axis);
        // The blends in order. First assuming none of them ends the pixel:
        // a skipped candidate enters with alpha 0, which leaves every value
        // as it was (its colour and depth are finite). T only falls, so if
        // it ends at or above the threshold no blend of the step crossed
        // it; otherwise the step is applied again from the values before
        // it, as selects that stop at the crossing blend (the same
        // operations on the same values).
        const PixelSum in = s;
        if (!kRecord) {
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const bool take = v[q] && !(cd[q].alpha < (float)(1.0 / 255.0));
            const float alpha = take ? cd[q].alpha : 0.0f;
            const float contribution = s.t * alpha;
            s.r = s.r + cd[q].color.x * contribution;
            s.g = s.g + cd[q].color.y * contribution;
            s.b = s.b + cd[q].color.z * contribution;
            s.d = s.d + cd[q].depth * contribution;
            s.t = s.t * (1.0f - alpha);
            s.blended += take ? 1u : 0u;
          }
        }
        if (kRecord || s.t < threshold) {
          s = in;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const float alpha = cd[q].alpha;
            const bool take = v[q] && !done && !(alpha < (float)(1.0 / 255.0));
            const float contribution = s.t * alpha;
            if (kRecord && take) atomicMax(&record[L.gid[j[q]]], gs_f2b(contribution));
            const float r = s.r + cd[q].color.x * contribution;
            const float g = s.g + cd[q].color.y * contribution;
            const float b = s.b + cd[q].color.z * contribution;
            const float d = s.d + cd[q].depth * contribution;
            const float nt = s.t * (1.0f - alpha);
            s.r = take ? r : s.r;
            s.g = take ? g : s.g;
            s.b = take ? b : s.b;
            s.d = take ? d : s.d;
            s.blended += take ? 1u : 0u;
            s.t = take ? nt : s.t;
            const bool fin = take && nt < threshold;
            stop = fin ? j[q] + 1u : stop;
            done = done || fin;
          }
        }
        if (done) break;
      }
      s.checked += stop;  // pairs the reference's loop evaluates
    }
    if (__syncthreads_and(done || !pixel_thread)) break;
  }
  return s;
}

// DPP broadcast of lane k of each group of kLanes lanes (2 or 4) to the
// group: quad_perm [k, k, k, k], or [k, k, k + 2, k + 2] for pairs.
template <int kLanes, int k>
GS_DEV float group_bcast(float v) {
  constexpr int kSel = kLanes == 4 ? k * 85 : (k == 0 ? 0xa0 : 0xf5);
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kSel, 0xf, 0xf, false));
}
template <int kLanes>
GS_DEV void group_bcast_all(float v, float (&o)[kLanes]) {
  o[0] = group_bcast<kLanes, 0>(v);
  o[1] = group_bcast<kLanes, 1>(v);
  if constexpr (kLanes == 4) {
    o[2] = group_bcast<4, 2>(v);
    o[3] = group_bcast<4, 3>(v);
  }
}
template <int kLanes>
GS_DEV uint32_t group_or(uint32_t v) {
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 177, 0xf, 0xf, false);  // lane ^ 1
  if constexpr (kLanes == 4)
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 78, 0xf, 0xf, false);  // lane ^ 2
  return v;
}

// Heavy-unit slabs with kLanes (2 or 4) lanes per pixel (tile edges 8 and
// 16). The block covers kSlabRows rows of the tile starting at row
// slab * kSlabRows. Records are staged kBatch per round, kBatch / kThreads
// per thread, and tested against their support rectangles (clipped to the
// slab's rows) exactly as in blend_tile_range. The blend of pixel p is
// shared by lanes kLanes p .. kLanes p + kLanes - 1: per step the group
// takes the pixel's next kLanes candidates in record order, lane q
// evaluates the q-th one (power, exponential and alpha, the expensive part),
// the group exchanges the results (DPP broadcasts) and every lane applies
// the same blends in the same order, so the lanes hold the same sum, bit
// for bit the one-lane result. Next nonzero mask words come from a per-pixel summary word, so no
// lane scans empty words. The record ordinals of round r + 2 are loaded
// while round r runs, the records of round r + 1 gathered. Needs cb < ce.
template <int kEdge, int kSlabRows, int kLanes, int kBatch, bool kRecord, bool kPeriodic>
GS_DEV PixelSum blend_heavy_group(const FrameParams& fp, uint32_t tile, uint32_t slab,
                                  uint32_t cb, uint32_t ce, const uint2* __restrict__ sorted,
                                  const Splat* __restrict__ splat,
                                  RasterLds<kEdge * kSlabRows, kBatch>& L,
                                  uint32_t& x, uint32_t& y, bool& active,
                                  uint32_t* __restrict__ record) {
  static_assert(kLanes == 2 || kLanes == 4, "pairs or quads of lanes per pixel");
  constexpr int kPixels = kEdge * kSlabRows;
  constexpr int kThreads = kLanes * kPixels;
  constexpr int kPer = kBatch / kThreads;  // records staged per thread per round
  constexpr int kMaskWords = kBatch / 32;
  static_assert(kBatch % kThreads == 0, "whole records per thread");
  static_assert(kMaskWords <= 32, "the summary word covers 32 mask words");
  const uint32_t t = threadIdx.x, q = t & (uint32_t)(kLanes - 1), pix = t >> (kLanes / 2);
  int tx0 = (int)((tile % fp.tiles_x) * kEdge);
  int sy0 = (int)((tile / fp.tiles_x) * kEdge + slab * kSlabRows);  // first row of the slab
  x = (uint32_t)tx0 + (pix & (uint32_t)(kEdge - 1));
  y = (uint32_t)sy0 + (pix >> log2_edge(kEdge));
  active = x < fp.width && y < fp.height;
  float pxc = (float)x + 0.5f, pyc = (float)y + 0.5f;
  float vw = fp.viewport[0], vh = fp.viewport[1];
  // The plain policy's instance carries no boundary arithmetic.
  const uint32_t axis = kPeriodic ? fp.boundary_axis : 0u;
  float threshold = fp.transmittance_threshold;
  PixelSum s = {0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0u, 0u};
  constexpr int kChannels = 4 / kLanes;  // of r, g, b, depth, accumulated by this lane
  float acc[kChannels];
#pragma unroll
  for (int h = 0; h < kChannels; h++) acc[h] = 0.0f;
  bool done = !active;
  // Per staged record: its ordinal two rounds ahead, its splat one round
  // ahead. Slots past the unit load the unit's last record (and are never
  // staged): unconditional loads land in the loop-carried registers, where
  // a conditional one makes the compiler wait for the gather as soon as it
  // is issued (to move its result at the join).
  const uint32_t last = ce - 1u;
  uint32_t cg[kPer], cg_next[kPer];
  float4 m[kPer], c[kPer], k[kPer], sp[kPer];
#pragma unroll
  for (int e = 0; e < kPer; e++) {
    uint32_t r = cb + e * kThreads + t;
    cg[e] = sorted[min(r, last)].y;
    gather_splat(splat, cg[e], m[e], c[e], k[e], sp[e]);
    cg_next[e] = sorted[min(r + kBatch, last)].y;
  }
  for (uint32_t batch = cb; batch < ce; batch += kBatch) {
    for (uint32_t i = t; i < (uint32_t)(kMaskWords * kPixels); i += kThreads)
      (&L.mask[0][0])[i] = 0u;
    __syncthreads();
    float4 cm[kPer], cc[kPer];
    Rect crc[kPer];
    bool hit[kPer];
#pragma unroll
    for (int e = 0; e < kPer; e++) {
      uint32_t slot = e * kThreads + t;
      hit[e] = batch + slot < ce;
      crc[e] = {0, -1, 0, -1};
      if (hit[e]) {
        crc[e] = record_rect<kEdge>(m[e], sp[e], tx0, sy0, sy0 + kSlabRows - 1, axis);
        hit[e] = crc[e].x0 <= crc[e].x1 && crc[e].y0 <= crc[e].y1;
      }
      if (hit[e]) {
        L.mean[slot] = m[e];
        L.conic[slot] = c[e];
        L.color[slot] = make_float4(k[e].x, k[e].y, k[e].z, m[e].z);  // r, g, b, depth
        if (kRecord) L.gid[slot] = cg[e];
      }
      cm[e] = m[e];
      cc[e] = c[e];
      cg[e] = cg_next[e];
      gather_splat(splat, cg[e], m[e], c[e], k[e], sp[e]);
      cg_next[e] = sorted[min(batch + 2u * kBatch + slot, last)].y;
    }
#pragma unroll
    for (int e = 0; e < kPer; e++)
      if (hit[e])
        mark_candidates<kEdge>(L, e * kThreads + t, crc[e], tx0, sy0, cm[e], cc[e], vw, vh, axis);
    __syncthreads();
    if (!done) {  // uniform over the group
      uint32_t sum = 0u;
      for (uint32_t w = q; w < (uint32_t)kMaskWords; w += (uint32_t)kLanes)
        if (L.mask[w][pix] != 0u) sum |= 1u << w;
      sum = group_or<kLanes>(sum);
      uint32_t n = min((uint32_t)kBatch, ce - batch);
      uint32_t stop = n;
      uint32_t w = 0u, bits = 0u;
      while (true) {
        uint32_t j[kLanes];
        bool v[kLanes];
#pragma unroll
        for (int qq = 0; qq < kLanes; qq++) {
          if (bits == 0u && sum != 0u) {
            w = (uint32_t)__builtin_ctz(sum);
            sum &= sum - 1u;
            bits = L.mask[w][pix];
          }
          v[qq] = bits != 0u;
          j[qq] = v[qq] ? w * 32u + (uint32_t)__builtin_ctz(bits) : (qq == 0 ? 0u : j[0]);
          bits &= bits - 1u;
        }
        if (!v[0]) break;
        uint32_t mj = j[0];
        bool mv = v[0];
#pragma unroll
        for (int qq = 1; qq < kLanes; qq++) {
          mj = q == (uint32_t)qq ? j[qq] : mj;
          mv = q == (uint32_t)qq ? v[qq] : mv;
        }
        // This lane's channels of the group's candidates (indices of absent
        // candidates stay inside the round; their values are not used).
        float ch[kLanes][kChannels];
#pragma unroll
        for (int qq = 0; qq < kLanes; qq++) {
          const float* src = reinterpret_cast<const float*>(&L.color[j[qq]]) + q * kChannels;
#pragma unroll
          for (int h = 0; h < kChannels; h++) ch[qq][h] = src[h];
        }
        float a = 0.0f;
        if (mv) {
          float4 gm = L.mean[mj];
          float power = pair_power(pxc, pyc, gm, L.conic[mj], vw, vh, axis);
          a = gs_clampf(gm.w * gs_expf_nonpos(power), 0.0f, 0.99f);
        }
        float qa[kLanes];
        group_bcast_all<kLanes>(a, qa);
        // The blends in order. First assuming none of them ends the pixel:
        // a skipped candidate enters with alpha 0, which leaves every value
        // as it was (its colour is finite: absent candidates read the
        // step's first one). T only falls, so if it ends at or above the
        // threshold no blend of the step crossed it; otherwise the step is
        // applied again from the values before it, as selects that stop at
        // the crossing blend (the same operations on the same values).
        const float t_in = s.t;
        const uint32_t blended_in = s.blended;
        float acc_in[kChannels];
#pragma unroll
        for (int h = 0; h < kChannels; h++) acc_in[h] = acc[h];
        if (!kRecord) {
#pragma unroll
          for (int qq = 0; qq < kLanes; qq++) {
            const bool take = v[qq] && !(qa[qq] < (float)(1.0 / 255.0));
            const float alpha = take ? qa[qq] : 0.0f;
            const float contribution = s.t * alpha;
#pragma unroll
            for (int h = 0; h < kChannels; h++) acc[h] = acc[h] + ch[qq][h] * contribution;
            s.t = s.t * (1.0f - alpha);
            s.blended += take ? 1u : 0u;
          }
        }
        if (kRecord || s.t < threshold) {
          s.t = t_in;
          s.blended = blended_in;
#pragma unroll
          for (int h = 0; h < kChannels; h++) acc[h] = acc_in[h];
#pragma unroll
          for (int qq = 0; qq < kLanes; qq++) {
            const float alpha = qa[qq];
            const bool take = v[qq] && !done && !(alpha < (float)(1.0 / 255.0));
            const float contribution = s.t * alpha;
            if (kRecord && take && q == (uint32_t)qq)
              atomicMax(&record[L.gid[j[qq]]], gs_f2b(contribution));
#pragma unroll
            for (int h = 0; h < kChannels; h++) {
              const float sum = acc[h] + ch[qq][h] * contribution;
              acc[h] = take ? sum : acc[h];
            }
            const float nt = s.t * (1.0f - alpha);
            s.blended += take ? 1u : 0u;
            s.t = take ? nt : s.t;
            const bool fin = take && nt < threshold;
            stop = fin ? j[qq] + 1u : stop;
            done = done || fin;
          }
        }
        if (done) break;
      }
      s.checked += stop;  // pairs the reference's loop evaluates
    }
    if (__syncthreads_and(done)) break;
  }
  // Lane 0 of each group collects the channels of the other lanes.
  if constexpr (kLanes == 4) {
    s.r = acc[0];
    s.g = group_bcast<4, 1>(acc[0]);
    s.b = group_bcast<4, 2>(acc[0]);
    s.d = group_bcast<4, 3>(acc[0]);
  } else {
    s.r = acc[0];
    s.g = acc[1];
    s.b = group_bcast<2, 1>(acc[0]);
    s.d = group_bcast<2, 1>(acc[1]);
  }
  return s;
}

GS_DEV void add_stats(DeviceState* state, uint32_t checked, uint32_t blended, bool heavy) {
  unsigned long long c64 = checked, b64 = blended;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    c64 += __shfl_xor(c64, d, 64);
    b64 += __shfl_xor(b64, d, 64);
  }
  if (lane_id() == 0) {
    atomicAdd(&state->checked, c64);
    atomicAdd(&state->blended, b64);
    if (heavy) atomicAdd(&state->checked_heavy, c64);
  }
}

// storeFinalPixel (TileRasterizer.ts:930-964) + finalizeGaussianDepth.ts:18-68
// with the default depth node (GaussianContextNodes.ts:181-189).
GS_DEV void store_pixel(const FrameParams& fp, uint32_t x, uint32_t y, float r, float g, float b,
                        float d, float t, uint16_t* __restrict__ color_out,
                        float* __restrict__ depth_out) {
  float bga = gs_clampf(fp.background[3], 0.0f, 1.0f);
  float o0 = r + fp.background[0] * t * bga;
  float o1 = g + fp.background[1] * t * bga;
  float o2 = b + fp.background[2] * t * bga;
  float o3 = 1.0f - t * (1.0f - bga);
  size_t p = (size_t)y * fp.width + x;
  uint2 packed = make_uint2(gs_float_to_half(o0) | (gs_float_to_half(o1) << 16),
                            gs_float_to_half(o2) | (gs_float_to_half(o3) << 16));
  reinterpret_cast<uint2*>(color_out)[p] = packed;
  if (depth_out) {
    float galpha = 1.0f - t;
    float od = 0.0f;
    if (galpha > 0.0f) od = d / galpha;
    bool scalar_valid = galpha > 0.0f && d == d && gs_fabsf(d) <= kMaxF32 && od == od &&
                        gs_fabsf(od) <= kMaxF32 && od > 0.0f;
    bool valid = scalar_valid && galpha > fp.depth_alpha_threshold;
    float dev = fp.invalid_depth_value;
    if (valid) {
      float near = fp.camera->near_plane, far = fp.camera->far_plane;
      float vz = -od;
      float enc = gs_clampf((near + vz) * far / ((far - near) * vz), 0.0f, 1.0f);
      if (enc == enc && gs_fabsf(enc) <= kMaxF32) dev = enc;
    }
    depth_out[p] = dev;
  }
}

// Light tiles (up to kLightMax records): a grid of at most kLightBlocks
// blocks walks the tiles in strides (most tiles of a frame can be empty, and
// a block per tile would spend the raster stage dispatching them). An empty
// tile stores the background directly.
constexpr uint32_t kLightBlocks = 2048;

// One light tile (at most kLightMax records) with one thread per pixel, or
// the background for an empty tile.
template <int kEdge, int kBatch, bool kRecord>
GS_DEV void light_tile(const FrameParams& fp, DeviceState* __restrict__ state, uint32_t tile,
                       uint32_t begin, uint32_t end, const uint2* __restrict__ sorted,
                       const Splat* __restrict__ splat, RasterLds<kEdge * kEdge, kBatch>& L,
                       uint16_t* __restrict__ color_out, float* __restrict__ depth_out,
                       uint32_t* __restrict__ record) {
  constexpr uint32_t kPixels = kEdge * kEdge;
  if (begin == end) {
    uint32_t t = threadIdx.x;
    uint32_t x = (tile % fp.tiles_x) * kEdge + (t & (uint32_t)(kEdge - 1));
    uint32_t y = (tile / fp.tiles_x) * kEdge + ((t >> log2_edge(kEdge)) & (uint32_t)(kEdge - 1));
    if (t < kPixels && x < fp.width && y < fp.height)
      store_pixel(fp, x, y, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, color_out, depth_out);
    return;
  }
  uint32_t x, y;
  bool active;
  PixelSum s = blend_tile_range<kEdge, light_threads(kEdge, kBatch), kBatch, kRecord>(
      fp, tile, begin, end, sorted, splat, L, x, y, active, record);
  if (fp.raster_stats) add_stats(state, s.checked, s.blended, false);
  if (active) store_pixel(fp, x, y, s.r, s.g, s.b, s.d, s.t, color_out, depth_out);
}

// Light tiles, beside the whole-tile heavy kernel: the grid walks the tiles
// in strides. In frames whose heavy units run with two lanes per pixel
// (edge 16: two-lane whole tiles or row slabs) that shape's kernel takes the
// light tiles and the empty ones too, and this kernel exits at once (as
// does the four-lane whole-tile kernel), ahead of the two-lane kernel whose
// persistent grid takes every CU.
struct RasterArgs {
  FrameParams fp;
  DeviceState* state;
  const uint2* sorted;
  const uint32_t* off;
  const Splat* splat;
  const uint2* units;
  const uint32_t* chunk_first;
  float4* partials;
  float* partial_depth;
  uint16_t* color_out;
  float* depth_out;
  uint32_t* record;
};

// Register budget: 5 waves per SIMD (96 VGPRs, a few spilled outside the
// blend loop) instead of the compiler's 113 VGPRs at 4 waves: the 1M/1080p
// frame 0.422 -> 0.416 ms, where this kernel runs beside the whole-tile one.
template <int kEdge, int kBatch, bool kRecord, int kN>
__global__ __launch_bounds__(light_threads(kEdge, kBatch)) __attribute__((amdgpu_waves_per_eu(5, 8)))
void k_raster_light(
    const ViewArgs<RasterArgs, kN> grp) {
  __shared__ RasterLds<kEdge * kEdge, kBatch> lds;
  for (uint32_t view = 0; view < (kN == 1 ? 1u : grp.n); view++) {
  const RasterArgs& a = grp.v[view];
  const FrameParams& fp = a.fp;
  DeviceState* __restrict__ state = a.state;
  const uint2* __restrict__ sorted = a.sorted;
  const uint32_t* __restrict__ off = a.off;
  const Splat* __restrict__ splat = a.splat;
  uint16_t* __restrict__ color_out = a.color_out;
  float* __restrict__ depth_out = a.depth_out;
  uint32_t* __restrict__ record = a.record;
  if (state->overflow) continue;  // keep the previous image (intersections.ts:114-134)
  if (kEdge == 16 && state->heavy_split != 0u) continue;
  for (uint32_t tile = blockIdx.x; tile < fp.tile_count; tile += gridDim.x) {
    uint32_t begin = off[tile], end = off[tile + 1];
    // Heavy units and chunked tiles (chunk sizes below kLightMax make some
    // tiles of at most kLightMax records chunked) are not light (block-uniform).
    if (end - begin > kLightMax || (fp.chunk_size != 0u && end - begin > fp.chunk_size))
      continue;
    light_tile<kEdge, kBatch, kRecord>(fp, state, tile, begin, end, sorted, splat, lds,
                                       color_out, depth_out, record);
  }
  }
}

// rasterChunks.ts:3-65 in one block, extended into the heavy work list:
// first the chunk tasks of tiles above chunk_size (the reference's exact
// chunks, each blended from T = 1 and composited later; chunk_first holds
// each tile's first task), then the remaining tiles above kLightMax,
// heaviest size class first. Units are (tile, chunk) with chunk = ~0u for a
// whole tile.
struct HeavyUnitArgs {
  FrameParams fp;
  DeviceState* state;
  const uint32_t* off;
  uint32_t* chunk_first;
  uint2* units;
  uint32_t unit_capacity;
};

template <int kN>
__global__ __launch_bounds__(1024) void k_heavy_units(const ViewArgs<HeavyUnitArgs, kN> grp) {
  const HeavyUnitArgs& a = grp.v[kN == 1 ? 0u : blockIdx.x];
  const FrameParams& fp = a.fp;
  DeviceState* __restrict__ state = a.state;
  const uint32_t* __restrict__ off = a.off;
  uint32_t* __restrict__ chunk_first = a.chunk_first;
  uint2* __restrict__ units = a.units;
  const uint32_t unit_capacity = a.unit_capacity;
  // Tiles are visited in strides of the block; the waves reserve their chunk
  // tasks with one LDS add each, so no iteration waits on another (a tile's
  // chunks are contiguous and start at chunk_first[tile]; the order between
  // tiles is free). Heavy tiles then take slots after all chunk tasks,
  // heaviest size class first.
  __shared__ uint32_t chunk_total, light_total;
  __shared__ uint32_t class_count[5], class_next[5];
  if (threadIdx.x == 0) chunk_total = light_total = 0u;
  if (threadIdx.x < 5) class_count[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t tiles = state->overflow ? 0u : fp.tile_count;
  const uint32_t cs = fp.chunk_size, lane = lane_id();
#pragma unroll 4
  for (uint32_t base = threadIdx.x & ~63u; base < tiles; base += 1024u) {
    uint32_t t = base + lane;
    uint32_t n = 0u;
    if (t < tiles) {
      uint32_t count = off[t + 1] - off[t];
      if (cs != 0 && count > cs) {
        n = (count + cs - 1) / cs;
      } else if (count > kLightMax) {
        uint32_t c = 31u - (uint32_t)__builtin_clz((count - 1u) >> 10);
        atomicAdd(&class_count[min(c, 4u)], 1u);
      }
    }
    uint32_t incl = wave_inclusive_scan(n);
    uint32_t wave_sum = __shfl(incl, 63, 64);
    uint32_t first = 0u;
    if (wave_sum != 0u) {  // uniform
      uint32_t wbase = 0u;
      if (lane == 0) wbase = atomicAdd(&chunk_total, wave_sum);
      first = __shfl(wbase, 0, 64) + incl - n;
    }
    if (t < tiles) chunk_first[t] = first;
    for (uint32_t c = 0; c < n; c++)
      if (first + c < unit_capacity) units[first + c] = make_uint2(t, c);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t next = chunk_total;
    state->chunk_tasks = min(chunk_total, unit_capacity);
    for (int c = 4; c >= 0; c--) {
      class_next[c] = next;
      next += class_count[c];
    }
    state->heavy_units = min(next, unit_capacity);
    state->heavy_split = next <= fp.split_units       ? (uint32_t)kShapeWhole
                         : next > fp.fine_split_units ? (uint32_t)kShapeSlab4
                         : next > fp.pair_units       ? (uint32_t)kShapeSlab8
                                                      : (uint32_t)kShapePair;
    for (int x = 0; x < 8; x++) state->heavy_next[x] = 0u;
  }
  __syncthreads();
  // Two-lane frames (tile edge 16) also list the non-empty light tiles that
  // are not chunked, after the heavy units, for the two-lane kernel.
  const bool list_light = fp.tile_size == 16 && state->heavy_split != 0u;
  const uint32_t heavy_total = state->heavy_units;
#pragma unroll 4
  for (uint32_t t = threadIdx.x; t < tiles; t += 1024u) {
    uint32_t count = off[t + 1] - off[t];
    bool chunked = cs != 0 && count > cs;
    if (!chunked && count > kLightMax) {
      uint32_t c = min(31u - (uint32_t)__builtin_clz((count - 1u) >> 10), 4u);
      uint32_t slot = atomicAdd(&class_next[c], 1u);
      if (slot < unit_capacity) units[slot] = make_uint2(t, ~0u);
    } else if (list_light && !chunked && count != 0u) {
      uint32_t slot = heavy_total + atomicAdd(&light_total, 1u);
      if (slot < unit_capacity) units[slot] = make_uint2(t, ~0u);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t n = list_light ? light_total : 0u;
    state->light_units = min(n, unit_capacity - heavy_total);
  }
}

// Heavy units: per chunk task or heavy tile, S = heavy_slabs(E, kShape)
// work items, one per row slab (S = 1 for whole tiles, with four or two
// lanes per pixel); the launches of every shape run and those
// the frame did not choose (DeviceState::heavy_split, set by
// k_heavy_units) exit at once. The grid is persistent (as many blocks
// as fit on the GPU at once, a multiple of 8). Workgroups go to the eight
// XCDs in turn, so block b works for XCD x = b % 8: its items are
// k = 0, 1, 2, ... of that XCD's sequence, where item k is unit
// (k / S) * 8 + x, slab k % S. A unit's slabs therefore run on one XCD and
// share its L2, and units start in list order (heaviest first). The
// two-lane kernels then take the light tiles listed after the heavy units,
// as items of the same kind, and last the background of the empty tiles, 64
// tiles per item. A block starts with item b / 8 and then takes the XCD's
// next unclaimed item (one atomic per item), so blocks that finish early
// keep working.
// Register budget: 6 waves per SIMD (at most 80 VGPRs, no spills; the row
// slabs' 4- and 2-wave blocks took 81-85 at 5 waves); the four-lane
// whole-tile block (16 waves) fills its CU's LDS and runs at 4 waves per
// SIMD either way.
template <int kEdge, int kShape, bool kRecord, int kN>
__global__ __launch_bounds__(heavy_threads(kEdge, kShape)) __attribute__((amdgpu_waves_per_eu(6, 8)))
void k_raster_heavy(
    const ViewArgs<RasterArgs, kN> grp) {
  constexpr int kPixels = kEdge * kEdge;
  constexpr bool kGroup = kEdge != 32;  // lane groups per pixel (edges 8, 16)
  constexpr bool kSplit = kShape != kShapeWhole;  // two-lane shapes, which also take the light tiles
  constexpr uint32_t kSlabs = heavy_slabs(kEdge, kShape);
  constexpr int kBatch = heavy_batch(kEdge, kShape);
  __shared__ RasterLds<heavy_slab_pixels(kEdge, kShape), kBatch> lds;
  __shared__ uint32_t next_item;
  for (uint32_t view = 0; view < (kN == 1 ? 1u : grp.n); view++) {
  const RasterArgs& a = grp.v[view];
  const FrameParams& fp = a.fp;
  DeviceState* __restrict__ state = a.state;
  const uint2* __restrict__ sorted = a.sorted;
  const uint32_t* __restrict__ off = a.off;
  const Splat* __restrict__ splat = a.splat;
  const uint2* __restrict__ units = a.units;
  float4* __restrict__ partials = a.partials;
  float* __restrict__ partial_depth = a.partial_depth;
  uint16_t* __restrict__ color_out = a.color_out;
  float* __restrict__ depth_out = a.depth_out;
  uint32_t* __restrict__ record = a.record;
  if (kEdge == 16 && state->heavy_split != (uint32_t)kShape) continue;
  const uint32_t xcd = blockIdx.x & 7u, per_xcd = gridDim.x >> 3;
  const uint32_t n_heavy = state->heavy_units;
  const uint32_t n_light = kSplit ? state->light_units : 0u;
  const uint32_t heavy_items = ((n_heavy + 7u) >> 3) * kSlabs;  // per XCD
  const uint32_t unit_items = heavy_items + ((n_light + 7u) >> 3) * kSlabs;
  constexpr uint32_t kBgTiles = 64;  // empty tiles per background item
  const uint32_t tiles = (kSplit && !state->overflow) ? fp.tile_count : 0u;
  const uint32_t bg_groups = (tiles + kBgTiles - 1u) / kBgTiles;
  const uint32_t items = unit_items + ((bg_groups + 7u) >> 3);
  for (uint32_t item = blockIdx.x >> 3; item < items;) {
    if (item >= unit_items) {
      // Background of the empty tiles of group (item - unit_items) * 8 + x.
      const uint32_t g = (item - unit_items) * 8u + xcd;
      for (uint32_t i = 0; i < kBgTiles && g < bg_groups; i++) {
        uint32_t tile = g * kBgTiles + i;
        if (tile >= tiles) break;
        if (off[tile + 1] != off[tile]) continue;  // block-uniform
        for (uint32_t px = threadIdx.x; px < (uint32_t)kPixels; px += blockDim.x) {
          uint32_t x = (tile % fp.tiles_x) * kEdge + (px & (uint32_t)(kEdge - 1));
          uint32_t y = (tile / fp.tiles_x) * kEdge + (px >> log2_edge(kEdge));
          if (x < fp.width && y < fp.height)
            store_pixel(fp, x, y, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, color_out, depth_out);
        }
      }
    } else {
      const bool light = item >= heavy_items;
      const uint32_t k = light ? item - heavy_items : item;
      const uint32_t slab = k % kSlabs;
      const uint32_t u = (light ? n_heavy : 0u) + (k / kSlabs) * 8u + xcd;
      if (u < n_heavy + (light ? n_light : 0u)) {
        uint2 unit = units[u];
        uint32_t tile = unit.x;
        uint32_t begin = off[tile], end = off[tile + 1];
        uint32_t cb = begin, ce = end;
        bool chunk = unit.y != ~0u;
        if (chunk) {
          cb = begin + unit.y * fp.chunk_size;
          ce = min(cb + fp.chunk_size, end);
        }
        uint32_t x, y;
        bool active;
        PixelSum s;
        uint32_t pix = threadIdx.x;  // pixel within the tile
        bool writer = threadIdx.x < (uint32_t)kPixels;
        if constexpr (kGroup) {
          constexpr int kLanes = heavy_lanes(kEdge, kShape);
          constexpr int kRows = heavy_slab_rows(kEdge, kShape);
          if (fp.boundary_axis == 0u)
            s = blend_heavy_group<kEdge, kRows, kLanes, kBatch, kRecord, false>(
                fp, tile, slab, cb, ce, sorted, splat, lds, x, y, active, record);
          else
            s = blend_heavy_group<kEdge, kRows, kLanes, kBatch, kRecord, true>(
                fp, tile, slab, cb, ce, sorted, splat, lds, x, y, active, record);
          pix = slab * heavy_slab_pixels(kEdge, kShape) + (threadIdx.x >> (kLanes / 2));
          writer = (threadIdx.x & (uint32_t)(kLanes - 1)) == 0u;
        } else {
          s = blend_tile_range<kEdge, kHeavyThreads, kBatch, kRecord>(
              fp, tile, cb, ce, sorted, splat, lds, x, y, active, record);
        }
        if (fp.raster_stats)
          add_stats(state, writer ? s.checked : 0u, writer ? s.blended : 0u, !light);
        if (writer) {
          if (chunk) {
            size_t p = (size_t)u * kPixels + pix;  // chunk tasks lead the list
            partials[p] = make_float4(s.r, s.g, s.b, s.t);
            partial_depth[p] = s.d;
          } else if (active) {
            store_pixel(fp, x, y, s.r, s.g, s.b, s.d, s.t, color_out, depth_out);
          }
        }
      }
    }
    __syncthreads();  // the LDS and next_item are free again
    if (threadIdx.x == 0) next_item = per_xcd + atomicAdd(&state->heavy_next[xcd], 1u);
    __syncthreads();
    item = next_item;
  }
  }
}

// Persistent grid of k_raster_heavy, a multiple of 8 blocks (at least 8):
// whole-tile blocks (16 waves) one per CU, leaving room for the light-tile
// kernel beside them; the two-lane shapes' blocks (8, 4 or 2 waves) 24
// waves per CU, the waves their register budget allows (measured faster
// than leaving the light kernel room: their units outnumber the CUs and
// gain most from occupancy).
template <int kEdge, int kShape>
inline uint32_t heavy_grid(uint32_t cu_count) {
  uint32_t waves = (uint32_t)heavy_threads(kEdge, kShape) / 64u;
  uint32_t per_cu = waves >= 16u ? 1u : 24u / waves;
  uint32_t g = (cu_count * per_cu) & ~7u;
  return g < 8u ? 8u : g;
}

// TileRasterizer.ts:807-902: ordered over-composite of a tile's chunks.
template <int kEdge, int kN>
__global__ __launch_bounds__(kEdge * kEdge) void k_composite(const ViewArgs<RasterArgs, kN> grp) {
  const uint32_t view = kN == 1 ? 0u : blockIdx.x / grp.bpv;
  const RasterArgs& a = grp.v[view];
  const FrameParams& fp = a.fp;
  const DeviceState* __restrict__ state = a.state;
  const uint32_t* __restrict__ off = a.off;
  const uint32_t* __restrict__ chunk_first = a.chunk_first;
  const float4* __restrict__ partials = a.partials;
  const float* __restrict__ partial_depth = a.partial_depth;
  uint16_t* __restrict__ color_out = a.color_out;
  float* __restrict__ depth_out = a.depth_out;
  constexpr int kPixels = kEdge * kEdge;
  if (state->overflow) return;
  uint32_t tile = blockIdx.x - view * grp.bpv;
  if (tile >= fp.tile_count || fp.chunk_size == 0) return;
  uint32_t count = off[tile + 1] - off[tile];
  uint32_t cs = fp.chunk_size;
  if (count <= cs) return;
  uint32_t n = (count + cs - 1) / cs;
  uint32_t t = threadIdx.x;
  uint32_t x = (tile % fp.tiles_x) * kEdge + (t & (uint32_t)(kEdge - 1));
  uint32_t y = (tile / fp.tiles_x) * kEdge + (t >> log2_edge(kEdge));
  if (x >= fp.width || y >= fp.height) return;
  float r = 0.0f, g = 0.0f, b = 0.0f, d = 0.0f, tr = 1.0f;
  uint32_t first = chunk_first[tile];
  for (uint32_t k = 0; k < n; k++) {
    size_t p = (size_t)(first + k) * kPixels + t;
    float4 c = partials[p];
    r = r + c.x * tr;
    g = g + c.y * tr;
    b = b + c.z * tr;
    d = d + partial_depth[p] * tr;
    tr = tr * c.w;
    if (tr < fp.transmittance_threshold) break;
  }
  store_pixel(fp, x, y, r, g, b, d, tr, color_out, depth_out);
}

// --------------------------------------------------------------- driver ----

namespace {

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// A sort of one view of a group: its ping-pong buffers, key count, workspace
// and state.
struct SortView {
  uint2* a;
  uint2* b;
  const uint32_t* count_ptr;
  SortWorkspace ws;
  DeviceState* state;
  uint32_t* in_place;
};

// The first view's arguments of a group, for the kN = 1 instances.
template <class A>
ViewArgs<A, 1> first_view(const ViewArgs<A>& g) {
  ViewArgs<A, 1> o{};
  o.v[0] = g.v[0];
  o.n = 1;
  o.bpv = g.bpv;
  return o;
}

template <class A>
ViewArgs<A> view_args(uint32_t n, uint32_t bpv) {
  ViewArgs<A> g{};
  g.n = n;
  g.bpv = bpv ? bpv : 1u;
  return g;
}

// Sort pairs in a <-> b for `passes` kBits-bit digits; the result lands in
// a for an even pass count, in b otherwise. Every view of the group runs
// the same passes (a pass whose digit is zero for every key only copies).
template <int kBits, int kItems>
void sort_passes(const SortView* sv, uint32_t n, uint32_t max_count, uint32_t passes,
                 hipStream_t s) {
  uint32_t blocks = div_up(max_count, (uint64_t)kSortThreads * kItems);
  if (blocks == 0) blocks = 1;
  for (uint32_t p = 0; p < passes; p++) {
    ViewArgs<SortPassArgs> grp = view_args<SortPassArgs>(n, blocks);
    for (uint32_t v = 0; v < n; v++)
      grp.v[v] = {(p % 2 == 0) ? sv[v].a : sv[v].b, (p % 2 == 0) ? sv[v].b : sv[v].a,
                sv[v].count_ptr, p, sv[v].ws, sv[v].state,
                p + 1 == passes ? sv[v].in_place : nullptr};
    if (n == 1)
      k_sort_pass<kBits, kItems, 1><<<blocks, kSortThreads, 0, s>>>(first_view(grp));
    else
      k_sort_pass<kBits, kItems, kGroupMax><<<n * blocks, kSortThreads, 0, s>>>(grp);
  }
}

template <int kBits>
void radix_sort_bits(const SortView* sv, uint32_t n, uint32_t max_count, uint32_t passes,
                     hipStream_t s) {
  uint32_t parts = div_up(max_count, kSortPartition);
  if (parts == 0) parts = 1;
  uint32_t hist_blocks = parts / 4u;
  hist_blocks = hist_blocks < kHistBlocksMin ? kHistBlocksMin
                                             : (hist_blocks > kHistBlocksMax ? kHistBlocksMax
                                                                             : hist_blocks);
  ViewArgs<HistArgs> grp = view_args<HistArgs>(n, hist_blocks);
  for (uint32_t v = 0; v < n; v++) grp.v[v] = {sv[v].a, sv[v].count_ptr, passes, sv[v].ws};
  if (n == 1)
    k_sort_histograms<kBits, 1><<<hist_blocks, kBlock, 0, s>>>(first_view(grp));
  else
    k_sort_histograms<kBits, kGroupMax><<<n * hist_blocks, kBlock, 0, s>>>(grp);
  if (max_count >= kLargeSort)
    sort_passes<kBits, 2 * kSortItems>(sv, n, max_count, passes, s);
  else
    sort_passes<kBits, kSortItems>(sv, n, max_count, passes, s);
}

// Sort pairs in a <-> b (see radix_sort_bits) for each view of a group,
// sized for max_count keys per view. With in_place, a last pass whose digit
// is one value for every key of its view is skipped and *in_place records
// it; that view's sorted pairs are then in the last pass's input buffer.
void radix_sort(const SortView* sv, uint32_t n, uint32_t max_count, uint32_t passes,
                uint32_t bits, hipStream_t s) {
  if (bits == 4)
    radix_sort_bits<4>(sv, n, max_count, passes, s);
  else if (bits == 6)
    radix_sort_bits<6>(sv, n, max_count, passes, s);
  else
    radix_sort_bits<8>(sv, n, max_count, passes, s);
}

template <int kEdge, int kShape, bool kRecord>
void launch_heavy(const ViewArgs<RasterArgs>& ra, uint32_t cu_count, hipStream_t s) {
  const uint32_t grid = heavy_grid<kEdge, kShape>(cu_count);
  if (ra.n == 1)
    k_raster_heavy<kEdge, kShape, kRecord, 1>
        <<<grid, heavy_threads(kEdge, kShape), 0, s>>>(first_view(ra));
  else
    k_raster_heavy<kEdge, kShape, kRecord, kGroupMax>
        <<<grid, heavy_threads(kEdge, kShape), 0, s>>>(ra);
}

// Tile records of view b after a tile sort of `passes` passes.
const uint2* tile_records_after(const FrameBuffers& b, uint32_t passes) {
  return (passes % 2 == 0) ? b.records_a : b.records_b;
}

// The raster stage of a group: the heavy work lists on s, then heavy units
// on aux concurrently with light tiles on s, then the chunk composite.
template <int kEdge, int kBatch, bool kRecord>
hipError_t launch_raster(const FrameBuffers* views, uint32_t n, hipStream_t s,
                         const RasterStreams& rs, hipEvent_t* marks) {
  const FrameBuffers& b0 = views[0];
  uint32_t max_tiles = 1;
  ViewArgs<HeavyUnitArgs> hu = view_args<HeavyUnitArgs>(n, 1);
  ViewArgs<RasterArgs> ra = view_args<RasterArgs>(n, 1);
  for (uint32_t v = 0; v < n; v++) {
    const FrameBuffers& b = views[v];
    max_tiles = std::max(max_tiles, b.max_tiles);
    hu.v[v] = {b.params, b.state, b.tile_offsets, b.chunk_first, b.tasks, b.unit_capacity};
    ra.v[v] = {b.params,      b.state,          tile_records_after(b, b.tile_passes),
               b.tile_offsets, b.proj.splat,    b.tasks,
               b.chunk_first,  b.chunk_partials, b.partial_depth,
               b.color_out,    b.depth_out,      b.record};
  }
  if (n == 1)
    k_heavy_units<1><<<1, 1024, 0, s>>>(first_view(hu));
  else
    k_heavy_units<kGroupMax><<<n, 1024, 0, s>>>(hu);
  // Without a second stream (rs.aux null) the heavy kernels run on s, ahead
  // of the light-tile kernel, and no event is recorded.
  const hipStream_t aux = rs.aux ? rs.aux : s;
  hipError_t e = hipSuccess;
  if (rs.aux) {
    e = hipEventRecord(rs.fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(rs.aux, rs.fork, 0);
    if (e != hipSuccess) return e;
  }
  // The whole-tile kernel on the second branch; on s the light tiles, then
  // (edge 16) the three two-lane shapes. Each kernel takes the views of its
  // shape, and one whose views all chose another shape exits at once. So in
  // whole-tile frames the other shapes' exits follow the light tiles (which
  // finish before the heavy units) instead of delaying the join, and in
  // two-lane frames the whole-tile and light launches exit beside or ahead
  // of the working kernel.
  if (marks) (void)hipEventRecord(marks[kMarkHeavyBegin], aux);
  launch_heavy<kEdge, kShapeWhole, kRecord>(ra, b0.cu_count, aux);
  if (marks) (void)hipEventRecord(marks[kMarkHeavyBegin + 1], aux);
  if (rs.aux) {
    e = hipEventRecord(rs.join, rs.aux);
    if (e != hipSuccess) return e;
  }
  uint32_t light_blocks = max_tiles < kLightBlocks ? max_tiles : kLightBlocks;
  if (n == 1)
    k_raster_light<kEdge, kBatch, kRecord, 1>
        <<<light_blocks, light_threads(kEdge, kBatch), 0, s>>>(first_view(ra));
  else
    k_raster_light<kEdge, kBatch, kRecord, kGroupMax>
        <<<light_blocks, light_threads(kEdge, kBatch), 0, s>>>(ra);
  // Mark pairs after the whole tiles' (every edge records them; other
  // edges launch nothing between): the 8-row slabs, the 4-row slabs, the
  // two-lane whole tiles.
  if (marks) (void)hipEventRecord(marks[kMarkHeavyBegin + 6], s);
  if constexpr (kEdge == 16) launch_heavy<kEdge, kShapePair, kRecord>(ra, b0.cu_count, s);
  if (marks) (void)hipEventRecord(marks[kMarkHeavyBegin + 7], s);
  if (marks) (void)hipEventRecord(marks[kMarkHeavyBegin + 2], s);
  if constexpr (kEdge == 16) launch_heavy<kEdge, kShapeSlab8, kRecord>(ra, b0.cu_count, s);
  if (marks) (void)hipEventRecord(marks[kMarkHeavyBegin + 3], s);
  if (marks) (void)hipEventRecord(marks[kMarkHeavyBegin + 4], s);
  if constexpr (kEdge == 16) launch_heavy<kEdge, kShapeSlab4, kRecord>(ra, b0.cu_count, s);
  if (marks) (void)hipEventRecord(marks[kMarkHeavyBegin + 5], s);
  if (rs.aux) {
    e = hipStreamWaitEvent(s, rs.join, 0);
    if (e != hipSuccess) return e;
  }
  if (b0.params.chunk_size != 0) {
    ra.bpv = max_tiles;
    if (n == 1)
      k_composite<kEdge, 1><<<max_tiles, kEdge * kEdge, 0, s>>>(first_view(ra));
    else
      k_composite<kEdge, kGroupMax><<<n * max_tiles, kEdge * kEdge, 0, s>>>(ra);
  }
  return hipGetLastError();
}

// The instrumented raster (contribution record) uses the 256-record batch.
template <int kEdge>
hipError_t launch_raster_edge(const FrameBuffers* views, uint32_t n, hipStream_t s,
                              const RasterStreams& rs, hipEvent_t* marks) {
  const FrameBuffers& b = views[0];
  if (b.record) return launch_raster<kEdge, 256, true>(views, n, s, rs, marks);
  if (b.raster_batch == 128) return launch_raster<kEdge, 128, false>(views, n, s, rs, marks);
  return launch_raster<kEdge, 256, false>(views, n, s, rs, marks);
}

}  // namespace

hipError_t launch_group(const FrameBuffers* views, uint32_t n, hipStream_t s,
                        const RasterStreams& rs, hipEvent_t* marks) {
  auto mark = [&](int stage) {
    if (marks) (void)hipEventRecord(marks[stage], s);
  };
  if (n == 0 || n > (uint32_t)kGroupMax) return hipErrorInvalidValue;
  // The views of a group share the launch shape, the mode (plain, cull,
  // collect), the scene and the sort pass counts.
  const FrameBuffers& b0 = views[0];
  uint32_t max_n = 0, max_k = 0, max_tiles = 0;
  for (uint32_t v = 0; v < n; v++) {
    const FrameBuffers& b = views[v];
    if (b.params.tile_size != b0.params.tile_size || b.radix_bits != b0.radix_bits ||
        b.raster_batch != b0.raster_batch || (b.record != nullptr) != (b0.record != nullptr) ||
        (b.cull_net != nullptr) != (b0.cull_net != nullptr) ||
        b.depth_passes != b0.depth_passes || b.tile_passes != b0.tile_passes ||
        b.params.chunk_size != b0.params.chunk_size)
      return hipErrorInvalidValue;
    max_n = std::max(max_n, b.max_n);
    max_k = std::max(max_k, b.max_k);
    max_tiles = std::max(max_tiles, b.max_tiles);
  }
  uint32_t n_blocks = div_up(max_n, kBlock);
  if (n_blocks == 0) n_blocks = 1;
  uint32_t scan_blocks = div_up(max_n, kScanPartition);
  if (scan_blocks == 0) scan_blocks = 1;
  mark(kStageProject);
  {
    ViewArgs<ProjectArgs> grp = view_args<ProjectArgs>(n, n_blocks);
    for (uint32_t v = 0; v < n; v++) {
      const FrameBuffers& b = views[v];
      grp.v[v] = {b.params, b.state, b.scene, b.proj, b.cull_feats, b.record, b.cull_net};
    }
    if (n == 1) {
      if (b0.cull_net)
        k_project_cull<1><<<n_blocks, kBlock, 0, s>>>(first_view(grp));
      else if (b0.record)
        k_project<kModeCollect, 1><<<n_blocks, kBlock, 0, s>>>(first_view(grp));
      else
        k_project<kModePlain, 1><<<n_blocks, kBlock, 0, s>>>(first_view(grp));
    } else {
      if (b0.cull_net)
        k_project_cull<kGroupMax><<<n * n_blocks, kBlock, 0, s>>>(grp);
      else if (b0.record)
        k_project<kModeCollect, kGroupMax><<<n * n_blocks, kBlock, 0, s>>>(grp);
      else
        k_project<kModePlain, kGroupMax><<<n * n_blocks, kBlock, 0, s>>>(grp);
    }
  }
  mark(kStageVisible);

  // Visible flags -> exclusive offsets -> compacted (depth key, ordinal).
  {
    ViewArgs<ScanArgs<LoadVisible, StoreCompact>> grp =
        view_args<ScanArgs<LoadVisible, StoreCompact>>(n, scan_blocks);
    for (uint32_t v = 0; v < n; v++) {
      const FrameBuffers& b = views[v];
      const FrameParams& fp = b.params;
      grp.v[v] = {LoadVisible{b.proj.mean},
                StoreCompact{b.proj.mean, fp.depth_mode, fp.camera, b.visible_offsets, b.depth_a},
                LenSpec{nullptr, 0u, 1u, fp.active_count},
                ScanWorkspace{reinterpret_cast<unsigned long long*>(b.partials)},
                kScanVisible,
                0u,
                fp.camera,
                b.state};
    }
    if (n == 1)
      k_scan<LoadVisible, StoreCompact, 1><<<scan_blocks, kBlock, 0, s>>>(first_view(grp));
    else
      k_scan<LoadVisible, StoreCompact, kGroupMax><<<n * scan_blocks, kBlock, 0, s>>>(grp);
  }

  // Depth sort over V.
  mark(kStageDepthSort);
  SortView sv[kGroupMax];
  for (uint32_t v = 0; v < n; v++) {
    const FrameBuffers& b = views[v];
    DeviceState* st = b.state;
    uint32_t* depth_totals = b.hist;
    uint32_t* tile_totals = b.hist + kMaxPasses * kHistShards * 256;
    uint32_t view_scan_blocks = div_up(b.max_n, kScanPartition);
    sv[v] = {b.depth_a,
             b.depth_b,
             &st->visible,
             {depth_totals, tile_totals, b.hist + kSortTotalsWords, b.sort_parts,
              reinterpret_cast<unsigned long long*>(b.partials),
              view_scan_blocks ? view_scan_blocks : 1u},
             st,
             &st->depth_in_place};
  }
  radix_sort(sv, n, max_n, b0.depth_passes, b0.radix_bits, s);

  // Depth-ordered tile counts -> write offsets; K, overflow, output state.
  mark(kStageScan);
  {
    ViewArgs<ScanArgs<LoadOrdered, StoreOrdered>> grp =
        view_args<ScanArgs<LoadOrdered, StoreOrdered>>(n, scan_blocks);
    for (uint32_t v = 0; v < n; v++) {
      const FrameBuffers& b = views[v];
      DeviceState* st = b.state;
      grp.v[v] = {LoadOrdered{sorted_depth_records(b, 0u), sorted_depth_records(b, 1u),
                            &st->depth_in_place, b.proj.hits, b.ordered_hits},
                StoreOrdered{b.ordered_counts, b.intersection_offsets},
                LenSpec{&st->visible, 0u, 1u, 0u},
                ScanWorkspace{reinterpret_cast<unsigned long long*>(b.partials)},
                kScanIntersections,
                b.params.capacity,
                b.params.camera,
                st};
    }
    if (n == 1)
      k_scan<LoadOrdered, StoreOrdered, 1><<<scan_blocks, kBlock, 0, s>>>(first_view(grp));
    else
      k_scan<LoadOrdered, StoreOrdered, kGroupMax><<<n * scan_blocks, kBlock, 0, s>>>(grp);
  }

  mark(kStageEmit);
  {
    ViewArgs<EmitArgs> grp = view_args<EmitArgs>(n, n_blocks);
    for (uint32_t v = 0; v < n; v++) {
      const FrameBuffers& b = views[v];
      grp.v[v] = {b.params,         b.state,          sorted_depth_records(b, 0u),
                sorted_depth_records(b, 1u), b.proj.splat, b.ordered_hits,
                b.intersection_offsets, b.records_a};
    }
    if (n == 1)
      k_emit<1><<<n_blocks, kBlock, 0, s>>>(first_view(grp));
    else
      k_emit<kGroupMax><<<n * n_blocks, kBlock, 0, s>>>(grp);
  }

  mark(kStageTileSort);
  for (uint32_t v = 0; v < n; v++) {
    const FrameBuffers& b = views[v];
    std::swap(sv[v].ws.ghist, sv[v].ws.ghist_other);
    sv[v].a = b.records_a;
    sv[v].b = b.records_b;
    sv[v].count_ptr = &b.state->count;
    sv[v].in_place = nullptr;
  }
  radix_sort(sv, n, max_k, b0.tile_passes, b0.radix_bits, s);

  mark(kStageTileOffsets);
  {
    uint32_t off_blocks = div_up((uint64_t)max_k + 1, kBlock);
    if (off_blocks > 2048) off_blocks = 2048;
    ViewArgs<TileOffsetArgs> grp = view_args<TileOffsetArgs>(n, off_blocks);
    for (uint32_t v = 0; v < n; v++) {
      const FrameBuffers& b = views[v];
      grp.v[v] = {b.params, b.state, sorted_tile_records(b), b.tile_offsets};
    }
    if (n == 1)
      k_tile_offsets<1><<<off_blocks, kBlock, 0, s>>>(first_view(grp));
    else
      k_tile_offsets<kGroupMax><<<n * off_blocks, kBlock, 0, s>>>(grp);
  }
  mark(kStageRaster);
  hipError_t e;
  if (b0.params.tile_size == 8)
    e = launch_raster_edge<8>(views, n, s, rs, marks);
  else if (b0.params.tile_size == 32)
    e = launch_raster_edge<32>(views, n, s, rs, marks);
  else
    e = launch_raster_edge<16>(views, n, s, rs, marks);
  if (e != hipSuccess) return e;
  mark(kStageCount);
  return hipGetLastError();
}

hipError_t launch_frame(const FrameBuffers& b, hipStream_t s, const RasterStreams& rs,
                        hipEvent_t* marks) {
  return launch_group(&b, 1, s, rs, marks);
}

hipError_t launch_sh_top(const uint32_t* sh, uint32_t n, uint32_t sh_degree, uint8_t* out,
                         hipStream_t s) {
  if (n == 0) return hipSuccess;
  k_sh_top<<<div_up(n, kBlock), kBlock, 0, s>>>(sh, n, sh_degree, out);
  return hipGetLastError();
}

const uint2* sorted_depth_records(const FrameBuffers& b, uint32_t depth_in_place) {
  // Pass p reads a for even p; the result is the last pass's output, or its
  // input when that pass was skipped.
  uint32_t filled = depth_in_place ? b.depth_passes - 1u : b.depth_passes;
  return (filled % 2 == 0) ? b.depth_a : b.depth_b;
}
const uint2* sorted_tile_records(const FrameBuffers& b) {
  return tile_records_after(b, b.tile_passes);
}

}  // namespace gs
