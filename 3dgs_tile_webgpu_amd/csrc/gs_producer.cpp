// gs_producer.cpp - compact/rgb8e8 encoder and the seeded scene generator.
//
// Host code (no device work). The encoder follows the reference's Rust
// producer operation by operation in binary32; see include/gs_producer.h.
#include <math.h>
#include <string.h>

#include <algorithm>
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
