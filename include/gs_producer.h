/* gs_producer.h - packed-buffer producer: the encoder that feeds the
 * GaussianStore contract, and a seeded synthetic scene generator.
 *
 * The encoder restates the reference's Rust producer so synthetic and real
 * scenes reach the frame pipeline in exactly the layout the store publishes:
 *   gs_encode_compact   <- rust/gaussian-backend/src/engine.rs:1801-1832
 *                          (+ model.rs:225-228 render_scale_opacity,
 *                           spark-lib splat_encode.rs:127-136, 234-252)
 *   gs_encode_sh_rgb8e8 <- engine.rs:1857-1872, 1940-1949 (pack_sh)
 * The generator has no reference counterpart; it produces the synthetic
 * workloads (clustered means, log-normal scales, uniform rotations,
 * sigmoid-normal opacity, rgb8e8 SH) that tests and benchmarks render.
 */
#ifndef GS_PRODUCER_H
#define GS_PRODUCER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compact shape + rotation for n Gaussians. scales: n*3 linear scales;
 * opacity: n; quats: n*4 (x, y, z, w), normalised by the encoder.
 * out_shape: n*2 u32; out_rotation: n u32. */
void gs_encode_compact(size_t n, const float* scales, const float* opacity,
                       const float* quats, uint32_t* out_shape, uint32_t* out_rotation);

/* Quaternion -> oct101012 (splat_encode.rs:234-252). */
uint32_t gs_encode_quat_oct101012(const float quat_xyzw[4]);

/* One signed RGB SH coefficient -> rgb8e8 (engine.rs:1940-1949). */
uint32_t gs_pack_sh_rgb8e8(const float rgb[3]);
/* n coefficients (n*3 floats) -> n packed words. */
void gs_encode_sh_rgb8e8(size_t n, const float* rgb, uint32_t* out);

/* IEEE binary16 round-to-nearest-even (half::f16::from_f32). */
uint16_t gs_f32_to_f16(float v);

/* Canonical 3DGS PLY reader (src/streaming-backend-impl/
 * CanonicalGaussianPlyLoader.ts:84-350): ascii, binary_little_endian or
 * binary_big_endian, vertex element first, scalar properties x y z,
 * scale_0..2 (log), rot_0..3 (w x y z), opacity (logit), f_dc_0..2 and a
 * contiguous f_rest_0..(3K-1) block (channel-major), K in {0, 3, 8, 15}.
 * Outputs, as the loader computes them: means (count * 3), scales
 * max(exp(s), 1e-6) (count * 3), opacity sigmoid(o) (count), quaternions
 * normalised, (x, y, z, w), identity when degenerate (count * 4), SH
 * coefficients (count * (deg+1)^2 * 3: DC first, rgb per coefficient).
 * Call with every output NULL to read only *info. Returns 0, or -1 with the
 * reference's message in gs_ply_last_error() (per thread). */
typedef struct gs_ply_info {
  uint32_t count;
  uint32_t sh_degree;
  uint32_t format; /* 0 ascii, 1 binary little endian, 2 binary big endian */
} gs_ply_info;
int gs_ply_parse(const void* data, size_t size, gs_ply_info* info, float* means, float* scales,
                 float* opacity, float* quats, float* sh);
const char* gs_ply_last_error(void);

/* The store's slot buffers for one cloud (engine.rs:1801-1949, hole rows
 * :1917-1923): `capacity` slots, the first n holding the Gaussians in order
 * (means.w = object_id; compact shape and oct101012 rotation as
 * gs_encode_compact; SH as gs_encode_sh_rgb8e8), the rest holes with
 * means.w = -1 and every other word 0. out_means: capacity * 4 f32;
 * out_shape: capacity * 2; out_rotation: capacity; out_sh: capacity *
 * (deg+1)^2. Returns 0, or -1 when n > capacity or sh_degree > 3. */
int gs_pack_scene(size_t n, size_t capacity, uint32_t object_id, uint32_t sh_degree,
                  const float* means3, const float* scales, const float* opacity,
                  const float* quats, const float* sh, float* out_means, uint32_t* out_shape,
                  uint32_t* out_rotation, uint32_t* out_sh);

typedef struct gs_scene_spec {
  uint64_t seed;
  uint32_t count;
  uint32_t clusters;         /* cluster centres, uniform in [-extent, extent]^3 */
  float extent;
  float cluster_sigma;       /* isotropic spread of each cluster */
  float log_scale_mean, log_scale_sigma, log_scale_min, log_scale_max;
  uint32_t anisotropic;      /* one random axis major, the other two minor */
  float major_mean, major_sigma, minor_mean, minor_sigma;
  float opacity_sigma;       /* opacity = sigmoid(N(0, sigma)) */
  uint32_t sh_degree;
  float dc_mean, dc_sigma, band_sigma;
  /* Near-plane clusters: a fraction of points around centres placed at view
   * depth [near_depth_min, near_depth_max] inside the camera frustum. */
  float near_fraction;
  uint32_t near_clusters;
  float near_depth_min, near_depth_max;
  float near_sigma;
  double camera_position[3];
  double camera_target[3];
  double camera_fov_deg;     /* vertical */
  double camera_aspect;
} gs_scene_spec;

void gs_scene_spec_default(gs_scene_spec* spec);

/* Generate count Gaussians (object id 0). means: count*4 f32; shape:
 * count*2; rotation: count; sh: count*(deg+1)^2. The optional float
 * outputs (may be NULL) receive the pre-encoding values: scales count*3,
 * opacity count, quats count*4. Returns 0, or -1 on invalid spec. */
int gs_generate_scene(const gs_scene_spec* spec, float* means, uint32_t* shape,
                      uint32_t* rotation, uint32_t* sh, float* scales_out,
                      float* opacity_out, float* quats_out);

#ifdef __cplusplus
}
#endif

#endif
