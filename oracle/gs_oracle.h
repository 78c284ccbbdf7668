/* gs_oracle.h - CPU oracle for the tiled Gaussian-splatting frame path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference pipeline
 * (vinneyto/3dgs_tile_webgpu, the kernels under src/renderer/kernels and
 * the stages under src/renderer/pipeline) on the CPU, so the GPU product
 * can be checked against it. Only tests, the smoke check and the benchmark's
 * CPU-baseline leg may load it; the product never does.
 *
 * Arithmetic: IEEE binary32 in the reference's operation order, no FMA
 * contraction, with the elementary functions of gs_math.h standing in for the
 * WGSL builtins (whose precision WGSL leaves implementation-defined).
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Inputs of one frame. Buffers follow the reference GaussianData contract
 * (GaussianData.ts:10-33): compact geometry and rgb8e8 SH. */
typedef struct gs_oracle_scene {
  uint32_t count;        /* slots in the packed buffers */
  uint32_t active_count; /* active ordinals to project */
  uint32_t sh_degree;    /* 0..3 */
  uint32_t object_count;
  const float* means;             /* count * 4: xyz, objectId (<0 = hole) */
  const uint32_t* scales_opacity; /* count * 2: f16 ln sx|ln sy, ln sz|opacity */
  const uint32_t* rotations;      /* count: oct101012 */
  const uint32_t* sh;             /* count * (deg+1)^2: rgb8e8 */
  const uint32_t* active_slots;   /* active_count source slots, or NULL */
  const float* objects;           /* object_count * 40 (10 vec4 per object) */
  float projection[16];           /* column-major */
  float width, height, near_plane, far_plane;
  float filter_variance;
  uint32_t boundary_axis;         /* 0 clip, 1 repeat-x, 2 repeat-y */
  uint32_t subpixel_culling;      /* bool */
  uint32_t antialias_compensated; /* bool */
  uint32_t depth_mode;            /* 0 float32 keys, 1 packed16 keys */
  uint32_t raster_chunk_size;     /* 0 = no chunking */
  uint32_t capacity;              /* intersection capacity */
  float transmittance_threshold;
  float background[4];
  uint32_t threads;               /* worker threads (0 = all) */
  uint32_t tile_size;             /* tile edge 8, 16 or 32; 0 = 16 (constants.ts:1) */
  uint32_t cull_outputs;          /* also produce cull_features (n * 16 f32) and
                                     max_contribution (n u32: binary32 bits of each
                                     Gaussian's largest alpha * T) */
  uint32_t output_depth;          /* also produce depth_target (W*H f32): the
                                     finalized scalar of finalizeGaussianDepth.ts
                                     with the default depth node */
  float depth_alpha_threshold;    /* default 0.95 (GaussianPass.ts:243) */
  float invalid_depth_value;      /* default 1 */
} gs_oracle_scene;

typedef struct gs_oracle_counts {
  uint32_t visible;      /* V */
  uint32_t total;        /* requested intersections (saturating) */
  uint32_t count;        /* K actually emitted (0 on overflow) */
  uint32_t overflow;     /* 1 if total > capacity */
  uint32_t tiles_x, tiles_y;
  uint32_t sentinels;    /* padded records (must be 0) */
  uint32_t chunk_tasks;
  uint64_t checked;      /* raster (pixel, Gaussian) pairs evaluated: the
                            reference's `checked` counter, every chunk task
                            of a chunked tile counted in full */
  uint64_t blended;      /* pairs that blended */
  double stage_seconds[8]; /* project, visible, depth sort, scan, emit,
                              tile sort, offsets, raster */
  uint64_t checked_heavy; /* the part of `checked` in tiles above 1024
                             records or above the chunk size (the split the
                             GPU's heavy-unit counter uses) */
} gs_oracle_counts;

typedef struct gs_oracle_result gs_oracle_result;

/* Run the whole frame. Returns NULL on invalid input. */
gs_oracle_result* gs_oracle_run(const gs_oracle_scene* scene);
void gs_oracle_counts_of(const gs_oracle_result* r, gs_oracle_counts* out);
/* Copy a named intermediate into dst (bytes must equal the array size).
 * Names: projected_mean, projected_conic, projected_color, tile_counts,
 * visible_offsets, depth_records, ordered_counts, intersection_offsets,
 * emitted_records, sorted_records, tile_offsets, image_f32, image_f16,
 * depth_f32 (per-pixel weighted view depth), final_t, depth_target
 * (finalized depth, with output_depth).
 * Returns the byte size, or 0 if unknown; with dst == NULL only the size. */
size_t gs_oracle_get(const gs_oracle_result* r, const char* name, void* dst,
                     size_t bytes);
void gs_oracle_free(gs_oracle_result* r);

/* FP64 blend of the same sorted per-tile lists and projected values: the
 * exact-arithmetic image the f32 paths are measured against. out: W*H*4. */
void gs_oracle_blend_f64(const gs_oracle_result* r, double* out);

/* Single projection helpers, exposed for known-answer tests. */
/* perspective_gaussian_projection + gaussian_projection_result: out[16] is
 * the column-major mat4 the reference returns. */
void gs_oracle_perspective_projection(const float view[3],
                                      const float cov_view[9],
                                      const float projection[16],
                                      const float viewport[4],
                                      float filter_variance, float out[16]);
void gs_oracle_projection_result(const float center[2], float depth,
                                 const float covariance[3],
                                 float filter_variance, float out[16]);
int gs_oracle_projection_is_valid(const float p[16]);
void gs_oracle_view_covariance(const float scale[3], const float rotation[4],
                               const float model_view[16], float out[9]);
int gs_oracle_tile_hit(const float center[2], const float conic[3],
                       float power_threshold, const float radius[2],
                       int tile_x, int tile_y, const float viewport[2],
                       uint32_t boundary_axis);
uint32_t gs_oracle_count_tiles(const float center[2], const float conic[3],
                               float power_threshold, const float radius[2],
                               int tile_min_x, int tile_min_y, int tile_max_x,
                               int tile_max_y, const float viewport[2],
                               uint32_t boundary_axis);
int gs_oracle_subpixel_has_sample(const float center[2], const float conic[3],
                                  float power_threshold, const float extent[2],
                                  uint32_t viewport_x, uint32_t viewport_y,
                                  uint32_t boundary_axis);
void gs_oracle_boundary_delta(const float pixel[2], const float center[2],
                              const float conic[3], const float viewport[2],
                              uint32_t boundary_axis, float out[2]);
void gs_oracle_decode_scale_opacity(uint32_t x, uint32_t y, float out[4]);
void gs_oracle_decode_rotation(uint32_t packed, float out[4]);
void gs_oracle_decode_sh_rgb8e8(uint32_t packed, float out[3]);
/* evaluate_gaussian_sh_rgb8e8 for one Gaussian's (deg+1)^2 words and a unit
 * direction: clamp(sum + 0.5, 0, 1) per channel. */
void gs_oracle_evaluate_sh(const uint32_t* words, uint32_t degree, const float dir[3],
                           float out[3]);
/* finalizeGaussianDepth for one pixel. encoding 0 is the default depth node
 * (clamp((near + viewZ) far / ((far - near) viewZ), 0, 1), viewZ = -depth),
 * 1 the raw scalar (rasterOutputDepth), 2 a non-finite encoding
 * (rasterOutputDepth / 0), 3 the accumulated alpha. */
float gs_oracle_finalize_depth(float weighted_depth, float transmittance,
                               float depth_alpha_threshold, float invalid_value,
                               float near_plane, float far_plane, uint32_t encoding);
/* Elementwise shared elementary functions (accuracy tests). */
void gs_oracle_math(int fn, const float* a, const float* b, float* out, size_t n);

#ifdef __cplusplus
}
#endif

#endif
