// driver.cpp - exercises the CPU oracle and the host producer under
// AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitizers.py with g++, host code only). Exits non-zero, or the
// sanitizer aborts, on any failure.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "gs_oracle.h"
#include "gs_producer.h"

static int failures = 0;
#define CHECK(c)                                                 \
  do {                                                           \
    if (!(c)) {                                                  \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      failures++;                                                \
    }                                                            \
  } while (0)

static std::vector<unsigned char> read_file(const char* path) {
  std::vector<unsigned char> b;
  FILE* f = fopen(path, "rb");
  if (!f) return b;
  unsigned char tmp[4096];
  size_t n;
  while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) b.insert(b.end(), tmp, tmp + n);
  fclose(f);
  return b;
}

// Camera at (0, 0, dist) looking down -z, 60 degree fov: projection and one
// object frame (identity world, view = translation by -dist).
static void camera(float dist, float aspect, float proj[16], float objects[40]) {
  memset(proj, 0, 16 * sizeof(float));
  float f = 1.7320508f, n = 0.1f, fa = 1000.0f;
  proj[0] = f / aspect;
  proj[5] = f;
  proj[10] = -(fa + n) / (fa - n);
  proj[11] = -1.0f;
  proj[14] = -2.0f * fa * n / (fa - n);
  double cam[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, dist, 1};
  double world[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  // gs_object_frame lives in the device library; write the same table here.
  for (int i = 0; i < 16; i++) objects[i] = (float)world[i];
  for (int i = 0; i < 16; i++) objects[16 + i] = (float)(i == 14 ? -dist : world[i]);
  objects[32] = 0.0f;
  objects[33] = 0.0f;
  objects[34] = dist;
  objects[35] = 1.0f;
  objects[36] = 1.0f;
  objects[37] = objects[38] = objects[39] = 0.0f;
  (void)cam;
}

int main(int argc, char** argv) {
  // Producer: a seeded scene.
  gs_scene_spec spec;
  gs_scene_spec_default(&spec);
  spec.count = 20000;
  spec.sh_degree = 3;
  spec.extent = 5.0f;
  spec.camera_position[2] = 15.0;
  std::vector<float> means(spec.count * 4), scales(spec.count * 3), opac(spec.count),
      quats(spec.count * 4);
  std::vector<uint32_t> shape(spec.count * 2), rot(spec.count), sh(spec.count * 16);
  CHECK(gs_generate_scene(&spec, means.data(), shape.data(), rot.data(), sh.data(), scales.data(),
                          opac.data(), quats.data()) == 0);
  // Oracle frames over option variants.
  float proj[16], objects[40];
  camera(15.0f, 4.0f / 3.0f, proj, objects);
  struct Variant {
    uint32_t chunk, axis, depth_mode, output_depth, cull, tile;
  } variants[] = {{8192, 0, 0, 0, 0, 16}, {256, 0, 0, 1, 0, 16}, {0, 1, 0, 0, 0, 16},
                  {256, 2, 1, 1, 1, 16},  {512, 0, 0, 0, 0, 8},  {2048, 0, 0, 0, 0, 32}};
  for (const Variant& v : variants) {
    gs_oracle_scene s;
    memset(&s, 0, sizeof s);
    s.count = s.active_count = spec.count;
    s.sh_degree = 3;
    s.object_count = 1;
    s.means = means.data();
    s.scales_opacity = shape.data();
    s.rotations = rot.data();
    s.sh = sh.data();
    s.objects = objects;
    memcpy(s.projection, proj, sizeof proj);
    s.width = 320.0f;
    s.height = 240.0f;
    s.near_plane = 0.1f;
    s.far_plane = 1000.0f;
    s.filter_variance = 0.3f;
    s.boundary_axis = v.axis;
    s.subpixel_culling = 1;
    s.antialias_compensated = 1;
    s.depth_mode = v.depth_mode;
    s.raster_chunk_size = v.chunk;
    s.capacity = 1u << 22;
    s.transmittance_threshold = 1e-4f;
    s.threads = 4;
    s.tile_size = v.tile;
    s.cull_outputs = v.cull;
    s.output_depth = v.output_depth;
    s.depth_alpha_threshold = 0.95f;
    s.invalid_depth_value = 1.0f;
    gs_oracle_result* r = gs_oracle_run(&s);
    CHECK(r != nullptr);
    if (!r) continue;
    gs_oracle_counts c;
    gs_oracle_counts_of(r, &c);
    CHECK(c.visible > 0 && c.count > 0 && c.sentinels == 0);
    const char* names[] = {"projected_mean", "sorted_records", "tile_offsets", "image_f16",
                           "depth_target", "max_contribution"};
    for (const char* n : names) {
      size_t bytes = gs_oracle_get(r, n, nullptr, 0);
      std::vector<unsigned char> buf(bytes ? bytes : 1);
      CHECK(gs_oracle_get(r, n, buf.data(), bytes) == bytes);
    }
    std::vector<double> f64((size_t)320 * 240 * 4);
    gs_oracle_blend_f64(r, f64.data());
    gs_oracle_free(r);
  }
  // Producer: PLY parsing of the fixture, every truncation of it and a few
  // corruptions; packing with hole rows.
  std::vector<unsigned char> ply = argc > 1 ? read_file(argv[1]) : std::vector<unsigned char>();
  CHECK(!ply.empty());
  gs_ply_info info;
  CHECK(gs_ply_parse(ply.data(), ply.size(), &info, nullptr, nullptr, nullptr, nullptr, nullptr) ==
        0);
  size_t n = info.count, coefs = (size_t)(info.sh_degree + 1) * (info.sh_degree + 1);
  std::vector<float> pm(n * 3), ps(n * 3), po(n), pq(n * 4), psh(n * coefs * 3);
  CHECK(gs_ply_parse(ply.data(), ply.size(), &info, pm.data(), ps.data(), po.data(), pq.data(),
                     psh.data()) == 0);
  for (size_t cut = 0; cut < ply.size(); cut += 7) {
    std::vector<unsigned char> part(ply.begin(), ply.begin() + cut);
    gs_ply_info i2;
    if (gs_ply_parse(part.data(), part.size(), &i2, nullptr, nullptr, nullptr, nullptr, nullptr) ==
            0 &&
        i2.count == info.count && i2.sh_degree == info.sh_degree)
      gs_ply_parse(part.data(), part.size(), &i2, pm.data(), ps.data(), po.data(), pq.data(),
                   psh.data());
  }
  for (size_t at = 0; at < 1200 && at < ply.size(); at += 3) {
    std::vector<unsigned char> bad(ply);
    bad[at] = (unsigned char)(bad[at] ^ 0x5a);
    gs_ply_info i3;
    if (gs_ply_parse(bad.data(), bad.size(), &i3, nullptr, nullptr, nullptr, nullptr, nullptr) ==
            0 &&
        i3.count == info.count && i3.sh_degree == info.sh_degree)
      gs_ply_parse(bad.data(), bad.size(), &i3, pm.data(), ps.data(), po.data(), pq.data(),
                   psh.data());
  }
  size_t cap = n + 17;
  std::vector<float> om(cap * 4);
  std::vector<uint32_t> os(cap * 2), orot(cap), osh(cap * coefs);
  CHECK(gs_pack_scene(n, cap, 1, info.sh_degree, pm.data(), ps.data(), po.data(), pq.data(),
                      psh.data(), om.data(), os.data(), orot.data(), osh.data()) == 0);
  CHECK(om[(cap - 1) * 4 + 3] == -1.0f);
  printf("sanitized driver: %d failures\n", failures);
  return failures ? 1 : 0;
}
