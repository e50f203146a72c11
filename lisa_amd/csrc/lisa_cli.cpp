// `lisa -s scene.rto [-d] [--cache] [--dump-float file.pfm]`: the MI355X
// build of LiSA's command line (src/LiSA/src/main.cc:6-28).
//
//   -s path   scene file (required), parsed like SceneParser
//   -d        progressive mode: subframes of min(spp, 16) samples, each
//             lerped into the accum buffer (render.cc:75-131, shader.cu:160-164,
//             optix_wrapper.cc:430) -- without the GLFW window
//   --cache   progressive frames of the learned radiance cache instead
//             (lisa_render_cached, trained online: paths end at the cache
//             after 2 radiance rays; 4 samples per frame until the scene's
//             sample count; a biased estimate, see lisa.h)
//   --dump-float path   also write the linear float accum buffer as PFM
// Render mode puts every sample in one launch (render.cc:133-148), saves the
// PPM named by output_image and prints the wall time in minutes.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lisa.h"

namespace {

const char* get_option(int argc, char** argv, const char* opt) {  // parse_args.hh:3-11
  for (int i = 1; i < argc; ++i)
    if (std::strcmp(argv[i], opt) == 0) return i + 1 < argc ? argv[i + 1] : nullptr;
  return nullptr;
}
bool has_option(int argc, char** argv, const char* opt) {  // parse_args.hh:13-16
  for (int i = 1; i < argc; ++i)
    if (std::strcmp(argv[i], opt) == 0) return true;
  return false;
}

int die(const char* what, lisa_ctx* ctx) {
  std::fprintf(stderr, "%s: %s\n", what, lisa_last_error(ctx));
  return 1;
}

bool write_pfm(const char* path, const std::vector<float>& rgba, uint32_t w, uint32_t h) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  std::fprintf(f, "PF\n%u %u\n-1.0\n", w, h);  // little-endian, bottom row first
  std::vector<float> row(3 * (size_t)w);
  for (uint32_t y = 0; y < h; ++y) {
    for (uint32_t x = 0; x < w; ++x)
      for (int c = 0; c < 3; ++c) row[3 * x + c] = rgba[4 * ((size_t)y * w + x) + c];
    std::fwrite(row.data(), sizeof(float), row.size(), f);
  }
  std::fclose(f);
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  const char* scene_path = has_option(argc, argv, "-s") ? get_option(argc, argv, "-s") : nullptr;
  if (!scene_path) {
    std::fprintf(stderr, "Missing scene path.\nUsage: %s -s scene_path\n", argv[0]);
    return 1;
  }
  lisa_scene_file* sf = nullptr;
  if (lisa_scene_file_load(scene_path, &sf) != LISA_OK) {
    std::fprintf(stderr, "%s\n", lisa_last_error(nullptr));
    return 1;
  }
  lisa_scene_info info;
  lisa_scene_file_info(sf, &info);

  lisa_ctx* ctx = nullptr;
  if (lisa_create(&ctx, 0) != LISA_OK) return die("lisa_create", nullptr);
  if (lisa_set_scene(ctx, lisa_scene_file_vertices(sf), lisa_scene_file_normals(sf),
                     lisa_scene_file_mat_indices(sf), info.num_tris, lisa_scene_file_materials(sf),
                     info.num_materials) != LISA_OK)
    return die("lisa_set_scene", ctx);

  lisa_render_params p;
  std::memset(&p, 0, sizeof(p));
  p.width = info.width;
  p.height = info.height;
  p.bounces = info.num_bounces;
  p.mode = LISA_MODE_PLAIN;
  for (int c = 0; c < 3; ++c) p.eye[c] = info.eye[c];
  if (lisa_camera_uvw(info.eye, info.look_at, info.fov, (float)info.width / (float)info.height, p.U,
                      p.V, p.W) != LISA_OK)
    return die("lisa_camera_uvw", nullptr);

  std::printf("Starting rendering...\n");
  const size_t npx = (size_t)info.width * info.height;
  std::vector<float> accum(4 * npx);
  std::vector<uint8_t> rgba8(4 * npx);
  const auto t0 = std::chrono::system_clock::now();
  if (has_option(argc, argv, "--cache")) {
    lisa_cache* cache = nullptr;
    if (lisa_cache_create(&cache, 0, 1, 1e-3f) != LISA_OK) return die("lisa_cache_create", nullptr);
    lisa_cache_frame_params cp;
    std::memset(&cp, 0, sizeof(cp));
    cp.prefix_bounces = 2;
    cp.train_paths = 262144;
    cp.train_bounces = std::max<uint32_t>(1u, std::min<uint32_t>(info.num_bounces, 64u));
    cp.train_max = 1u << 16;
    cp.train_steps = 8;
    cp.flags = LISA_CACHE_EMISSION;
    cp.nee_loops = 8;
    cp.nee_loops_deep = 1;
    cp.seed = 1;
    p.spp = std::max<uint32_t>(1u, std::min<uint32_t>(info.num_samples, 4u));
    for (p.subframe = 0;; ++p.subframe) {
      lisa_cache_frame_stats st;
      if (lisa_render_cached(ctx, cache, &p, &cp, &st) != LISA_OK) return die("lisa_render_cached", ctx);
      std::printf("nb sample   : %8u\n", (p.subframe + 1) * p.spp);
      if ((p.subframe + 1) * p.spp >= info.num_samples) break;
    }
    p.subframe = 0;
    lisa_cache_destroy(cache);
  } else if (has_option(argc, argv, "-d")) {
    p.spp = std::min<uint32_t>(info.num_samples, 16u);
    if (p.spp == 0) p.spp = 1;
    for (p.subframe = 0;; ++p.subframe) {
      if (lisa_render(ctx, &p, nullptr, nullptr) != LISA_OK) return die("lisa_render", ctx);
      std::printf("nb sample   : %8u\n", (p.subframe + 1) * p.spp);
      if ((p.subframe + 1) * p.spp >= info.num_samples) break;
    }
    p.subframe = 0;
  } else {
    p.spp = info.num_samples;
    if (lisa_render(ctx, &p, nullptr, nullptr) != LISA_OK) return die("lisa_render", ctx);
  }
  if (lisa_synchronize(ctx) != LISA_OK) return die("lisa_synchronize", ctx);
  if (lisa_copy_accum(ctx, accum.data(), (int64_t)npx) != LISA_OK) return die("lisa_copy_accum", ctx);
  lisa_srgb_quantize(accum.data(), (int64_t)npx, rgba8.data());
  if (lisa_save_ppm(lisa_scene_file_output_image(sf), rgba8.data(), (int32_t)info.width,
                    (int32_t)info.height) != LISA_OK)
    return die("save_image", nullptr);
  const char* pfm = get_option(argc, argv, "--dump-float");
  if (pfm && !write_pfm(pfm, accum, info.width, info.height)) {
    std::fprintf(stderr, "could not write %s\n", pfm);
    return 1;
  }
  const std::chrono::duration<float> dt = std::chrono::system_clock::now() - t0;
  std::printf("Rendering finished in %.2f mn.\n", dt.count() / 60.0f);
  lisa_destroy(ctx);
  lisa_scene_file_free(sf);
  return 0;
}
