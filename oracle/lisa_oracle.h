/*
 * CPU oracle for the LiSA path-tracing hot path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker; the product (liblisa.so)
 * never links or calls it.
 *
 * A scalar C++ restatement of
 *   src/LiSA/src/shader.cu            (raygen / closest-hit / miss path loop, NEE)
 *   src/LiSA/src/maths.cu             (rng remap, hemisphere, fresnel, refract, bary normal)
 *   src/LiSA/src/bsdfs/lambertian.cu  (bounce / BRDF / BTDF)
 *   src/cuda/random.h                 (tea<16>, lcg, rnd)
 *   src/sutil/vec_math.h              (float3 algebra, exact op order)
 *   src/cuda/helpers.h                (sRGB + 8-bit quantisation)
 *   src/sutil/Camera.cpp:34-45        (UVW frame)
 *   src/sutil/sutil.cpp:97-117,523-600 (PPM writer, row flip)
 * plus the accelerator the reference gets from OptiX (closed): an LBVH
 * (30-bit Morton codes, stable sort, Karras 2012 hierarchy, refit) and a
 * stack traversal with the Woop/Benthin/Wald (JCGT 2013) watertight test.
 *
 * Decisions where the reference is unspecified (shared with the HIP kernels):
 *  - occlusion rays report the CLOSEST hit (ties -> lowest triangle id), not
 *    OptiX's first-found hit (shader.cu:69);
 *  - a ray whose direction is zero (total internal reflection, maths.cu:29)
 *    or not finite counts as a ray but misses without traversal;
 *  - make_float3(rng,rng,rng) / make_float2(rng,rng) draw x, then y, then z;
 *  - pow() in fresnel is r*r and ((x*x)*(x*x))*x;
 *  - IEEE FP32 everywhere (no FMA contraction, correctly rounded / and sqrt),
 *    the reference's --use_fast_math rounding is not reproduced.
 * Parity status of every function is listed in the header of each test file.
 */
#ifndef LISA_ORACLE_H
#define LISA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same byte layout as lisa_material in include/lisa.h (structs.hh:16-23). */
typedef struct orc_material {
  float roughness;
  float alpha;
  float n;
  float diffuse[3];
  int32_t emit;
  float emission[3];
} orc_material;

typedef struct orc_render_params {
  uint32_t width, height, spp, bounces, subframe;
  float eye[3], U[3], V[3], W[3];
  int32_t record_bounces; /* max bounces stored per sample in rad_hits (0 = none) */
} orc_render_params;

/* counters[0..7]: radiance rays, occlusion rays, node visits, triangle tests,
 * radiance hits (shading gathers), samples, zero/invalid-direction rays, pixels */
enum { ORC_NCOUNTERS = 8 };

typedef struct orc_scene orc_scene;

/* --- random.h / maths.cu / lambertian.cu restatements (exposed for KATs) --- */
uint32_t orc_tea16(uint32_t v0, uint32_t v1);
uint32_t orc_lcg(uint32_t* state);
float orc_rnd(uint32_t* state);
float orc_rng(uint32_t* state);
void orc_hemisphere(const float* N, uint32_t* state, float* out3);
float orc_fresnel(float cosT, float eta);
void orc_refract(float cosI, const float* d, const float* N, float eta, float* out3);
void orc_reflect(const float* i, const float* n, float* out3);
float orc_brdf(const float* N, const float* L);
void orc_bounce(const float* d, const float* N, uint32_t* state, float roughness, float* out3);
/* n bounces from one RNG stream; the sampler's density at n unit directions. */
void orc_bounce_many(const float* d, const float* N, uint32_t* state, float roughness, int32_t n,
                     float* out3);
void orc_bsdf_pdf(const float* d, const float* N, const float* L, int32_t n, float roughness, float* out);
void orc_bary_normal(const float* P, const float* n1, const float* n2, const float* n3,
                     const float* v1, const float* v2, const float* v3, float* out3);
void orc_camera_uvw(const float* eye, const float* lookat, float fov, float aspect,
                    float* U, float* V, float* W);
void orc_srgb_quantize(const float* rgba, int n_pixels, uint8_t* rgba8);
int orc_save_ppm(const char* path, const uint8_t* rgba8, int width, int height);
uint32_t orc_morton3(float x, float y, float z); /* x,y,z in [0,1] */
/* The NEE loop of shader.cu:196-209 driven by scripted occlusion outcomes
 * (-1 miss, 0 non-emitter, m > 0 emitter material m; past n: non-emitter).
 * Returns +tries when a light was found, -tries otherwise; the final hit
 * flag and light material are returned through hit_out / light_out. */
int32_t orc_nee_scripted(const int32_t* outcomes, int32_t n, int32_t hit_in, int32_t light_in,
                         int32_t* hit_out, int32_t* light_out);

/* --- scene / LBVH --- */
orc_scene* orc_scene_create(const float* verts /*9T*/, const float* normals /*9T*/,
                            const int32_t* mat_idx /*T*/, int32_t T,
                            const orc_material* mats, int32_t M);
void orc_scene_destroy(orc_scene* s);
/* New positions (and optionally normals) for the same T triangles: the LBVH
 * keeps its Morton order and topology and is refit (topology reuse). */
void orc_scene_update_vertices(orc_scene* s, const float* verts, const float* normals);
int32_t orc_num_tris(const orc_scene* s);
int32_t orc_num_emitters(const orc_scene* s);
// Switch the emitter-query culls off (enabled = 0) or back on: the culls
// change no result, only the node-visit counters (a test of that).
void orc_scene_set_emitter_cull(orc_scene* s, int32_t enabled);
// The NEE loop's cull of a try on its raw draws (line_culled):
// culled[i] = 1 when the line through O[i] along d[i] certainly misses every emitter.
void orc_raw_cull(const orc_scene* s, const float* O, const float* d, int32_t n, int32_t* culled);
/* which = 0: the BVH over all triangles, 1: over the emitting triangles (N of them).
 * morton_sorted[N], prim_sorted[N] (global triangle ids), child refs [2*max(N-1,1)],
 * child boxes [12*max(N-1,1)] as (c0.lo.xyz, c0.hi.xyz, c1.lo.xyz, c1.hi.xyz) */
int orc_get_bvh(const orc_scene* s, int32_t which, uint32_t* morton_sorted,
                uint32_t* prim_sorted, uint32_t* child_refs, float* child_boxes);
/* NEE occlusion record codes of n rays with the given hit flags, from the
 * emitter-BVH + any-hit query used by the path loop (code_fast) and from one
 * full closest-hit query (code_ref).  Counters are those of the fast query. */
void orc_occlusion(const orc_scene* s, const float* O, const float* D, int32_t n,
                   const int32_t* hit_before, int32_t* code_fast, int32_t* code_ref,
                   uint64_t* counters);

/* Closest hit of n rays; prim_out = -1 on miss; counters optional (ORC_NCOUNTERS). */
void orc_closest_hit(const orc_scene* s, const float* O, const float* D, int32_t n,
                     float tmin, float tmax, int32_t* prim_out, float* t_out,
                     uint64_t* counters, int threads);
void orc_closest_hit_brute(const orc_scene* s, const float* O, const float* D, int32_t n,
                           float tmin, float tmax, int32_t* prim_out, float* t_out);

/* --- radiance-cache path modes (lisa_render_cached) ---
 * Scene box of the position features: min corner and 1 / extent per axis. */
void orc_cache_bounds(const orc_scene* s, float* bmin, float* binv);
/* Paths of p->spp samples on the listed pixels, prefix_bounces + 1 radiance
 * rays at most: the surface hit of radiance ray prefix_bounces + 1 ends the
 * path in a query (12 features, throughput).  accum_out (4n): per-pixel mean
 * of the radiance gathered before the queries.  Queries in (pixel, sample)
 * order; q_item = index into pixel_ids; *n_queries may exceed q_cap.
 * emit_expect != 0: at each roughness-1 Lambertian vertex with a bounce left,
 * one light sample (by area over the emitters) adds the bounce's emission in
 * expectation, the throughput is scaled by (1 - the sample's hit estimate)
 * and the bounce is drawn on the condition that it hits no emitter (at most
 * 64 redraws, then the path ends).  Same expectation as without it.
 * nee_loops > 1 (needs emit_expect, at most 16): the NEE at the camera ray's
 * hit is the mean of that many runs from its entry hit flag and light, each
 * added as it ends; the last run's flag and light go on (nee_loops_deep: the
 * same at later vertices). */
int orc_cache_queries(const orc_scene* s, const orc_render_params* p, uint32_t prefix_bounces,
                      const int32_t* pixel_ids, int32_t n, float* accum_out, float* q_feat,
                      float* q_att, int32_t* q_item, int32_t q_cap, int32_t* n_queries, int threads,
                      int emit_expect, int nee_loops, int nee_loops_deep);
/* The variance-reduced plain tracer (the estimators of orc_cache_queries'
 * emit_expect / nee_loops on whole paths of p->bounces radiance rays, no
 * query): per-pixel means into accum_out (4n).  Without emit_expect these are
 * the plain tracer's samples. */
int orc_render_full_paths(const orc_scene* s, const orc_render_params* p, const int32_t* pixel_ids,
                          int32_t n, float* accum_out, int threads, int emit_expect, int nee_loops,
                          int nee_loops_deep);
/* Training rows of plain paths (p->bounces) on the listed pixels: one row per
 * surface hit (depth 0) or per surface hit of radiance ray number `depth`,
 * 12 features and the target (final - gathered before the hit) / throughput
 * on arrival; in (pixel, sample, vertex) order, row_path = index into
 * pixel_ids (may be NULL); *n_rows may exceed cap. */
int orc_cache_train_rows(const orc_scene* s, const orc_render_params* p, const int32_t* pixel_ids,
                         int32_t n, uint32_t depth, float* feat, float* tgt, int32_t* row_path,
                         int32_t cap, int32_t* n_rows, int threads);

/* Render pixels (all if pixel_ids == NULL, else n_pixels listed ids).
 * accum_out: float4 per listed pixel (rgb, 1).  accum_prev (same layout), when
 * given and subframe > 0, is blended as a running mean; without it the
 * subframe only seeds the RNG.  Per-sample records are optional (NULL to skip):
 *   sample_rgb[n*spp*3], sample_bounces[n*spp], sample_occl[n*spp],
 *   sample_hash[n*spp], rad_hits[n*spp*record_bounces]. */
int orc_render(const orc_scene* s, const orc_render_params* p,
               const int32_t* pixel_ids, int32_t n_pixels,
               float* accum_out, const float* accum_prev,
               float* sample_rgb, int32_t* sample_bounces, int32_t* sample_occl,
               uint32_t* sample_hash, int32_t* rad_hits,
               uint64_t* counters, int threads);

#ifdef __cplusplus
}
#endif
#endif
