/*
 * lisa.h -- C ABI of the MI355X-native LiSA path tracer (liblisa.so).
 *
 * This is the drop-in boundary for LiSA's OptiX render call.  Plain C types
 * only: pointers + sizes, integer status codes, no exception crosses it.
 * Every call returns LISA_OK (0) or a negative lisa_status; the message of
 * the last failure on a context is lisa_last_error(ctx) (lisa_last_error(NULL)
 * gives the last failure of a call made without a context).
 *
 * Threading: one HIP stream per context; calls are synchronous unless noted;
 * a context is not thread-safe; distinct contexts are independent.
 * Ownership: inputs are copied at call time, the caller keeps them; output
 * buffers are caller-owned; the context owns all device memory.
 */
#ifndef LISA_H
#define LISA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t lisa_status;
#define LISA_OK 0
#define LISA_ERR_INVALID_ARG (-1)
#define LISA_ERR_HIP (-2)
#define LISA_ERR_ALLOC (-3)
#define LISA_ERR_NO_SCENE (-4)
#define LISA_ERR_PARSE (-5)
#define LISA_ERR_IO (-6)
#define LISA_ERR_UNSUPPORTED (-7)

/* Material, same fields and meaning as `Material` in
 * src/LiSA/include/structs.hh:16-23: alpha < 1 => dielectric with index n;
 * otherwise opaque Lambertian with roughness; emit != 0 => light with
 * emission (other fields unused on that branch).  40 bytes. */
typedef struct lisa_material {
  float roughness;
  float alpha;
  float n;
  float diffuse[3];
  int32_t emit;
  float emission[3];
} lisa_material;

/* Launch parameters; the fields of `OptixParams`
 * (src/LiSA/include/structs.hh:60-75) minus the buffers and the handle,
 * which the context owns.  U, V, W come from lisa_camera_uvw. */
typedef struct lisa_render_params {
  uint32_t width, height;
  uint32_t spp;      /* samples_per_launch (render.cc:140 puts all spp in one launch) */
  uint32_t bounces;  /* num_bounces */
  uint32_t subframe; /* subframe_index: > 0 lerps with the previous accum buffer */
  float eye[3], U[3], V[3], W[3];
  int32_t mode;      /* LISA_MODE_* */
  uint32_t flags;    /* LISA_FLAG_* */
} lisa_render_params;

#define LISA_MODE_PLAIN 0        /* one work-item per pixel megakernel (the drop-in) */
#define LISA_FLAG_FP64_ACCUM 1u  /* per-pixel FP64 accumulation (reference renders >= 16k spp): the
                                    image's FP64 sums and sample counts are kept across subframes, so
                                    subframe k's accum is the exact mean of every sample of subframes
                                    0..k (instead of shader.cu:160-164's FP32 running lerp);
                                    resumable with lisa_accum64_save / lisa_accum64_load */
#define LISA_FLAG_COUNT 2u       /* also count rays / node visits / triangle tests */
#define LISA_FLAG_PERSISTENT 4u  /* one persistent launch whose lanes pull pixels from a queue
                                    (same image; not with COUNT or FP64_ACCUM) */
#define LISA_FLAG_WAVEFRONT 8u   /* the ray-level wavefront form: path states in HBM, rounds of a
                                    persistent traversal kernel (per-wave LDS request queues, lanes
                                    refilled as their walks end) and a shading kernel (ballot /
                                    prefix-sum compaction of the next requests); the same image
                                    bit for bit (not with COUNT, FP64_ACCUM or PERSISTENT) */

/* Counters (lisa_get_counters), in order. */
#define LISA_NCOUNTERS 8
/* 0 radiance rays, 1 occlusion rays, 2 BVH node visits, 3 triangle tests,
 * 4 radiance hits shaded (vertex+normal gathers), 5 samples,
 * 6 zero/non-finite-direction rays (count as rays, miss without traversal),
 * 7 pixels */

typedef struct lisa_stats {
  float render_ms;    /* last render/trace kernel time (HIP events on the context stream) */
  float bvh_build_ms; /* last lisa_set_scene device build time */
  int32_t num_tris;
  int32_t num_nodes;  /* internal BVH nodes (64 B each) */
  int32_t num_emitters; /* triangles whose material emits (the NEE light BVH) */
  uint64_t device_bytes; /* device memory held by the context */
  int32_t wide_tree;  /* 1: the scene's tree also has its 4-wide form, which the uncounted path
                         kernels walk (lisa_set_wide_threshold) */
  int32_t reserved;
} lisa_stats;

typedef struct lisa_ctx lisa_ctx;

/* Replaces OptixWrapper::init_optix (src/LiSA/src/optix_wrapper.cc:39-52). */
lisa_status lisa_create(lisa_ctx** out, int32_t device);

/* Replaces OptixWrapper::~OptixWrapper / clean_state (optix_wrapper.cc:35-37,448-465). */
void lisa_destroy(lisa_ctx* ctx);

const char* lisa_last_error(const lisa_ctx* ctx);

/* Replaces OptixWrapper::create_mesh_handler + the GAS build and the SBT's
 * per-material records (optix_wrapper.cc:54-149, 327-417): uploads the
 * de-indexed soup (verts, normals: 9*T floats; mat_idx: T ints, as
 * RendererParams in structs.hh:90-103 with num_vertices = 3*T) and builds
 * the LBVH on the device. T == 0 is allowed (every ray misses). */
lisa_status lisa_set_scene(lisa_ctx* ctx, const float* verts, const float* normals,
                           const int32_t* mat_idx, int32_t T,
                           const lisa_material* mats, int32_t M);

/* New vertex positions (9*T floats) and optionally normals (NULL keeps them)
 * for the T triangles of the current scene: triangle records are rewritten
 * in the build's Morton order and both LBVHs are refit bottom-up, keeping
 * topology and Morton codes (animated / re-posed geometry; cheaper than a
 * rebuild, box quality degrades with large motion); a 4-wide form of the
 * tree is rebuilt from the refitted one.  No counterpart in the
 * reference, whose GAS is built once (optix_wrapper.cc:104-106). */
lisa_status lisa_update_vertices(lisa_ctx* ctx, const float* verts, const float* normals);

/* lisa_update_vertices for callers that cannot allocate device memory (code
 * under HIP graph capture, tight animation / serving loops): the same refit,
 * bit-identical trees, with the new positions (9*T floats) and optional
 * normals already in DEVICE memory (caller-owned), and no hipMalloc / hipFree
 * of its own -- the refit's union tables live in a device array of the library
 * (64 MB per device, held until the process ends).  At most 1048576
 * triangles (LISA_ERR_INVALID_ARG beyond, before anything runs).  A 4-wide
 * form of the tree (lisa_set_wide_threshold) is rewritten in place.  Every
 * call on one device shares that array: calls that overlap on one GPU (from
 * several threads or contexts) race.  No counterpart in the reference. */
lisa_status lisa_update_vertices_noalloc(lisa_ctx* ctx, const float* d_verts, const float* d_normals);

/* Replaces render() -> launchSubframe() -> optixLaunch(w, h, 1)
 * (src/LiSA/src/render.cc:35-58, 133-148).  The float4 accum buffer stays on
 * the device (as OptixParams::accum_buffer); accum_rgba_out (w*h*4 floats)
 * and rgba8_out (w*h*4 bytes, make_color of helpers.h:57-66 done on the host)
 * are optional host copies. */
lisa_status lisa_render(lisa_ctx* ctx, const lisa_render_params* p,
                        float* accum_rgba_out, uint8_t* rgba8_out);

/* Progressive subframes p->subframe .. p->subframe + n_subframes - 1 in one
 * hipGraph launch: the chain of n_subframes plain launches is built on the
 * first call for an image size and n (and for FP64 or not) and replayed by
 * later calls, with each launch's subframe index and buffers set on the
 * instantiated graph.  The accum buffer ends as after n_subframes lisa_render
 * calls with those subframes, bit for bit: the running-mean lerp of
 * src/LiSA/src/shader.cu:160-164, or with LISA_FLAG_FP64_ACCUM (the only flag
 * taken) the FP64 sums.  The reference's display loop (render.cc:75-131)
 * launches one subframe per frame; this is the same sequence for callers that
 * refine an image offline, with one host call and one wait per run. */
lisa_status lisa_render_progressive(lisa_ctx* ctx, const lisa_render_params* p, int32_t n_subframes);

/* Copies the device accum buffer (float4 per pixel, n_pixels = w*h of the
 * last render) to rgba_out.  No counterpart in the reference, whose
 * frame_buffer is mapped host memory (CUDAOutputBuffer ZERO_COPY). */
lisa_status lisa_copy_accum(lisa_ctx* ctx, float* rgba_out, int64_t n_pixels);

/* Checkpoint / resume of an FP64 progressive render (LISA_FLAG_FP64_ACCUM):
 * save copies the image's FP64 state, 4 doubles per pixel (r, g, b sums,
 * samples so far), after an FP64 render; load puts such a state into a
 * context (the accum buffer becomes its mean) so that the next FP64 render
 * with subframe k + 1 continues the sequence exactly where subframe k left
 * it -- bit for bit as if it had not been interrupted.  No counterpart in the
 * reference (its display loop restarts from subframe 0). */
lisa_status lisa_accum64_save(lisa_ctx* ctx, double* sums_out, int64_t n_pixels);
lisa_status lisa_accum64_load(lisa_ctx* ctx, const double* sums, int64_t n_pixels);

/* Waits for the context's stream (the reference's CUDA_SYNC_CHECK after a
 * launch, render.cc:57,143; every call here already waits before it returns). */
lisa_status lisa_synchronize(lisa_ctx* ctx);
/* Scene size, BVH build time and the last render's kernel time (HIP events).
 * No counterpart in the reference, which prints the render's wall time only
 * (render.cc:146-147). */
lisa_status lisa_get_stats(const lisa_ctx* ctx, lisa_stats* out);
/* Scenes set from now on with at least min_triangles triangles (default
 * 2^20) also get the 4-wide form of their tree (128 bytes per inner node: each
 * node's children, inner children replaced by their own two), and the path
 * kernels that count nothing walk it: fewer, wider visits, which pay where
 * the tree streams from HBM.  The hits are the same (closest hit, equal t to
 * the lower triangle id); node-visit counters (LISA_FLAG_COUNT) stay those of
 * the binary tree.  0 gives every scene the wide form.  No counterpart in
 * the reference (OptiX chooses its acceleration structure). */
lisa_status lisa_set_wide_threshold(lisa_ctx* ctx, int64_t min_triangles);
/* Counters of the last call made with LISA_FLAG_COUNT (or lisa_trace_rays). */
lisa_status lisa_get_counters(const lisa_ctx* ctx, uint64_t* counters /* LISA_NCOUNTERS */);
/* The hipStream_t the context launches on (for external event timing). */
void* lisa_get_stream(const lisa_ctx* ctx);

/* ---- batched wavefront form (no counterpart in the reference) ----
 * n scenes, each already set on its own context (all on one device), are
 * rendered by ONE persistent launch: the pixels of every scene, in 8x8-tile
 * order, form a device work queue; each lane takes `chunk` pixels at a time,
 * renders them completely and takes more until the queue is empty.  Every
 * pixel is computed exactly as lisa_render computes it (bit-identical
 * images); the work of short and long paths and of small and large scenes
 * is balanced over the whole GPU.  The queue reset and the launch are
 * captured once in a hipGraph and replayed by every lisa_batch_render.
 * Results land in each context's accum buffer (lisa_copy_accum).  The batch
 * must be re-created when a context's scene or image size changes. */
typedef struct lisa_batch lisa_batch;
typedef struct lisa_batch_config {
  int32_t grid_blocks; /* persistent workgroups (64 work-items each) */
  int32_t chunk;       /* form 0: pixels a lane takes from the queue per atomic */
  int32_t form;        /* 0: every lane pulls pixels (any scene); 1: every wave pulls 8x8 tiles;
                          2: ray-level wavefront (LISA_FLAG_WAVEFRONT's form over the batch's work
                          list: grid_blocks = the traversal kernel's persistent grid, chunk =
                          paths in flight per traversal lane) */
  int32_t graph;       /* 1: the launch is captured once into a hipGraph and replayed; 0: plain
                          launches on the stream.  A capture that fails (e.g. the stream is
                          already being captured by the caller) falls back to plain launches
                          and lisa_batch_get_config then reports graph = 0. */
} lisa_batch_config;

lisa_status lisa_batch_create(lisa_batch** out, lisa_ctx* const* ctxs,
                              const lisa_render_params* params, int32_t n);
void lisa_batch_destroy(lisa_batch* b);
lisa_status lisa_batch_set_config(lisa_batch* b, const lisa_batch_config* cfg);
lisa_status lisa_batch_get_config(const lisa_batch* b, lisa_batch_config* cfg);
/* Replays the graph; elapsed_ms (optional) = HIP-event time of the replay. */
lisa_status lisa_batch_render(lisa_batch* b, float* elapsed_ms);
/* Autotuner: times every (form, grid, chunk) candidate on a copy of the batch
 * with spp reduced to tune_spp and keeps the fastest configuration. */
lisa_status lisa_batch_autotune(lisa_batch* b, int32_t tune_spp, lisa_batch_config* best_out);

/* ---- learned radiance cache (no counterpart in the reference) ----
 * A small MLP maps 12 inputs in [0, 1] (position 3, incoming direction 2,
 * normal 2 as (theta/pi, phi/2pi), albedo 3, roughness 1, dielectric flag 1),
 * each encoded as (sin, cos)(pi 2^b x) for b = 0..5 (144 features), through
 * four ReLU layers of width 64 to 3 linear outputs (RGB radiance), no biases.
 * BF16 weights and activations, FP32 accumulation (MFMA); FP32 master weights
 * trained with Adam on the relative L2 loss (y - t)^2 / (y^2 + 0.01).
 * Weights are LISA_CACHE_PARAMS floats: per layer l, W_l^T[out_pad][in_pad]
 * row-major at offsets 0, 10240, 14336, 18432, 22528 (in_pad = 160, 64, 64,
 * 64, 64; out_pad = 64, 64, 64, 64, 16); padding entries are zero. */
typedef struct lisa_cache lisa_cache;
#define LISA_CACHE_INPUTS 12
#define LISA_CACHE_PARAMS 23552

lisa_status lisa_cache_create(lisa_cache** out, int32_t device, uint64_t seed, float lr);
void lisa_cache_destroy(lisa_cache* c);
const char* lisa_cache_last_error(const lisa_cache* c);
lisa_status lisa_cache_set_weights(lisa_cache* c, const float* w);
lisa_status lisa_cache_get_weights(lisa_cache* c, float* w);
/* out[n][3] = network(feat[n][12]) */
lisa_status lisa_cache_infer(lisa_cache* c, const float* feat, int64_t n, float* out);
/* Gradient of the mean relative-L2 loss over n samples (no update). */
lisa_status lisa_cache_gradients(lisa_cache* c, const float* feat, const float* target, int64_t n,
                                 float* grad /* LISA_CACHE_PARAMS, optional */,
                                 float* loss /* optional */);
/* One Adam step on n samples; loss (optional) is the loss before the step. */
lisa_status lisa_cache_train(lisa_cache* c, const float* feat, const float* target, int64_t n,
                             float* loss);
/* HIP-event time of the last infer / gradients / train call. */
float lisa_cache_last_ms(const lisa_cache* c);

/* One frame of the cached renderer on ctx's scene (all on one device):
 *  1. online training: train_paths full-length paths (train_bounces, the
 *     plain path loop, pixels drawn with `seed`) record the surface hit of
 *     radiance ray prefix_bounces + 1, where the frame's paths query the
 *     cache; its suffix radiance (final - gathered before it) / throughput
 *     is a training row (at most train_max rows), a target with the same
 *     remaining bounce budget as the queries; train_steps Adam steps on them;
 *  2. the frame: p->spp samples per pixel whose radiance ray number
 *     prefix_bounces + 1 ends at a surface hit by adding
 *     throughput * max(cache(features), 0) (misses and lights stay exact) --
 *     or, when fallback_loss > 0 and this frame's training loss exceeds it,
 *     the plain tracer with p->bounces;
 *  3. the frame is blended into the context's accum buffer as a running
 *     mean with weight 1 / (p->subframe + 1) (frame p->subframe). */
typedef struct lisa_cache_frame_params {
  uint32_t prefix_bounces; /* radiance rays traced before the cache answers (0: at the camera ray's hit) */
  uint32_t train_paths;    /* training paths per frame (0: no training; at most 2^24) */
  uint32_t train_bounces;  /* bounces of a training path */
  uint32_t train_max;      /* at most this many training rows (suffixes) per frame; 0: no cap */
  uint32_t train_steps;    /* Adam steps per frame over the frame's rows (at most 64) */
  float fallback_loss;     /* > 0: plain frame when the last training loss exceeds it */
  uint32_t flags;          /* LISA_CACHE_* */
  uint32_t nee_loops;      /* with LISA_CACHE_EMISSION: NEE runs averaged at the camera ray's
                              hit (0 or 1: the reference's one; at most 16) */
  uint32_t nee_loops_deep; /* the same at the later prefix vertices */
  uint32_t judge_pixels;   /* > 0: the plain tracer judges the cached frames on this many seeded
                              pixels (at most 2^20): its own samples there (an RNG stream apart)
                              against the mean of the cached frames of the current block of 8;
                              once the bias estimate (squared difference less both means'
                              variances) exceeds judge_rmse, from frame 16 on with >= 4 frames
                              in the block, the sequence continues with plain frames, their
                              mean restarted */
  float judge_rmse;
  uint64_t seed;           /* draws the training pixels */
} lisa_cache_frame_params;
#define LISA_CACHE_PERSISTENT 1u /* the frame's paths in a persistent launch */
/* The prefix paths add the expected emission of each Lambertian (roughness 1)
 * bounce from one light sample and draw the bounce on the condition that it
 * hits no emitter, weighted by 1 - (the sampled chance that it would): the
 * same expectation as the plain tracer's bounces, with less variance.  With
 * nee_loops = M > 1 the NEE at the camera ray's hit is the mean of M
 * independent runs of the reference's NEE loop from the vertex's hit flag
 * and light (nee_loops_deep at later prefix vertices); the last run's flag
 * and light go on down the path (same expectation). */
#define LISA_CACHE_EMISSION 2u
/* Train before the frame's paths instead of beside them.  Without a
 * fallback_loss the frame's prefix paths (which do not read the network) run
 * while the training steps run on the cache's stream; then train_ms and
 * train_kernel_ms include the wait for the paths.  This flag keeps the two
 * apart (for timing the training kernels alone). */
#define LISA_CACHE_SERIAL 8u
/* No network: the frame's paths run all p->bounces radiance rays with the
 * prefix estimators (LISA_CACHE_EMISSION, nee_loops, nee_loops_deep) at every
 * vertex and end without a query; no training, no inference.  The frames are
 * the variance-reduced plain tracer -- the method the cache is measured
 * against to tell what the network itself gains.  Without LISA_CACHE_EMISSION
 * they are the plain tracer's frames bit for bit (another RNG stream per
 * subframe, as in progressive rendering). */
#define LISA_CACHE_FULL_PATHS 16u

typedef struct lisa_cache_frame_stats {
  float trace_train_ms, train_ms, render_ms, infer_ms, total_ms;  /* host wall time */
  float loss;
  int32_t queries, train_samples, fell_back;
  float train_kernel_ms; /* HIP-event time of the training steps (forward, backward, Adam) */
  float infer_kernel_ms; /* HIP-event time of the inference launch */
  float judge_rmse_est;  /* the judge's bias estimate after this frame (-1: not judged) */
  int32_t judge_plain;   /* 1: this frame was a plain frame because the judge took over */
} lisa_cache_frame_stats;

lisa_status lisa_render_cached(lisa_ctx* ctx, lisa_cache* cache, const lisa_render_params* p,
                               const lisa_cache_frame_params* cp, lisa_cache_frame_stats* stats);

/* The two path forms of lisa_render_cached on a pixel list, for checking
 * against a CPU model.  Cache features of a surface hit (12, each in [0, 1]):
 * position in the scene box (the union of the BVH root's child boxes), the
 * incoming ray direction and the shading normal as (acos(y)/pi,
 * (atan2(z, x) + pi)/2pi), albedo rgb, roughness (0 for a dielectric),
 * dielectric flag.
 *
 * lisa_cache_trace_queries: the prefix paths of a frame with cp's
 * prefix_bounces, flags (LISA_CACHE_EMISSION, LISA_CACHE_FULL_PATHS: all
 * p->bounces rays, no query), nee_loops and nee_loops_deep
 * (the other fields are not read): p->spp samples
 * per listed pixel (seeded by p->subframe), at most prefix_bounces + 1
 * radiance rays; the surface hit of radiance ray prefix_bounces + 1 ends the
 * path in a query (features,
 * throughput, index into `pixels`).  accum_out (4n) = per-pixel mean of the
 * radiance gathered before the queries.  Queries come in the order of
 * `pixels`, a pixel's in sample order; *n_queries may exceed q_cap (only the
 * first q_cap are copied).
 *
 * lisa_cache_trace_rows: plain paths (p->spp, p->bounces) on the listed
 * pixels; one row per surface hit (depth = 0) or per surface hit of radiance
 * ray number `depth`: features and the target (path radiance - radiance
 * gathered before the hit) / throughput on arrival.  Rows come in the order
 * of `pixels`, then sample, then depth along the path; *n_rows may exceed
 * cap (only the first cap are copied). */
lisa_status lisa_cache_trace_queries(lisa_ctx* ctx, const lisa_render_params* p,
                                     const lisa_cache_frame_params* cp, const int32_t* pixels, int32_t n,
                                     float* accum_out, float* q_feat, float* q_att, int32_t* q_item,
                                     int32_t q_cap, int32_t* n_queries);
lisa_status lisa_cache_trace_rows(lisa_ctx* ctx, const lisa_render_params* p, const int32_t* pixels,
                                  int32_t n, uint32_t depth, float* feat, float* tgt, int32_t cap,
                                  int32_t* n_rows);

/* ---- debug / parity hooks (no counterpart in the reference: OptiX hides these) ---- */

/* An LBVH as built (which = 0: all T triangles; 1: the N emitting ones):
 * Morton codes and global triangle ids in sorted order (N each), child refs
 * (2 per internal node; bit 31 = leaf, low bits = sorted index;
 * 0xFFFFFFFF = empty) and child boxes (12 floats per internal node:
 * c0.lo.xyz c0.hi.xyz c1.lo.xyz c1.hi.xyz).  Internal nodes = max(N-1, 1). */
lisa_status lisa_get_bvh(lisa_ctx* ctx, int32_t which, uint32_t* morton, uint32_t* prim,
                         uint32_t* child_refs, float* child_boxes);

/* Closest hit of n rays (O, D: 3n floats): prim_out = triangle id or -1,
 * t_out = hit distance.  Closest-hit semantics of optixTrace
 * (shader.cu:77-98) with ties resolved to the lower triangle id. */
lisa_status lisa_trace_rays(lisa_ctx* ctx, const float* O, const float* D, int32_t n,
                            float tmin, float tmax, int32_t* prim_out, float* t_out);

/* The plain path loop on a list of pixels, recording per sample: radiance
 * (3 floats), radiance rays traced (bounce count), occlusion rays traced,
 * an FNV-1a hash of every ray's hit id, and the hit id of the first
 * record_bounces radiance rays (-1 miss, -2 zero/non-finite direction,
 * -3 not traced).  Any record pointer may be NULL.  accum_rgba_out gets each
 * pixel's mean over these spp samples; p->subframe only seeds the RNG
 * (shader.cu:141), there is no running mean. */
lisa_status lisa_trace_subset(lisa_ctx* ctx, const lisa_render_params* p,
                              const int32_t* pixels, int32_t n_pixels, float* accum_rgba_out,
                              float* sample_rgb, int32_t* sample_bounces,
                              int32_t* sample_occl, uint32_t* sample_hash,
                              int32_t* rad_hits, int32_t record_bounces);

/* ---- host-side pieces of the drop-in (no device needed) ---- */

/* sutil::Camera::UVWFrame (src/sutil/Camera.cpp:34-45) with up = (0,1,0),
 * fov vertical in degrees, aspect = width / height (optix_wrapper.cc:434-442). */
lisa_status lisa_camera_uvw(const float eye[3], const float look_at[3], float fov,
                            float aspect, float U[3], float V[3], float W[3]);

/* make_color (src/cuda/helpers.h:35-66) on n float4 pixels. */
lisa_status lisa_srgb_quantize(const float* rgba, int64_t n, uint8_t* rgba8);

/* sutil::saveImage for UNSIGNED_BYTE4 (src/sutil/sutil.cpp:523-600, 97-117):
 * rows flipped, P6, extension must be .ppm / .PPM. */
lisa_status lisa_save_ppm(const char* path, const uint8_t* rgba8, int32_t width, int32_t height);

/* ---- .rto scene file (SceneParser, src/LiSA/src/scene_parser.cc) ---- */
typedef struct lisa_scene_file lisa_scene_file;

typedef struct lisa_scene_info {
  int32_t num_tris, num_materials;
  uint32_t width, height, num_samples, num_bounces;
  float eye[3], look_at[3], fov;
} lisa_scene_info;

/* Parses path like SceneParser::SceneParser (scene_parser.cc:35-57) and the
 * .obj files it names (parse_obj.cc:24-69).  On failure returns
 * LISA_ERR_PARSE / LISA_ERR_IO with the reference's message in
 * lisa_last_error(NULL) instead of exit(1). */
lisa_status lisa_scene_file_load(const char* path, lisa_scene_file** out);
void lisa_scene_file_free(lisa_scene_file* f);
lisa_status lisa_scene_file_info(const lisa_scene_file* f, lisa_scene_info* info);
/* Borrowed pointers valid until lisa_scene_file_free. */
const float* lisa_scene_file_vertices(const lisa_scene_file* f);  /* 9*T */
const float* lisa_scene_file_normals(const lisa_scene_file* f);   /* 9*T */
const int32_t* lisa_scene_file_mat_indices(const lisa_scene_file* f); /* T */
const lisa_material* lisa_scene_file_materials(const lisa_scene_file* f);
const char* lisa_scene_file_output_image(const lisa_scene_file* f);

#ifdef __cplusplus
}
#endif

#endif /* LISA_H */
