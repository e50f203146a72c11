// liblisa.so device half: LBVH builder, traversal and the plain path-tracing
// megakernel for gfx950 (MI355X), plus the C-ABI entry points that drive
// them (include/lisa.h).
//
// HBM layout per scene (all 16-byte aligned, built once by lisa_set_scene):
//   nodes  float4[4 * nI]   one 64-byte line per internal node:
//                           c0.lo.xyz c0.hi.xyz c1.lo.xyz c1.hi.xyz | ref0 ref1 - -
//                           (child boxes live in the parent; ref bit 31 = leaf,
//                           low bits = sorted triangle / node index)
//   wide   float4[8 * nI]   scenes of >= 2^20 triangles: the 4-wide form (k_wide),
//                           one 128-byte line per internal node
//   tris   float4[3 * T]    48 bytes per triangle in Morton order:
//                           v0.xyz|prim id, v1.xyz|material id, v2.xyz|0
//   nrm    float4[3 * T]    vertex normals in the same order (shading gather)
//   mats   DevMat[M]        48 bytes per material
//   accum  float4[w * h]    OptixParams::accum_buffer (structs.hh:62)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lisa.h"
#include "lisa_internal.h"
#include "lisa_math.h"

using namespace lisa;

namespace {

constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
// 1 + 2*gamma(3) (Ize, "Robust BVH Ray Traversal", JCGT 2013).
constexpr float kRobust = 0x1.000006p+0f;
// Boxes are culled against the best hit distance widened by 2^-16, so a
// triangle whose hit lies on its box boundary (axis-aligned or coplanar
// geometry, where the slab entry can round a few ulp above the triangle's t)
// is still tested and equal-t hits resolve to the lower triangle id.
constexpr float kWiden = 0x1.0001p+0f;
constexpr float kTmin = 1e-4f;  // shader.cu:114, :202
constexpr int kMaxRedraws = 64;  // conditioned bounces drawn again at most this often
constexpr float kTmax = 1e16f;  // shader.cu:115, :202
constexpr int kNeeTries = 30;   // shader.cu:199
// Per-lane traversal stack: the first kStackLds entries in LDS, the rest in
// a private spill array.  The depth of a Karras tree over index-augmented
// 30-bit keys is below 64, so kStackMax entries always suffice.
#ifndef LISA_STACK_LDS
#define LISA_STACK_LDS 20
#endif
#ifndef LISA_PATH_BLOCK
#define LISA_PATH_BLOCK 64
#endif
// Waves per SIMD the plain path kernel k_path is compiled for: 6 with a
// 20-entry LDS stack (24 LDS words per lane with the lane state: 6 KB per
// 64-lane workgroup) at 80 VGPRs and no spills.  The cache-mode paths run at
// 5, except the tiled prefix-emission mode (6: its spills sit around the
// walks and the island's frames ran 2 % faster than at 5; its persistent
// form would spill far more at 6 and stays at 5), the counting and recording
// builds at 4, and the batched forms keep the registers they need.
#ifndef LISA_PATH_MIN_BLOCKS
#define LISA_PATH_MIN_BLOCKS 6
#endif
#ifndef LISA_WIDE_WAVES
#define LISA_WIDE_WAVES 6  // k_path over the 4-wide tree (5 measured 8 % slower on C3)
#endif
#ifndef LISA_MODE_WAVES
#define LISA_MODE_WAVES 5  // the cache-mode path kernels
#endif
#ifndef LISA_EMIT_WAVES
#define LISA_EMIT_WAVES 6  // the query kernel with the prefix-emission light samples
#endif
#ifndef LISA_TILE_WAVES
#define LISA_TILE_WAVES 6  // the batched wave-tile form
#endif
// The BSDF of opaque materials, chosen at compile time as the reference
// chooses it with an #include (shader.cu:4): a type with
// sample(wi, N, seed&, roughness) -> wo and eval(N, L) -> the NEE weight
// (lisa_math.h: Lambertian, bsdfs/lambertian.cu).
#ifndef LISA_BSDF
#define LISA_BSDF Lambertian
#endif
constexpr int kStackLds = LISA_STACK_LDS;
constexpr int kStackMax = 64;
// LDS words per lane of a path kernel: the traversal stack, then the pixel's
// running sum of sample radiance (3 floats, touched once per sample) and the
// NEE light material (written when a NEE ray finds a light) -- in LDS so
// that they hold no register through the path loop.  At 20 + 4 words a
// 64-lane workgroup takes 6 KB, 24 of them (6 waves per SIMD) 144 KB
// (6.5 KB per workgroup measured 7 % slower: fewer workgroups fit).  The
// benchmark scenes' traversals need at most 18 stack entries.
constexpr int kLaneLds = kStackLds + 4;
// Path workgroup: kPathBlock work-items on a kTileW x kTileH pixel tile
// (64: 8x8, 128: 16x8, 256: 16x16), so a wave covers a compact 8-row or
// 4-row block of pixels whose camera rays start coherent.
constexpr int kPathBlock = LISA_PATH_BLOCK;
constexpr int kTileW = kPathBlock >= 128 ? 16 : 8;
constexpr int kTileH = kPathBlock / kTileW;
constexpr int kRayBlock = 256;
static_assert(kStackLds >= 1 && kStackLds <= kStackMax, "LDS stack depth");
static_assert(kPathBlock == 64 || kPathBlock == 128 || kPathBlock == 256, "path block size");

struct DevMat {
  float4 diffuse_alpha;  // diffuse.xyz, alpha
  float4 emission_emit;  // emission.xyz, emit (1/0)
  float4 rough_n;        // roughness, n, -, -
};

// One device LBVH: 64-byte nodes + 48-byte Morton-ordered triangle records.
struct BvhView {
  const float4* __restrict__ nodes;
  const float4* __restrict__ tris;
  int32_t T;
  const float4* __restrict__ wide;  // the 4-wide form of `nodes` (null: none)
};

struct SceneView {
  BvhView all;   // every triangle
  BvhView emit;  // emitting triangles only (NEE light queries)
  const float4* __restrict__ nrm;  // vertex normals, Morton order of `all`
  const DevMat* __restrict__ mats;
  float4 ecull_c, ecull_h;  // emitter cull box: centre, inflated half extents (emit_culled)
};

// ---------------------------------------------------------------------------
// Traversal: stack-based BVH2 walk, near child first, leaves tested as soon
// as their box (stored in the parent) is hit; watertight FP32 ray/triangle
// test (Woop, Benthin & Wald, JCGT 2013).  Closest hit over t in
// (tmin, tmax], equal t resolved to the lower triangle id, so the answer does
// not depend on traversal order.
// ---------------------------------------------------------------------------
struct RayT {
  F3 O, inv;
  float Sx, Sy, Sz;
  int kx, ky, kz;
};

__device__ __forceinline__ bool ray_valid(F3 D) {
  return isfinite(D.x) && isfinite(D.y) && isfinite(D.z) &&
         (D.x != 0.0f || D.y != 0.0f || D.z != 0.0f);
}

__device__ __forceinline__ RayT make_ray(F3 O, F3 D) {
  RayT r;
  r.O = O;
  r.inv = mk(1.0f / D.x, 1.0f / D.y, 1.0f / D.z);
  const float ax = fabsf(D.x), ay = fabsf(D.y), az = fabsf(D.z);
  int kz = 0;
  float m = ax;
  if (ay > m) {
    kz = 1;
    m = ay;
  }
  if (az > m) kz = 2;
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  const float dz = sel(D, kz);
  if (dz < 0.0f) {
    const int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.Sx = sel(D, kx) / dz;
  r.Sy = sel(D, ky) / dz;
  r.Sz = sel(r.inv, kz);  // 1 / D[kz], already rounded once for the slab tests
  return r;
}

__device__ __forceinline__ bool watertight(const RayT& r, F3 a, F3 b, F3 c, float& t) {
  const F3 A = a - r.O, B = b - r.O, C = c - r.O;
  const float Akz = sel(A, r.kz), Bkz = sel(B, r.kz), Ckz = sel(C, r.kz);
  const float Ax = sel(A, r.kx) - r.Sx * Akz;
  const float Ay = sel(A, r.ky) - r.Sy * Akz;
  const float Bx = sel(B, r.kx) - r.Sx * Bkz;
  const float By = sel(B, r.ky) - r.Sy * Bkz;
  const float Cx = sel(C, r.kx) - r.Sx * Ckz;
  const float Cy = sel(C, r.ky) - r.Sy * Ckz;
  // FP32 edge functions only (no double-precision recomputation of a zero
  // edge function): the edge shared by two triangles gets exactly negated
  // values in the two (the same per-vertex products, subtracted in the other
  // order), so a ray on either side is accepted by one of them and a ray on
  // the edge (a zero) by both -- none falls through.
  const float U = Cx * By - Cy * Bx;
  const float V = Ax * Cy - Ay * Cx;
  const float W = Bx * Ay - By * Ax;
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = U + V + W;
  if (det == 0.0f) return false;
  const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  const float T = U * Az + V * Bz + W * Cz;
  t = T / det;
  return true;
}

__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                     const RayT& r, float tmin, float tmax, float& tn_out) {
  const float tx0 = (lx - r.O.x) * r.inv.x, tx1 = (hx - r.O.x) * r.inv.x;
  const float ty0 = (ly - r.O.y) * r.inv.y, ty1 = (hy - r.O.y) * r.inv.y;
  const float tz0 = (lz - r.O.z) * r.inv.z, tz1 = (hz - r.O.z) * r.inv.z;
  const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  const float tfb = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1)) * kRobust;
  const float tf = fminf(tfb, tmax);
  tn_out = tn;
  return tn <= tf;
}

// Division-free cull of NEE's emitter query.  The box (c, H) holds every
// emitter, inflated by twice a margin m = S * 2^-12 per side, where S is the
// largest |coordinate| of the scene box.  A ray whose origin lies in the
// scene box and that misses the inflated box by a separating axis test
// evaluated in FP32 misses the box inflated by m in exact arithmetic; the
// slab tests of the emitter root's child boxes (each inside the box) then
// miss in FP32 too -- their rounding error is below 16 S u / |D_i| (u = 2^-24)
// per t value against a gap of at least m / |D_i| + m / |D_j| -- so the query
// has the traversal's answer, "no emitter", without the ray's divisions.
//
// line_culled: the line through O along d misses the box (separating axes
// d x e_x, d x e_y, d x e_z).  It does not depend on the length or sign of
// d, so it can run on hemisphere()'s raw draws before they are normalised:
// the normalised direction is parallel to them up to a few ulps, an angle
// whose offset over the scene (< 16 S u) the margin absorbs.  A direction
// with a zero component is never culled.
__device__ __forceinline__ bool line_culled(float4 c, float4 H, F3 O, F3 d) {
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  if (ax == 0.0f || ay == 0.0f || az == 0.0f) return false;
  const float vx = c.x - O.x, vy = c.y - O.y, vz = c.z - O.z;
  if (fabsf(vy * d.z - vz * d.y) > H.y * az + H.z * ay) return true;
  if (fabsf(vz * d.x - vx * d.z) > H.z * ax + H.x * az) return true;
  if (fabsf(vx * d.y - vy * d.x) > H.x * ay + H.y * ax) return true;
  return false;
}
// emit_culled: the ray misses the box -- the box lies behind the origin along
// an axis the ray moves away on, or the ray's line misses it.
__device__ __forceinline__ bool emit_culled(float4 c, float4 H, F3 O, F3 D) {
  if (D.x == 0.0f || D.y == 0.0f || D.z == 0.0f) return false;
  const float vx = c.x - O.x, vy = c.y - O.y, vz = c.z - O.z;
  if (D.x > 0.0f ? vx < -H.x : vx > H.x) return true;
  if (D.y > 0.0f ? vy < -H.y : vy > H.y) return true;
  if (D.z > 0.0f ? vz < -H.z : vz > H.z) return true;
  return line_culled(c, H, O, D);
}

// The prefix light sample's estimate p^ = q |cos_y| A / r^2 stays below 1/2
// at O: q <= 3 sqrt(3) / 12 < 0.4331 (the cube-hemisphere density at the
// cube's diagonal), |cos_y| <= 1, and every emitter point y lies in the box
// (c, H), so r^2 >= d^2, d = O's distance to the box; 0.9 A <= d^2 gives
// p^ <= 0.4331 / 0.9 < 0.49.  A function of the vertex alone, so choosing
// the estimator by it changes no expectation.
__device__ __forceinline__ bool light_sample_bounded(float4 c, float4 H, F3 O, float A) {
  const float dx = fmaxf(fabsf(O.x - c.x) - H.x, 0.0f);
  const float dy = fmaxf(fabsf(O.y - c.y) - H.y, 0.0f);
  const float dz = fmaxf(fabsf(O.z - c.z) - H.z, 0.0f);
  const float d2 = dx * dx + dy * dy + dz * dz;
  return 0.9f * A <= d2;
}

struct TravCount {
  uint64_t visits = 0, tests = 0;
};

// Closest hit (any = false) or any hit (any = true) over `bv`: returns the
// sorted index of the accepted triangle (-1 when none); t_out its t.  A
// triangle is accepted when t > tmin and (t < best_t or (t == best_t and
// id < best)), starting from best_t = tlim, best = lim_prim: the closest hit
// resolves equal t to the lower triangle id; the any-hit query answers "is
// any triangle strictly before (tlim, lim_prim)" and stops at the first.
// `stk` points at this lane's first LDS stack slot; consecutive entries are
// kStride words apart, so a wave's 64 lanes hit 64 distinct banks.
template <bool kCount, int kStride>
__device__ __forceinline__ int32_t closest(const BvhView& bv, const RayT& r, float tmin,
                                           float tlim, uint32_t lim_prim, bool any, float& t_out,
                                           uint32_t* stk, TravCount& cnt) {
  // Entries past the LDS part of the stack.  volatile: the compiler keeps
  // the pop's LDS read and private read apart instead of merging them into
  // one generic (flat) load through a selected pointer.  (An empty asm with
  // a memory clobber does the same but makes the batched kernels re-read
  // their scene views, which live in global memory, after every pop.)
  volatile uint32_t spill[kStackMax - kStackLds];
  float best_t = tlim;
  float cull_t = tlim * kWiden;
  uint32_t best_prim = lim_prim;
  int32_t best_k = -1;
  int sp = 0;
  uint32_t node = 0;
  // A tree visits each internal node at most once; the bound only guards
  // against a corrupt hierarchy (it never triggers on a valid one).
  const uint32_t max_visits = (uint32_t)bv.T + 1u;
  for (uint32_t it = 0; it < max_visits; ++it) {
    if (kCount) cnt.visits++;
    const float4* np = bv.nodes + 4 * (size_t)node;
    const float4 a = np[0], b = np[1], c = np[2], d = np[3];
    const uint32_t r0 = __float_as_uint(d.x), r1 = __float_as_uint(d.y);
    float tn0 = 0.0f, tn1 = 0.0f;
    // both slab tests unconditionally (no short circuit), so the node's four
    // 16-byte loads are issued together and waited for once
    bool h0 = (r0 != kEmpty) & slab(a.x, a.y, a.z, a.w, b.x, b.y, r, tmin, cull_t, tn0);
    bool h1 = (r1 != kEmpty) & slab(b.z, b.w, c.x, c.y, c.z, c.w, r, tmin, cull_t, tn1);
#pragma unroll
    for (int ch = 0; ch < 2; ++ch) {
      const bool hit = ch == 0 ? h0 : h1;
      const uint32_t ref = ch == 0 ? r0 : r1;
      if (hit && (ref & kLeaf)) {
        if (kCount) cnt.tests++;
        const uint32_t k = ref & ~kLeaf;
        const float4* tp = bv.tris + 3 * (size_t)k;
        const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
        const uint32_t prim = __float_as_uint(t0.w);
        float t;
        if (watertight(r, xyz(t0), xyz(t1), xyz(t2), t) && t > tmin &&
            (t < best_t || (t == best_t && prim < best_prim))) {
          best_t = t;
          cull_t = t * kWiden;
          best_prim = prim;
          best_k = (int32_t)k;
        }
        if (ch == 0)
          h0 = false;
        else
          h1 = false;
      }
    }
    if (any && best_k >= 0) break;
    if (h0 && h1) {
      uint32_t nr = r0, fr = r1;
      if (tn1 < tn0) {
        nr = r1;
        fr = r0;
      }
      if (sp == kStackMax) break;  // unreachable for a valid tree (depth < 64)
      if (sp < kStackLds)
        stk[sp * kStride] = fr;
      else
        spill[sp - kStackLds] = fr;
      ++sp;
      node = nr;
    } else if (h0) {
      node = r0;
    } else if (h1) {
      node = r1;
    } else {
      if (sp == 0) break;
      --sp;
      if (sp < kStackLds)
        node = stk[sp * kStride];
      else
        node = spill[sp - kStackLds];
    }
  }
  t_out = best_t;
  return best_k;
}

// The closest / any-hit query of closest() over the 4-wide form of the tree
// (k_wide): up to four slab tests per visit, leaves tested at once, the
// nearest inner child visited next and the others pushed far to near.  The
// accepted triangle is closest()'s (the same acceptance rule, which does not
// depend on the order of the tests); the visits differ, so the counting
// builds keep closest().  The stack holds up to 3 entries per level of a
// tree of at most 32 levels.  Fewer, wider visits pay where the tree streams
// from HBM (large scenes); on a cache-resident tree the binary walk issues
// fewer instructions.
constexpr int kWideStackMax = 96;
template <int kStride>
__device__ __forceinline__ int32_t closest_wide(const BvhView& bv, const RayT& r, float tmin,
                                                float tlim, uint32_t lim_prim, bool any, float& t_out,
                                                uint32_t* stk) {
  volatile uint32_t spill[kWideStackMax - kStackLds];
  float best_t = tlim;
  float cull_t = tlim * kWiden;
  uint32_t best_prim = lim_prim;
  int32_t best_k = -1;
  int sp = 0;
  uint32_t node = 0;
  const uint32_t max_visits = (uint32_t)bv.T + 1u;
  for (uint32_t it = 0; it < max_visits; ++it) {
    const float4* np = bv.wide + 8 * (size_t)node;
    const float4 lx = np[0], ly = np[1], lz = np[2], hx = np[3], hy = np[4], hz = np[5], rf = np[6];
    uint32_t ref[4] = {__float_as_uint(rf.x), __float_as_uint(rf.y), __float_as_uint(rf.z),
                       __float_as_uint(rf.w)};
    float tn[4];
    bool h[4];
    h[0] = (ref[0] != kEmpty) & slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, r, tmin, cull_t, tn[0]);
    h[1] = (ref[1] != kEmpty) & slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, r, tmin, cull_t, tn[1]);
    h[2] = (ref[2] != kEmpty) & slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, r, tmin, cull_t, tn[2]);
    h[3] = (ref[3] != kEmpty) & slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, r, tmin, cull_t, tn[3]);
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
      if (h[ch] && (ref[ch] & kLeaf)) {
        const uint32_t k = ref[ch] & ~kLeaf;
        const float4* tp = bv.tris + 3 * (size_t)k;
        const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
        const uint32_t prim = __float_as_uint(t0.w);
        float t;
        if (watertight(r, xyz(t0), xyz(t1), xyz(t2), t) && t > tmin &&
            (t < best_t || (t == best_t && prim < best_prim))) {
          best_t = t;
          cull_t = t * kWiden;
          best_prim = prim;
          best_k = (int32_t)k;
        }
        h[ch] = false;
      }
    }
    if (any && best_k >= 0) break;
    // inner children still before the (possibly nearer) best hit, nearest first
    float key[4];
    int m = 0;
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
      h[ch] = h[ch] && tn[ch] <= cull_t;
      key[ch] = h[ch] ? tn[ch] : INFINITY;
      m += h[ch] ? 1 : 0;
    }
    if (m == 0) {
      if (sp == 0) break;
      --sp;
      if (sp < kStackLds)
        node = stk[sp * kStride];
      else
        node = spill[sp - kStackLds];
      continue;
    }
    // sorting network on (key, ref): (0,1) (2,3) (0,2) (1,3) (1,2)
    auto cswap = [&](int i, int j) {
      const bool sw = key[j] < key[i];
      const float ki = key[i], kj = key[j];
      const uint32_t ri = ref[i], rj = ref[j];
      key[i] = sw ? kj : ki;
      key[j] = sw ? ki : kj;
      ref[i] = sw ? rj : ri;
      ref[j] = sw ? ri : rj;
    };
    cswap(0, 1);
    cswap(2, 3);
    cswap(0, 2);
    cswap(1, 3);
    cswap(1, 2);
    if (sp + 3 > kWideStackMax) break;  // unreachable for a valid tree (depth < 64)
#pragma unroll
    for (int q = 3; q >= 1; --q) {
      if (q < m) {
        if (sp < kStackLds)
          stk[sp * kStride] = ref[q];
        else
          spill[sp - kStackLds] = ref[q];
        ++sp;
      }
    }
    node = ref[0];
  }
  t_out = best_t;
  return best_k;
}

// ---------------------------------------------------------------------------
// Ray batch kernel (hit-id parity, lisa_trace_rays).
// ---------------------------------------------------------------------------
template <bool kCount>
__global__ __launch_bounds__(kRayBlock) void k_trace_rays(SceneView sc, const float* __restrict__ O,
                                                          const float* __restrict__ D, int32_t n,
                                                          float tmin, float tmax,
                                                          int32_t* __restrict__ prim_out,
                                                          float* __restrict__ t_out,
                                                          unsigned long long* counters) {
  __shared__ uint32_t stk[kStackLds * kRayBlock];
  const int32_t i = blockIdx.x * kRayBlock + threadIdx.x;
  if (i >= n) return;
  const F3 o = mk(O[3 * i], O[3 * i + 1], O[3 * i + 2]);
  const F3 d = mk(D[3 * i], D[3 * i + 1], D[3 * i + 2]);
  TravCount cnt;
  int32_t prim = -1;
  float t = 0.0f;
  if (sc.all.T > 0 && ray_valid(d)) {
    const RayT r = make_ray(o, d);
    int32_t k = closest<kCount, kRayBlock>(sc.all, r, tmin, tmax, 0xFFFFFFFFu, false, t,
                                           stk + threadIdx.x, cnt);
    // with a 4-wide tree, the hit of its walk (which the path kernels use;
    // the counts stay the binary walk's)
    if (sc.all.wide) k = closest_wide<kRayBlock>(sc.all, r, tmin, tmax, 0xFFFFFFFFu, false, t, stk + threadIdx.x);
    if (k >= 0) prim = (int32_t)__float_as_uint(sc.all.tris[3 * (size_t)k].w);
  }
  prim_out[i] = prim;
  t_out[i] = prim >= 0 ? t : 0.0f;
  if (kCount) {
    atomicAdd(&counters[2], (unsigned long long)cnt.visits);
    atomicAdd(&counters[3], (unsigned long long)cnt.tests);
  }
}

// ---------------------------------------------------------------------------
// Plain megakernel: one work-item per pixel, all spp in one launch
// (shader.cu:126-167), the bounce loop (102-124), closest-hit / miss
// programs (172-255) and NEE (196-209) as one per-lane state machine.  Every
// iteration traces exactly one ray, radiance or occlusion, so lanes of a
// wave that are at different stages of their paths still traverse together.
// ---------------------------------------------------------------------------
struct PathArgs {
  uint32_t width, height, spp, bounces, subframe;
  F3 eye, U, V, W;
  const int32_t* pixels;  // nullptr: full image in 16x8 tiles
  int32_t n_items;
  uint32_t tiles_x;
  float4* accum;
  const float4* accum_prev;
  int32_t fp64;  // FP64 per-pixel sums: the launch picks the k_path<.., .., true> instantiation
  double4* sum64;  // FP64: the image's sums and sample counts across subframes (full renders)
  float* s_rgb;
  int32_t* s_b;
  int32_t* s_o;
  uint32_t* s_h;
  int32_t* rad_hits;
  int32_t rb;
  unsigned long long* counters;
  // learned radiance cache (path_lane's kMode 1 and 2)
  uint32_t cache_at;  // kMode 1: this radiance ray's surface hit queries the cache
  // kMode 1: one query slot per sample (item * spp + sample): 4 float4 (hit
  // point and material index bits, incoming direction, shading normal,
  // throughput) and a flag set to 1 when the sample queried (zeroed before
  // the launch)
  float4* q_rec;
  uint8_t* q_has;
  float* t_rec;       // kMode 2: 18 floats per vertex slot (12 features, 3 throughput, 3 radiance so far)
  uint32_t t_slots;   // kMode 2: slots per training path
  float* t_row;       // kMode 2: 15 floats (features, target) per row slot ((item * spp + sample) * t_slots + vertex)
  uint8_t* t_n;       // kMode 2: rows of each sample (zeroed before the launch)
  uint32_t t_depth;   // kMode 2: > 0 records only the hits of radiance ray number t_depth
  int32_t emit_expect;      // queries: path_lane's kMode 3 instead of kMode 1
  int32_t nee_loops;        // kMode 3: NEE runs averaged at the camera ray's hit (>= 1)
  int32_t nee_loops_deep;   // kMode 3: ... at the later vertices (>= 1)
  const float* emit_cdf;    // area prefix sums of the emitting triangles (emitter tree order)
  float emit_area;          // their total area (the last prefix sum)
  float emit_area_below;    // the largest float below emit_area
  F3 bmin, binv;      // position normalisation for the features
};

// One slot of a device counter for every active lane, with one atomic per
// wave: the lowest active lane reserves the whole group, each lane takes its
// rank in it.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter) {
  const uint64_t act = __ballot(1);
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)act) - 1u;
  const uint32_t rank = (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(act));
  base = __shfl(base, (int)leader);
  return base + rank;
}

// Cache features of a surface hit (lisa.h, lisa_cache_*): position in the
// scene box, incoming direction and shading normal as (theta/pi, phi/2pi),
// albedo, roughness, dielectric flag -- each in [0, 1].
__device__ __forceinline__ void cache_features(F3 bmin, F3 binv, F3 P, F3 D, F3 N, const DevMat& m,
                                               float* f) {
  f[0] = fminf(fmaxf((P.x - bmin.x) * binv.x, 0.0f), 1.0f);
  f[1] = fminf(fmaxf((P.y - bmin.y) * binv.y, 0.0f), 1.0f);
  f[2] = fminf(fmaxf((P.z - bmin.z) * binv.z, 0.0f), 1.0f);
  const float dl = sqrtf(dot(D, D));
  const float dy = dl > 0.0f ? D.y / dl : 0.0f;
  f[3] = acosf(fminf(fmaxf(dy, -1.0f), 1.0f)) * 0.31830988618f;
  f[4] = (atan2f(D.z, D.x) + 3.14159265359f) * 0.15915494309f;
  f[5] = acosf(fminf(fmaxf(N.y, -1.0f), 1.0f)) * 0.31830988618f;
  f[6] = (atan2f(N.z, N.x) + 3.14159265359f) * 0.15915494309f;
  const bool diel = m.diffuse_alpha.w < 1.0f;
  f[7] = fminf(fmaxf(m.diffuse_alpha.x, 0.0f), 1.0f);
  f[8] = fminf(fmaxf(m.diffuse_alpha.y, 0.0f), 1.0f);
  f[9] = fminf(fmaxf(m.diffuse_alpha.z, 0.0f), 1.0f);
  f[10] = diel ? 0.0f : fminf(fmaxf(m.rough_n.x, 0.0f), 1.0f);
  f[11] = diel ? 1.0f : 0.0f;
}

// A pixel of work for one lane: image coordinates and the index of its
// result in the launch's accum / record arrays.
struct Job {
  uint32_t x, y;
  int32_t item;
};

// The per-lane path loop over the pixels a source hands out.  `src.view()`
// and `src.args()` point at the scene and launch of the current pixel (read
// again after every `src.next(j)`, which gives the lane its next pixel;
// false: no work left).
// kMode 0: the plain path loop.  kMode 1: the radiance ray numbered
// p.cache_at ends at a surface hit by writing a cache query (the hit, the
// path throughput) at the sample's slot; misses and emitters are handled
// exactly.
// kMode 3: kMode 1 with the emission that the prefix's diffuse bounces would
// hit taken in expectation (a light sample at each Lambertian vertex, the
// bounce drawn on the condition that it hits no emitter) and each vertex's
// NEE the mean of nee_loops (first hit) or nee_loops_deep runs; the same
// expectation as kMode 1.
// kMode 2: every surface hit of the (full-length) path (or, with t_depth > 0,
// the hit of radiance ray number t_depth only) records its features,
// throughput and the radiance gathered so far; at the end of the sample the
// suffix radiance (final - so far) / throughput of each vertex becomes a
// training row at the sample's row slots.  Without accum_prev the launch writes its own mean (no
// running mean over subframes; the subframe then only picks the RNG stream).
template <bool kRecord, bool kCount, class Src, int kMode = 0, bool kF64 = false, bool kWide = false,
          class Bsdf = LISA_BSDF>
__device__ __forceinline__ void path_lane(Src& src, uint32_t* my_stk) {
  Job job;
  if (!src.next(job)) return;
  const SceneView* scp = src.view();  // the current pixel's scene and launch
  const PathArgs* pp = src.args();
  uint32_t x = job.x, y = job.y;
  int32_t item = job.item;

  uint32_t seed = tea16(y * pp->width + x, pp->subframe);  // shader.cu:141

  uint64_t c_rad = 0, c_occ = 0, c_rhit = 0, c_inv = 0, c_smp = 0, c_px = 0;
  TravCount tc;

  // the pixel's sum of sample radiance, in LDS after the lane's stack
  float* const acc = reinterpret_cast<float*>(my_stk + kStackLds * kPathBlock);
  acc[0] = 0.0f;
  acc[kPathBlock] = 0.0f;
  acc[2 * kPathBlock] = 0.0f;
  double ad0 = 0.0, ad1 = 0.0, ad2 = 0.0;

  // Per-lane path state (RayState, shader.cu:30-42, plus the NEE loop index).
  // O is the ray origin, and from a surface hit on the hit point (the origin
  // of the NEE rays and of the next bounce).
  F3 O = mk(0.0f, 0.0f, 0.0f), D = O, color = O, att = O, N = O, Din = O;
  bool hitf = false;
  int32_t qstage = 0;  // 0 radiance ray, 1 NEE emitter query, 2 NEE any-hit query
  int32_t tries = 0, matE = -1;
  // RayState::material of the NEE's light (shader.cu:38), in LDS after acc
  int32_t* const light_mat = reinterpret_cast<int32_t*>(my_stk + (kStackLds + 3) * kPathBlock);
  *light_mat = 0;
  uint32_t primE = 0xFFFFFFFFu;
  float tE = kTmax;
  float rough = 0.0f;
  uint32_t b = 0, s = 0;
  uint32_t hash = 0;
  int32_t occl_cnt = 0;
  size_t rec_base = (size_t)item * pp->spp;

  constexpr bool kQuery = kMode == 1 || kMode == 3;  // ends at the cache query
  constexpr bool kEmit = kMode == 3;                 // prefix emission in expectation
  uint32_t t_used = 0;  // kMode 2: vertex slots recorded in this sample
  // kMode 3: the current radiance ray is a bounce drawn on
  // the condition that it hits no emitter (redrawn when it does); lw is the
  // pending light sample's weight
  bool cond = false;
  int32_t rejects = 0, ls_k = 0;  // ls_k: the light sample's emitter triangle
  float lw = 0.0f;
  // kMode 3 with nee_loops > 1: the vertex's entry light and flag (light * 2
  // + flag) and the NEE runs done
  int32_t ent = 0, nl = 0;
  // accum_color += trace(...) (shader.cu:154) and the per-sample records.
  auto finish_sample = [&]() {
    if (kMode == 2 && t_used > 0) {
      // the sample's rows at its own slots; suffix radiance per unit throughput
      const size_t base = (size_t)item * pp->t_slots;
      float* row = pp->t_row + 15 * (rec_base + s) * pp->t_slots;
      for (uint32_t v = 0; v < t_used; ++v, row += 15) {
        const float* r = pp->t_rec + 18 * (base + v);
        for (int q = 0; q < 12; ++q) row[q] = r[q];
        row[12] = (color.x - r[15]) / r[12];
        row[13] = (color.y - r[16]) / r[13];
        row[14] = (color.z - r[17]) / r[14];
      }
      pp->t_n[rec_base + s] = (uint8_t)t_used;
      t_used = 0;
    }
    if (kCount) c_smp++;
    acc[0] = acc[0] + color.x;
    acc[kPathBlock] = acc[kPathBlock] + color.y;
    acc[2 * kPathBlock] = acc[2 * kPathBlock] + color.z;
    if (kF64) {
      ad0 += (double)color.x;
      ad1 += (double)color.y;
      ad2 += (double)color.z;
    }
    if (kRecord) {
      if (pp->s_rgb) {
        pp->s_rgb[3 * (rec_base + s)] = color.x;
        pp->s_rgb[3 * (rec_base + s) + 1] = color.y;
        pp->s_rgb[3 * (rec_base + s) + 2] = color.z;
      }
      if (pp->s_b) pp->s_b[rec_base + s] = (int32_t)b;
      if (pp->s_o) pp->s_o[rec_base + s] = occl_cnt;
      if (pp->s_h) pp->s_h[rec_base + s] = hash;
    }
    ++s;
  };
  // The camera ray of the next sample (shader.cu:147-152); samples with no
  // bounce finish at once.  False when the pixel has no sample left.
  auto begin_sample = [&]() -> bool {
    while (s < pp->spp) {
      const float jx = rng(seed);
      const float jy = rng(seed);
      const float dx = (2.0f * (float)x + jx) / (float)pp->width - 1.0f;
      const float dy = (2.0f * (float)y + jy) / (float)pp->height - 1.0f;
      D = normalize(dx * pp->U + dy * pp->V + pp->W);
      O = pp->eye;
      color = mk(0.0f, 0.0f, 0.0f);
      att = mk(1.0f, 1.0f, 1.0f);
      hitf = false;
      qstage = 0;
      b = 0;
      if (kEmit) cond = false;
      if (kRecord) {
        hash = 2166136261u;
        occl_cnt = 0;
        if (pp->rad_hits)
          for (int32_t q = 0; q < pp->rb; ++q) pp->rad_hits[(rec_base + s) * pp->rb + q] = -3;
      }
      if (pp->bounces > 0) return true;
      finish_sample();
    }
    return false;
  };
  // End of one shoot_ray_to_light: add its light, sample the BSDF for the
  // next bounce (shader.cu:251-252), end the sample after the last bounce.
  auto end_nee = [&](F3 nee) -> bool {
    if (kEmit) {
      // kMode 3: the light of one NEE run (the mean of nee_loops runs at the
      // camera ray's hit, nee_loops_deep at later vertices, each added as it
      // ends); the rest of the vertex follows the NEE loop below
      // (qstage 5), in one place for every way a run ends
      const int32_t runs = b > 1 ? pp->nee_loops_deep : pp->nee_loops;
      if (runs > 1)
        color = color + (nee * (1.0f / (float)runs)) * att;
      else
        color = color + nee * att;
      qstage = 5;
      return true;
    }
    color = color + nee * att;
    D = Bsdf::sample(Din, N, seed, rough);
    qstage = 0;
    if (b == pp->bounces) {
      finish_sample();
      return begin_sample();
    }
    return true;
  };
  // One occlusion ray resolved (record `code`): return the light, try again,
  // or give up after kNeeTries (shader.cu:200-208).
  auto nee_after = [&](int32_t code) -> bool {
    if (kRecord) hash = (hash ^ ((uint32_t)code ^ 0x5bd1e995u)) * 16777619u;
    if (hitf) return end_nee(xyz(scp->mats[*light_mat].emission_emit) * Bsdf::eval(N, D));
    if (++tries == kNeeTries) return end_nee(mk(0.0f, 0.0f, 0.0f));
    D = hemisphere_raw(seed);  // finished in the emitter-query loop
    qstage = 1;
    return true;
  };

  // shader.cu:156-166: the pixel's mean (and running mean over subframes).
  // With FP64 sums kept across subframes (sum64), the mean is over every
  // sample of subframes 0..subframe instead of a running FP32 lerp.
  auto write_pixel = [&]() {
    F3 out;
    bool exact_mean = false;
    if (kF64) {
      double s0 = ad0, s1 = ad1, s2 = ad2, n = (double)pp->spp;
      if (pp->sum64) {
        if (pp->subframe > 0) {
          const double4 prev = pp->sum64[item];
          s0 = prev.x + s0;
          s1 = prev.y + s1;
          s2 = prev.z + s2;
          n = prev.w + n;
        }
        pp->sum64[item] = make_double4(s0, s1, s2, n);
        exact_mean = true;
      }
      const double inv = 1.0 / n;
      out = mk((float)(s0 * inv), (float)(s1 * inv), (float)(s2 * inv));
    } else {
      const float inv = 1.0f / (float)pp->spp;
      out = mk(acc[0], acc[kPathBlock], acc[2 * kPathBlock]) * inv;
    }
    if (kMode == 0 && pp->subframe > 0 && pp->accum_prev && !exact_mean) {
      const float a = 1.0f / (float)(pp->subframe + 1);
      out = lerp(xyz(pp->accum_prev[item]), out, a);
    }
    pp->accum[item] = make_float4(out.x, out.y, out.z, 1.0f);
    if (kCount) c_px++;
  };

  bool alive = begin_sample();
  for (;;) {
    if (!alive) {
      write_pixel();
      if (!src.next(job)) break;
      scp = src.view();
      pp = src.args();
      x = job.x;
      y = job.y;
      item = job.item;
      seed = tea16(y * pp->width + x, pp->subframe);
      acc[0] = 0.0f;
      acc[kPathBlock] = 0.0f;
      acc[2 * kPathBlock] = 0.0f;
      ad0 = ad1 = ad2 = 0.0;
      s = 0;
      rec_base = (size_t)item * pp->spp;
      alive = begin_sample();
      continue;
    }
    // -- NEE occlusion rays, first query: the closest emitter (emitter BVH).
    // A miss and a non-emitting closest hit both leave a false hit flag
    // false (shader.cu:172-184), so with the flag false a ray that hits no
    // emitter is resolved here; these cheap tries loop without waiting for
    // the wave's full traversals.
    // D holds hemisphere()'s raw draws; with the flag false a try whose line
    // misses every emitter ends here without normalising them.
    while (qstage == 1) {
      if (kCount) c_occ++;
      if (kRecord) occl_cnt++;
      if (!hitf && line_culled(scp->ecull_c, scp->ecull_h, O, D)) {
        alive = nee_after(-5);
        if (!alive) break;
        continue;
      }
      D = hemisphere_dir(D, N);
      const bool valid = ray_valid(D);
      if (kCount && !valid) c_inv++;
      int32_t k = -1;
      float t = 0.0f;
      if (valid && scp->emit.T > 0 && !emit_culled(scp->ecull_c, scp->ecull_h, O, D)) {
        const RayT r = make_ray(O, D);
        k = closest<kCount, kPathBlock>(scp->emit, r, kTmin, kTmax, 0xFFFFFFFFu, false, t, my_stk, tc);
      }
      if (!valid) {  // counts as a ray, misses without traversal
        const int32_t code = hitf ? -1 : -5;
        hitf = false;
        alive = nee_after(code);
      } else if (k >= 0) {  // an emitter: is anything in front of it?
        const float4 e0 = scp->emit.tris[3 * (size_t)k], e1 = scp->emit.tris[3 * (size_t)k + 1];
        tE = t;
        primE = __float_as_uint(e0.w);
        matE = (int32_t)__float_as_uint(e1.w);
        qstage = 2;
      } else if (!hitf) {
        alive = nee_after(-5);
      } else {  // stale hit flag: a miss clears it, any other hit keeps it
        tE = kTmax;
        primE = 0xFFFFFFFFu;
        matE = -1;
        qstage = 2;
      }
      if (!alive) break;
    }
    if (!alive) continue;  // the pixel is written at the top of the loop
    if (kEmit && qstage == 5) {
      // kMode 3, the end of a vertex's NEE run: another run from the
      // vertex's entry flag and light (the last run's flag and light stay),
      // then the light sample or the bounce
      if (++nl < (b > 1 ? pp->nee_loops_deep : pp->nee_loops)) {
        hitf = (ent & 1) != 0;
        *light_mat = ent >> 1;
        tries = 0;
        D = hemisphere_raw(seed);
        qstage = 1;
        continue;
      }
      if (rough == 1.0f && scp->emit.T > 0 && pp->emit_area > 0.0f && b < pp->bounces &&
          light_sample_bounded(scp->ecull_c, scp->ecull_h, O, pp->emit_area)) {
        // Expected emission of this vertex's bounce ray.  The reference adds
        // Le * att when the bounce (a hemisphere draw at roughness 1, density
        // q = 1 / (12 max|w_i|^3) on N's side) hits an emitter, and ends the
        // path.  Here one light sample estimates p = P(the bounce's closest
        // hit emits) and a = E[Le; hit] without bias: p^ = V q |cos_y| A / r^2,
        // a^ = Le p^; the path adds a^ * att, continues with att * (1 - p^),
        // and its bounce is drawn again whenever it hits an emitter (a draw on
        // the condition that it does not).  The expectation is the reference's.
        // light_sample_bounded keeps p^ < 1/2 (so the throughput stays
        // positive and the redraw limit ends a path with probability < 2^-64);
        // vertices closer to the emitters take the reference's bounce.
        const float u0 = rnd(seed), u1 = rnd(seed), u2 = rnd(seed);
        const int32_t T = scp->emit.T;
        // u0 * A can round up to A: keep ta below the last prefix sum, so the
        // search ends on a triangle whose prefix sum grows (positive area)
        const float ta = fminf(u0 * pp->emit_area, pp->emit_area_below);
        int32_t lo = 0, hi = T - 1;  // first triangle whose prefix sum exceeds ta
        while (lo < hi) {
          const int32_t mid = (lo + hi) >> 1;
          if (ta < pp->emit_cdf[mid])
            hi = mid;
          else
            lo = mid + 1;
        }
        const float4* tp = scp->emit.tris + 3 * (size_t)lo;
        const float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
        const F3 v0 = xyz(p0), v1 = xyz(p1), v2 = xyz(p2);
        const float su = sqrtf(u1);
        const F3 yl = v0 * (1.0f - su) + v1 * (su * (1.0f - u2)) + v2 * (su * u2);
        const F3 wv = yl - O;
        const float r2 = dot(wv, wv);
        const F3 w = wv * (1.0f / sqrtf(r2));
        if (r2 > 0.0f && dot(N, w) > 0.0f && ray_valid(w)) {
          const F3 e1 = v1 - v0, e2 = v2 - v0;
          const F3 nt = mk(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
          const float cy = fabsf(dot(nt, w)) / sqrtf(dot(nt, nt));
          const float q = Bsdf::pdf(Din, N, w, 1.0f);  // the bounce's density at roughness 1
          lw = q * cy * pp->emit_area / r2;
          ls_k = lo;
          D = w;
          qstage = 4;  // the emitter query below: is lo the closest emitter?
        } else {
          D = Bsdf::sample(Din, N, seed, rough);  // p^ = 0
          cond = true;
          rejects = 0;
          qstage = 0;
        }
      } else {
        D = Bsdf::sample(Din, N, seed, rough);
        qstage = 0;
        if (b == pp->bounces) {
          finish_sample();
          alive = begin_sample();
          if (!alive) continue;
        }
      }
    }
    if (kEmit && qstage == 4) {
      // kMode 3's light sample: is the sampled triangle the closest emitter
      // along the line?  Then its any-hit test (qstage 3), else its bounce.
      float t = 0.0f;
      const int32_t k = closest<kCount, kPathBlock>(scp->emit, make_ray(O, D), kTmin, kTmax, 0xFFFFFFFFu,
                                                    false, t, my_stk, tc);
      if (k == ls_k) {
        const float4 e0 = scp->emit.tris[3 * (size_t)k], e1 = scp->emit.tris[3 * (size_t)k + 1];
        tE = t;
        primE = __float_as_uint(e0.w);
        matE = (int32_t)__float_as_uint(e1.w);
        qstage = 3;
      } else {  // p^ = 0
        D = Bsdf::sample(Din, N, seed, rough);
        cond = true;
        rejects = 0;
        qstage = 0;
      }
    }

    // -- one full traversal per iteration: the closest hit of a radiance ray
    // (qstage 0) or the any-hit query of an occlusion ray (qstage 2).
    const bool valid = ray_valid(D);
    int32_t k = -1;
    float t = 0.0f;
    if (valid && scp->all.T > 0) {
      const RayT r = make_ray(O, D);
      const bool any = qstage >= 2;
      if constexpr (kWide && !kCount)
        k = closest_wide<kPathBlock>(scp->all, r, kTmin, any ? tE : kTmax, any ? primE : 0xFFFFFFFFu,
                                     any, t, my_stk);
      else
        k = closest<kCount, kPathBlock>(scp->all, r, kTmin, any ? tE : kTmax, any ? primE : 0xFFFFFFFFu,
                                        any, t, my_stk, tc);
    }

    if (qstage == 0) {
      // radiance: miss (189-194) / closest hit (211-255)
      if (kCount) {
        c_rad++;
        if (!valid) c_inv++;
      }
      float4 t0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), t1 = t0;
      int32_t prim = valid ? -1 : -2, mat = -1;
      if (k >= 0) {
        t0 = scp->all.tris[3 * (size_t)k];
        t1 = scp->all.tris[3 * (size_t)k + 1];
        prim = (int32_t)__float_as_uint(t0.w);
        mat = (int32_t)__float_as_uint(t1.w);
      }
      if (kRecord) {
        hash = (hash ^ (uint32_t)prim) * 16777619u;
        if (pp->rad_hits && (int32_t)b < pp->rb) pp->rad_hits[(rec_base + s) * pp->rb + b] = prim;
      }
      ++b;
      bool finish = false;
      if (k < 0) {
        color = color + mk(0.0f, 0.0f, 0.0f);
        finish = true;
        if (kEmit) cond = false;
      } else {
        const DevMat m = scp->mats[mat];
        if (kEmit && cond && m.emission_emit.w != 0.0f) {
          // a conditioned bounce hit an emitter: draw it again (not a bounce)
          --b;
          if (++rejects == kMaxRedraws) {
            finish = true;
          } else {
            D = Bsdf::sample(Din, N, seed, rough);
            continue;
          }
        } else if (m.emission_emit.w != 0.0f) {
          color = color + xyz(m.emission_emit) * att;
          finish = true;
        } else {
          if (kEmit) cond = false;
          if (kCount) c_rhit++;
          O = O + t * D;
          Din = D;
          {
            const float4 t2 = scp->all.tris[3 * (size_t)k + 2];
            const float4* nq = scp->nrm + 3 * (size_t)k;
            N = bary_normal(O, xyz(nq[0]), xyz(nq[1]), xyz(nq[2]), xyz(t0), xyz(t1), xyz(t2));
          }
          bool cached = false;
          if (kQuery && b == pp->cache_at) {
            // the hit itself at the sample's slot; k_compact_queries turns
            // it into the features
            float4* r = pp->q_rec + 4 * (rec_base + s);
            r[0] = make_float4(O.x, O.y, O.z, __uint_as_float((uint32_t)mat));
            r[1] = make_float4(Din.x, Din.y, Din.z, 0.0f);
            r[2] = make_float4(N.x, N.y, N.z, 0.0f);
            r[3] = make_float4(att.x, att.y, att.z, 0.0f);
            pp->q_has[rec_base + s] = 1;
            cached = true;
            finish = true;
          }
          if (kMode == 2 && t_used < pp->t_slots && (pp->t_depth == 0 || b == pp->t_depth)) {
            float* r = pp->t_rec + 18 * ((size_t)item * pp->t_slots + t_used);
            cache_features(pp->bmin, pp->binv, O, Din, N, m, r);
            r[12] = att.x;
            r[13] = att.y;
            r[14] = att.z;
            r[15] = color.x;
            r[16] = color.y;
            r[17] = color.z;
            ++t_used;
          }
          if (cached) {
          } else if (m.diffuse_alpha.w < 1.0f) {
            D = Dielectric::scatter(D, N, m.rough_n.y, xyz(m.diffuse_alpha), att, seed);
            if (b == pp->bounces) finish = true;
          } else {
            att = att * xyz(m.diffuse_alpha);
            rough = m.rough_n.x;
            qstage = 1;
            tries = 0;
            if (kEmit) {  // the NEE loops of this vertex start from this flag and light
              ent = (*light_mat << 1) | (hitf ? 1 : 0);
              nl = 0;
            }
            D = hemisphere_raw(seed);
          }
        }
      }
      if (finish) {
        finish_sample();
        alive = begin_sample();
      }
    } else if (kEmit && qstage == 3) {
      // the light sample of a prefix bounce: visible unless blocked before (tE, primE)
      const float ph = k >= 0 ? 0.0f : lw;
      color = color + (xyz(scp->mats[matE].emission_emit) * ph) * att;
      att = att * (1.0f - ph);
      D = Bsdf::sample(Din, N, seed, rough);
      cond = true;
      rejects = 0;
      qstage = 0;
    } else {
      // occlusion, second query: blocked before (tE, primE)?
      const bool blocked = k >= 0;
      int32_t code;
      if (matE >= 0) {
        if (!blocked) {
          *light_mat = matE;
          hitf = true;
          code = (int32_t)primE;
        } else {
          code = hitf ? -4 : -5;
        }
      } else if (blocked) {
        code = -4;
      } else {
        code = -1;
        hitf = false;
      }
      alive = nee_after(code);
    }
  }

  if (kCount) {
    atomicAdd(&pp->counters[0], (unsigned long long)c_rad);
    atomicAdd(&pp->counters[1], (unsigned long long)c_occ);
    atomicAdd(&pp->counters[2], (unsigned long long)tc.visits);
    atomicAdd(&pp->counters[3], (unsigned long long)tc.tests);
    atomicAdd(&pp->counters[4], (unsigned long long)c_rhit);
    atomicAdd(&pp->counters[5], (unsigned long long)c_smp);
    atomicAdd(&pp->counters[6], (unsigned long long)c_inv);
    atomicAdd(&pp->counters[7], (unsigned long long)c_px);
  }
}

// One pixel per work-item: the listed pixels, or the full image in
// kTileW x kTileH tiles.
struct OnePixel {
  SceneView sc;
  PathArgs p;
  bool given;
  __device__ const SceneView* view() const { return &sc; }
  __device__ const PathArgs* args() const { return &p; }
  __device__ bool next(Job& j) {
    if (given) return false;
    given = true;
    if (p.pixels) {
      j.item = blockIdx.x * kPathBlock + threadIdx.x;
      if (j.item >= p.n_items) return false;
      const uint32_t pix = (uint32_t)p.pixels[j.item];
      j.x = pix % p.width;
      j.y = pix / p.width;
      return true;
    }
    const uint32_t tx = blockIdx.x % p.tiles_x, ty = blockIdx.x / p.tiles_x;
    j.x = tx * kTileW + threadIdx.x % kTileW;
    j.y = ty * kTileH + threadIdx.x / kTileW;
    if (j.x >= p.width || j.y >= p.height) return false;
    j.item = (int32_t)(j.y * p.width + j.x);
    return true;
  }
};

template <bool kRecord, bool kCount, bool kF64 = false, bool kWide = false>
__global__ __launch_bounds__(kPathBlock, (kRecord || kCount) ? 4 : (kWide ? LISA_WIDE_WAVES : LISA_PATH_MIN_BLOCKS)) void k_path(SceneView sc, PathArgs p) {
  __shared__ uint32_t stk[kLaneLds * kPathBlock];
  OnePixel src{sc, p, false};
  path_lane<kRecord, kCount, OnePixel, 0, kF64, kWide>(src, stk + threadIdx.x);
}

// Path loop in a cache mode (1: cached render of the full image in tiles,
// 2: training paths on a pixel list).
template <int kMode, bool kWide = false>
__global__ __launch_bounds__(kPathBlock, kMode == 3 ? LISA_EMIT_WAVES : LISA_MODE_WAVES) void k_path_mode(
    SceneView sc, PathArgs p) {
  __shared__ uint32_t stk[kLaneLds * kPathBlock];
  OnePixel src{sc, p, false};
  path_lane<false, false, OnePixel, kMode, false, kWide>(src, stk + threadIdx.x);
}

// One scene's pixels from a device counter, for persistent launches: the
// listed pixels (p.pixels) or the full image in 8x8-tile order.  A lane takes
// `chunk` work indices when it runs out; the lanes of a wave that ask
// together share one atomic.  Lanes whose paths end early pull more pixels
// instead of idling behind the wave's longest path.
struct PixelQueue {
  SceneView sc;
  PathArgs p;
  uint32_t* counter;
  uint32_t total, chunk, cur, end;
  __device__ const SceneView* view() const { return &sc; }
  __device__ const PathArgs* args() const { return &p; }
  __device__ bool next(Job& j) {
    for (;;) {
      if (cur >= end) {
        const uint64_t req = __ballot(1);
        const uint32_t lane = __lane_id();
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)req) - 1u;
        const uint32_t rank = (uint32_t)__popcll(req & ((1ull << lane) - 1ull));
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(req) * chunk);
        base = __shfl(base, (int)leader);
        const uint32_t mine = base + rank * chunk;
        if (mine >= total) return false;
        cur = mine;
        end = min(mine + chunk, total);
      }
      const uint32_t idx = cur++;
      if (p.pixels) {
        const uint32_t pix = (uint32_t)p.pixels[idx];
        j.x = pix % p.width;
        j.y = pix / p.width;
        j.item = (int32_t)idx;
        return true;
      }
      const uint32_t tile = idx >> 6, w = idx & 63u;
      const uint32_t x = (tile % p.tiles_x) * 8 + (w & 7u), y = (tile / p.tiles_x) * 8 + (w >> 3);
      if (x >= p.width || y >= p.height) continue;  // outside a partial edge tile
      j.x = x;
      j.y = y;
      j.item = (int32_t)(y * p.width + x);
      return true;
    }
  }
};

// Persistent form of the path loop (kMode as in k_path_mode, 0 = plain):
// `total` work indices (pixels of the list, or 64 per 8x8 tile) handed out
// by `counter`, which must be zero at launch.
template <int kMode, bool kWide = false>
__global__ __launch_bounds__(kPathBlock, kMode == 3 ? LISA_MODE_WAVES : 1) void k_path_pq(
    SceneView sc, PathArgs p, uint32_t* counter, uint32_t total, uint32_t chunk) {
  __shared__ uint32_t stk[kLaneLds * kPathBlock];
  PixelQueue src{sc, p, counter, total, chunk, 0u, 0u};
  path_lane<false, false, PixelQueue, kMode, false, kWide>(src, stk + threadIdx.x);
}

// frame[item] += throughput * max(network output, 0) / spp for every query.
// The queries come in slot order, so those of one pixel are contiguous and in
// sample order: the first of each run adds them all, in that order (no
// atomics; the same sums on every run).
__global__ __launch_bounds__(256) void k_cache_add(const float* __restrict__ pred,
                                                   const float* __restrict__ att,
                                                   const int32_t* __restrict__ item, uint32_t n,
                                                   float inv_spp, float4* frame) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int32_t it = item[i];
  if (i > 0 && item[i - 1] == it) return;
  float4 f = frame[it];
  for (uint32_t j = i; j < n && item[j] == it; ++j) {
    f.x = __fadd_rn(f.x, __fmul_rn(__fmul_rn(att[3 * j], fmaxf(pred[3 * j], 0.0f)), inv_spp));
    f.y = __fadd_rn(f.y, __fmul_rn(__fmul_rn(att[3 * j + 1], fmaxf(pred[3 * j + 1], 0.0f)), inv_spp));
    f.z = __fadd_rn(f.z, __fmul_rn(__fmul_rn(att[3 * j + 2], fmaxf(pred[3 * j + 2], 0.0f)), inv_spp));
  }
  frame[it] = f;
}

// Running mean over frames: accum = frame (first) or lerp(accum, frame, a).
// out[i] = a[px[i]] for n listed pixels.
__global__ __launch_bounds__(256) void k_pick(const float4* __restrict__ a, const int32_t* __restrict__ px,
                                              int32_t n, float4* __restrict__ out) {
  const int32_t i = (int32_t)(blockIdx.x * 256 + threadIdx.x);
  if (i < n) out[i] = a[px[i]];
}

// accum = FP64 sums / sample count (the accum buffer of loaded FP64 sums).
__global__ __launch_bounds__(256) void k_mean64(const double4* __restrict__ sum64, int64_t n, float4* accum) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double4 v = sum64[i];
  const double inv = 1.0 / v.w;
  accum[i] = make_float4((float)(v.x * inv), (float)(v.y * inv), (float)(v.z * inv), 1.0f);
}

// The training paths' pixels: item i of the frame's splitmix64 stream from
// sm0 (the state before the first step), modulo the pixel count -- the same
// values as the sequential generator, computed independently per item.
__global__ __launch_bounds__(256) void k_train_pixels(int32_t* __restrict__ px, uint32_t n, uint64_t sm0,
                                                      uint32_t npx) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  uint64_t z = sm0 + (uint64_t)(i + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  px[i] = (int32_t)(z % npx);
}

__global__ __launch_bounds__(256) void k_blend(float4* accum, const float4* __restrict__ frame,
                                               int64_t n, float a, int first) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 f = frame[i];
  if (first) {
    accum[i] = make_float4(f.x, f.y, f.z, 1.0f);
  } else {
    const float4 o = accum[i];
    const F3 r = lerp(mk(o.x, o.y, o.z), mk(f.x, f.y, f.z), a);
    accum[i] = make_float4(r.x, r.y, r.z, 1.0f);
  }
}

// ---------------------------------------------------------------------------
// Batched wavefront form: one persistent launch renders many scenes.  Work is
// the concatenation of every scene's pixels in 8x8-tile order; each lane
// takes `chunk` work indices at a time from a device counter, renders those
// pixels completely (the same per-pixel computation as k_path, so the images
// are bit-identical) and takes more until the queue is empty.  Lanes whose
// paths end early keep pulling pixels instead of idling behind the wave's
// longest path, and small and large scenes share one launch.
// ---------------------------------------------------------------------------
struct BatchScene {
  SceneView sc;
  PathArgs p;      // the scene's launch (tiles_x counts 8x8 tiles)
  uint32_t first;  // first work index of the scene
};

struct QueueSrc {
  const BatchScene* __restrict__ scenes;
  int32_t n_scenes;
  uint32_t* counter;
  uint32_t total;
  uint32_t chunk;
  uint32_t cur, end;
  const BatchScene* bs;  // the scene of the current pixel (device table entry)
  __device__ const SceneView* view() const { return &bs->sc; }
  __device__ const PathArgs* args() const { return &bs->p; }
  __device__ bool next(Job& j) {
    for (;;) {
      if (cur >= end) {
        const uint32_t base = atomicAdd(counter, chunk);
        if (base >= total) return false;
        cur = base;
        end = min(base + chunk, total);
      }
      const uint32_t idx = cur++;
      int32_t lo = 0, hi = n_scenes - 1;
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (scenes[mid].first <= idx)
          lo = mid;
        else
          hi = mid - 1;
      }
      bs = scenes + lo;
      const uint32_t local = idx - bs->first;
      const uint32_t tile = local >> 6, w = local & 63u;
      const uint32_t tx = bs->p.tiles_x, width = bs->p.width;
      const uint32_t x = (tile % tx) * 8 + (w & 7u), y = (tile / tx) * 8 + (w >> 3);
      if (x >= width || y >= bs->p.height) continue;  // outside a partial edge tile
      j.x = x;
      j.y = y;
      j.item = (int32_t)(y * width + x);
      return true;
    }
  }
};

__global__ __launch_bounds__(kPathBlock) void k_path_queue(const BatchScene* scenes, int32_t n,
                                                           uint32_t* counter, uint32_t total,
                                                           uint32_t chunk) {
  __shared__ uint32_t stk[kLaneLds * kPathBlock];
  QueueSrc src;
  src.scenes = scenes;
  src.n_scenes = n;
  src.counter = counter;
  src.total = total;
  src.chunk = chunk;
  src.cur = src.end = 0;
  src.bs = scenes;
  path_lane<false, false>(src, stk + threadIdx.x);
}

// The wave-tile form of the batched launch: each wave takes one 8x8 tile of
// the concatenated work list at a time and renders it like k_path.  The
// tile's scene is the same for the whole wave, so its pointers and launch
// constants are scalar; lanes whose pixel is done wait for the tile.
struct TileSrc {
  const BatchScene* bs;  // wave-uniform: its fields are read through scalar loads
  uint32_t tile;
  bool given;
  __device__ const SceneView* view() const { return &bs->sc; }
  __device__ const PathArgs* args() const { return &bs->p; }
  __device__ bool next(Job& j) {
    if (given) return false;
    given = true;
    const uint32_t w = threadIdx.x & 63u;
    const uint32_t tx = bs->p.tiles_x;
    j.x = (tile % tx) * 8 + (w & 7u);
    j.y = (tile / tx) * 8 + (w >> 3);
    if (j.x >= bs->p.width || j.y >= bs->p.height) return false;
    j.item = (int32_t)(j.y * bs->p.width + j.x);
    return true;
  }
};

__global__ __launch_bounds__(kPathBlock, LISA_TILE_WAVES) void k_path_tiles(const BatchScene* scenes, int32_t n,
                                                           uint32_t* counter, uint32_t total) {
  __shared__ uint32_t stk[kLaneLds * kPathBlock];
  const uint32_t n_tiles = total / 64u;
  for (;;) {
    uint32_t t = 0;
    if (threadIdx.x == 0) t = atomicAdd(counter, 1u);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
    if (t >= n_tiles) break;
    int32_t lo = 0, hi = n - 1;  // the scene whose work range holds tile t
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (scenes[mid].first / 64u <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    TileSrc src{scenes + lo, t - scenes[lo].first / 64u, false};
    path_lane<false, false, TileSrc>(src, stk + threadIdx.x);
  }
}

// Queue reset inside the captured graph (a kernel, not a memset node).
__global__ void k_queue_reset(uint32_t* counter) { *counter = 0u; }

// ---------------------------------------------------------------------------
// Ray-level wavefront form (LISA_FLAG_WAVEFRONT, batch form 2).  The per-lane
// state machine of path_lane is split at its one full traversal: every path
// in flight keeps its state in a slot of HBM arrays (SoA, 16-byte records),
// and each round of the host loop runs
//   k_wf_trace: persistent; each wave holds a queue of traversal requests in
//     LDS, filled 64 at a time from the round's request list, and every lane
//     whose traversal ends takes the next request from it at once (dynamic
//     refill), so lanes do not idle behind the wave's longest traversal;
//   k_wf_shade: the closest-hit / miss / NEE logic of the request's result,
//     then the path's cheap work up to its next full traversal (the NEE
//     tries resolved by the emitter tree, the next sample or the next pixel
//     of the work list); the paths that need a traversal are packed by a
//     wave ballot and prefix sum into LDS and appended to the next round's
//     list with one atomic per wave.
// Each pixel's samples still run in order through its own RNG stream, so
// every image is bit-identical to k_path's.
// ---------------------------------------------------------------------------
constexpr int kWfBlock = 64;   // one wave per workgroup
constexpr int kWfStack = 20;   // LDS stack entries per lane of k_wf_trace (8 waves/SIMD: 5 KB each)
constexpr uint32_t kWfAny = 0x80000000u;  // request bit: any-hit query (else closest hit)
constexpr uint32_t kWfIdle = 0xFFFFFFFFu;

// Slot state: seven float4 records (xyz + one 32-bit field in .w) and one
// uint4 of integer state; results of the round's traversal in `res`.
struct WfState {
  float4* O;      // ray origin / hit point | tE (the occlusion query's limit)
  float4* D;      // ray direction (raw NEE draws while trying) | primE
  float4* col;    // radiance of the sample | roughness of the vertex
  float4* att;    // throughput | matE
  float4* N;      // shading normal | light material (RayState::material)
  float4* Din;    // direction into the vertex | LCG state
  float4* acc;    // the pixel's sum of samples | samples done
  uint4* st;      // x: b | tries << 8 | qstage << 16 | hit flag << 20; y: work index (kWfIdle:
                  // none); z: scene index; w: 0
  uint2* res;     // sorted triangle index (or -1) and t of the round's traversal
  uint32_t* q[2]; // request lists (slot | kWfAny)
  uint32_t* cnt;  // [0], [1]: list lengths; [3]: next work index
  uint32_t slots;
};

__device__ __forceinline__ const BatchScene* wf_scene_of(const BatchScene* scenes, int32_t n, uint32_t idx) {
  int32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (scenes[mid].first <= idx)
      lo = mid;
    else
      hi = mid - 1;
  }
  return scenes + lo;
}

// One node visit of a resumable closest / any-hit walk (the loop body of
// closest(): the same order of visits and tests, so the same answer and
// counts).  Returns false when the walk has ended.
struct WfWalk {
  float best_t, cull_t;
  uint32_t best_prim;
  int32_t best_k;
  uint32_t node;
  int32_t sp;
  uint32_t any;
};

template <int kStride>
__device__ __forceinline__ bool wf_step(const BvhView& bv, const RayT& r, WfWalk& w, uint32_t* stk,
                                        uint32_t* spill) {
  const float4* np = bv.nodes + 4 * (size_t)w.node;
  const float4 a = np[0], b = np[1], c = np[2], d = np[3];
  const uint32_t r0 = __float_as_uint(d.x), r1 = __float_as_uint(d.y);
  float tn0 = 0.0f, tn1 = 0.0f;
  bool h0 = (r0 != kEmpty) & slab(a.x, a.y, a.z, a.w, b.x, b.y, r, kTmin, w.cull_t, tn0);
  bool h1 = (r1 != kEmpty) & slab(b.z, b.w, c.x, c.y, c.z, c.w, r, kTmin, w.cull_t, tn1);
#pragma unroll
  for (int ch = 0; ch < 2; ++ch) {
    const bool hit = ch == 0 ? h0 : h1;
    const uint32_t ref = ch == 0 ? r0 : r1;
    if (hit && (ref & kLeaf)) {
      const uint32_t k = ref & ~kLeaf;
      const float4* tp = bv.tris + 3 * (size_t)k;
      const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
      const uint32_t prim = __float_as_uint(t0.w);
      float t;
      if (watertight(r, xyz(t0), xyz(t1), xyz(t2), t) && t > kTmin &&
          (t < w.best_t || (t == w.best_t && prim < w.best_prim))) {
        w.best_t = t;
        w.cull_t = t * kWiden;
        w.best_prim = prim;
        w.best_k = (int32_t)k;
      }
      if (ch == 0)
        h0 = false;
      else
        h1 = false;
    }
  }
  if (w.any && w.best_k >= 0) return false;
  if (h0 && h1) {
    uint32_t nr = r0, fr = r1;
    if (tn1 < tn0) {
      nr = r1;
      fr = r0;
    }
    if (w.sp == kStackMax) return false;  // unreachable for a valid tree (depth < 64)
    if (w.sp < kWfStack) {
      stk[w.sp * kStride] = fr;
    } else {
      asm volatile("" ::: "memory");
      spill[w.sp - kWfStack] = fr;
    }
    ++w.sp;
    w.node = nr;
  } else if (h0) {
    w.node = r0;
  } else if (h1) {
    w.node = r1;
  } else {
    if (w.sp == 0) return false;
    --w.sp;
    if (w.sp < kWfStack) {
      w.node = stk[w.sp * kStride];
    } else {
      asm volatile("" ::: "memory");
      w.node = spill[w.sp - kWfStack];
    }
  }
  return true;
}

// k_wf_trace: the round's requests q[cur][0 .. cnt[cur]), split evenly over
// the grid's waves (no shared counter: thousands of waves taking requests
// from one atomic serialise on it).  A wave loads its share 64 requests at a
// time into its LDS queue; a lane with no walk takes the queue's next request
// (ballot + rank) whenever at least kRefill lanes are free or none is
// walking.
constexpr int kWfRefill = 16;
__global__ __launch_bounds__(kWfBlock, 8) void k_wf_trace(const BatchScene* __restrict__ scenes, int32_t n_scenes,
                                                          WfState ws, uint32_t cur) {
  __shared__ uint32_t stk[kWfStack * kWfBlock];
  __shared__ uint32_t lq[kWfBlock];
  uint32_t spill[kStackMax - kWfStack];
  uint32_t* my_stk = stk + threadIdx.x;
  const uint32_t lane = threadIdx.x;
  if (blockIdx.x == 0 && lane == 0) ws.cnt[cur ^ 1u] = 0u;  // the next list (k_wf_shade appends to it)
  const uint32_t n = ws.cnt[cur];
  const uint32_t* __restrict__ q = ws.q[cur];
  // wave-uniform: this wave's share [next, last) and the LDS queue's unread range
  uint32_t next = (uint32_t)((uint64_t)n * blockIdx.x / gridDim.x);
  const uint32_t last = (uint32_t)((uint64_t)n * (blockIdx.x + 1) / gridDim.x);
  uint32_t q_pos = 0, q_end = 0;
  bool exhausted = false;
  bool walking = false;
  uint32_t slot = 0;
  RayT r;
  WfWalk w;
  BvhView bv;
  for (;;) {
    const uint64_t free_lanes = __ballot(!walking);
    const uint32_t n_free = (uint32_t)__popcll(free_lanes);
    if (!exhausted && n_free > 0 && (n_free >= kWfRefill || n_free == 64u)) {
      if (q_pos >= q_end) {  // the LDS queue is empty: the share's next 64 requests
        if (next >= last) {
          exhausted = true;
        } else {
          q_end = min((uint32_t)kWfBlock, last - next);
          q_pos = 0;
          if (lane < q_end) lq[lane] = q[next + lane];
          next += q_end;
          __syncthreads();
        }
      }
      if (!exhausted) {
        const uint32_t rank = (uint32_t)__popcll(free_lanes & ((1ull << lane) - 1ull));
        const uint32_t avail = q_end - q_pos;
        if (!walking && rank < avail) {
          const uint32_t req = lq[q_pos + rank];
          slot = req & ~kWfAny;
          const float4 o4 = ws.O[slot], d4 = ws.D[slot];
          const uint4 st = ws.st[slot];
          const BatchScene* bs = scenes + st.z;
          bv = bs->sc.all;
          const F3 Dd = xyz(d4);
          w.any = req & kWfAny;
          w.best_t = w.any ? o4.w : kTmax;
          w.cull_t = w.best_t * kWiden;
          w.best_prim = w.any ? __float_as_uint(d4.w) : 0xFFFFFFFFu;
          w.best_k = -1;
          w.node = 0;
          w.sp = 0;
          if (ray_valid(Dd) && bv.T > 0) {
            r = make_ray(xyz(o4), Dd);
            walking = true;
          } else {  // a zero or non-finite direction misses without traversal
            ws.res[slot] = make_uint2(0xFFFFFFFFu, 0u);
          }
        }
        q_pos += min(n_free, avail);
      }
    }
    if (!__any(walking)) {
      if (exhausted) break;
      continue;
    }
    if (walking && !wf_step<kWfBlock>(bv, r, w, my_stk, spill)) {
      ws.res[slot] = make_uint2((uint32_t)w.best_k, __float_as_uint(w.best_t));
      walking = false;
    }
  }
}

// k_wf_shade (fresh = false): for each request of list `cur`, the result of
// its traversal, then the path up to its next traversal; k_wf_start
// (fresh = true): slots [0, slots) take their first pixel.  The steps and
// their order are path_lane's (plain mode): the same RNG draws, the same
// FP32 operations.
template <class Bsdf = LISA_BSDF>
__device__ __forceinline__ bool wf_advance(const BatchScene* __restrict__ scenes, int32_t n_scenes,
                                           uint32_t total, const WfState& ws, uint32_t slot, bool fresh,
                                           uint32_t* my_stk, uint32_t& req) {
  const float4 o4 = ws.O[slot], d4 = ws.D[slot], c4 = ws.col[slot], a4 = ws.att[slot];
  const float4 n4 = ws.N[slot], i4 = ws.Din[slot], ac4 = ws.acc[slot];
  const uint4 st = ws.st[slot];
  F3 O = xyz(o4), D = xyz(d4), color = xyz(c4), att = xyz(a4), N = xyz(n4), Din = xyz(i4);
  F3 acc = xyz(ac4);
  float tE = o4.w, rough = c4.w;
  uint32_t primE = __float_as_uint(d4.w), seed = __float_as_uint(i4.w), s = __float_as_uint(ac4.w);
  int32_t matE = (int32_t)__float_as_uint(a4.w), light_mat = (int32_t)__float_as_uint(n4.w);
  uint32_t b = st.x & 0xFFu, tries = (st.x >> 8) & 0xFFu, qstage = (st.x >> 16) & 0xFu;
  bool hitf = ((st.x >> 20) & 1u) != 0;
  uint32_t widx = st.y, scene = fresh ? 0u : st.z;  // a fresh slot's state is not yet written
  const BatchScene* bs = scenes + scene;
  const SceneView* scp = &bs->sc;
  const PathArgs* pp = &bs->p;
  TravCount tc;

  auto finish_sample = [&]() {
    acc = acc + color;
    ++s;
  };
  auto begin_sample = [&](uint32_t x, uint32_t y) -> bool {
    while (s < pp->spp) {
      const float jx = rng(seed);
      const float jy = rng(seed);
      const float dx = (2.0f * (float)x + jx) / (float)pp->width - 1.0f;
      const float dy = (2.0f * (float)y + jy) / (float)pp->height - 1.0f;
      D = normalize(dx * pp->U + dy * pp->V + pp->W);
      O = pp->eye;
      color = mk(0.0f, 0.0f, 0.0f);
      att = mk(1.0f, 1.0f, 1.0f);
      hitf = false;
      qstage = 0;
      b = 0;
      if (pp->bounces > 0) return true;
      finish_sample();
    }
    return false;
  };
  // the pixel of work index widx (x, y): false outside a partial edge tile
  // (bs's own launch: during a fetch bs is already the new pixel's scene)
  auto pixel_of = [&](uint32_t& x, uint32_t& y) -> bool {
    const uint32_t local = widx - bs->first;
    const uint32_t tile = local >> 6, wl = local & 63u;
    const uint32_t tx = bs->p.tiles_x;
    x = (tile % tx) * 8 + (wl & 7u);
    y = (tile / tx) * 8 + (wl >> 3);
    return x < bs->p.width && y < bs->p.height;
  };
  auto sample_xy = [&](uint32_t& x, uint32_t& y) { (void)pixel_of(x, y); };
  auto end_nee = [&](F3 nee) -> bool {
    color = color + nee * att;
    D = Bsdf::sample(Din, N, seed, rough);
    qstage = 0;
    if (b == pp->bounces) {
      finish_sample();
      uint32_t x, y;
      sample_xy(x, y);
      return begin_sample(x, y);
    }
    return true;
  };
  auto nee_after = [&]() -> bool {
    if (hitf) return end_nee(xyz(scp->mats[light_mat].emission_emit) * Bsdf::eval(N, D));
    if (++tries == (uint32_t)kNeeTries) return end_nee(mk(0.0f, 0.0f, 0.0f));
    D = hemisphere_raw(seed);
    qstage = 1;
    return true;
  };

  bool alive = true;
  if (fresh) {
    alive = false;
    widx = kWfIdle;
  } else {
    const uint2 rk = ws.res[slot];
    const int32_t k = (int32_t)rk.x;
    const float t = __uint_as_float(rk.y);
    if (qstage == 0) {  // radiance: miss (shader.cu:189-194) / closest hit (211-255)
      ++b;
      bool finish = false;
      if (k < 0) {
        color = color + mk(0.0f, 0.0f, 0.0f);
        finish = true;
      } else {
        const float4 t0 = scp->all.tris[3 * (size_t)k], t1 = scp->all.tris[3 * (size_t)k + 1];
        const DevMat m = scp->mats[__float_as_uint(t1.w)];
        if (m.emission_emit.w != 0.0f) {
          color = color + xyz(m.emission_emit) * att;
          finish = true;
        } else {
          O = O + t * D;
          Din = D;
          {
            const float4 t2 = scp->all.tris[3 * (size_t)k + 2];
            const float4* nq = scp->nrm + 3 * (size_t)k;
            N = bary_normal(O, xyz(nq[0]), xyz(nq[1]), xyz(nq[2]), xyz(t0), xyz(t1), xyz(t2));
          }
          if (m.diffuse_alpha.w < 1.0f) {
            D = Dielectric::scatter(D, N, m.rough_n.y, xyz(m.diffuse_alpha), att, seed);
            if (b == pp->bounces) finish = true;
          } else {
            att = att * xyz(m.diffuse_alpha);
            rough = m.rough_n.x;
            qstage = 1;
            tries = 0;
            D = hemisphere_raw(seed);
          }
        }
      }
      if (finish) {
        finish_sample();
        uint32_t x, y;
        sample_xy(x, y);
        alive = begin_sample(x, y);
      }
    } else {  // occlusion, second query: blocked before (tE, primE)?
      const bool blocked = k >= 0;
      if (matE >= 0) {
        if (!blocked) {
          light_mat = matE;
          hitf = true;
        }
      } else if (!blocked) {
        hitf = false;
      }
      alive = nee_after();
    }
  }
  for (;;) {
    if (!alive) {
      if (widx != kWfIdle) {  // shader.cu:156-166: the pixel's mean
        uint32_t x, y;
        sample_xy(x, y);
        const uint32_t item = y * pp->width + x;
        F3 out = acc * (1.0f / (float)pp->spp);
        if (pp->subframe > 0 && pp->accum_prev) {
          const float a = 1.0f / (float)(pp->subframe + 1);
          out = lerp(xyz(pp->accum_prev[item]), out, a);
        }
        pp->accum[item] = make_float4(out.x, out.y, out.z, 1.0f);
      }
      // the next pixel of the work list (one atomic per wave)
      uint32_t x = 0, y = 0;
      bool got = false;
      for (;;) {
        widx = wave_reserve(&ws.cnt[3]);
        if (widx >= total) break;
        bs = wf_scene_of(scenes, n_scenes, widx);
        if (pixel_of(x, y)) {
          got = true;
          break;
        }
      }
      if (!got) {  // no work left: the slot idles
        ws.st[slot] = make_uint4(0u, kWfIdle, 0u, 0u);
        return false;
      }
      scene = (uint32_t)(bs - scenes);
      scp = &bs->sc;
      pp = &bs->p;
      seed = tea16(y * pp->width + x, pp->subframe);  // shader.cu:141
      acc = mk(0.0f, 0.0f, 0.0f);
      s = 0;
      alive = begin_sample(x, y);
      continue;
    }
    // NEE occlusion rays, first query: the closest emitter (emitter tree);
    // with the flag false a ray that hits no emitter is resolved here
    while (qstage == 1) {
      if (!hitf && line_culled(scp->ecull_c, scp->ecull_h, O, D)) {
        alive = nee_after();
        if (!alive) break;
        continue;
      }
      D = hemisphere_dir(D, N);
      const bool valid = ray_valid(D);
      int32_t k = -1;
      float t = 0.0f;
      if (valid && scp->emit.T > 0 && !emit_culled(scp->ecull_c, scp->ecull_h, O, D)) {
        const RayT r = make_ray(O, D);
        k = closest<false, kWfBlock>(scp->emit, r, kTmin, kTmax, 0xFFFFFFFFu, false, t, my_stk, tc);
      }
      if (!valid) {  // counts as a ray, misses without traversal
        hitf = false;
        alive = nee_after();
      } else if (k >= 0) {  // an emitter: is anything in front of it?
        const float4 e0 = scp->emit.tris[3 * (size_t)k], e1 = scp->emit.tris[3 * (size_t)k + 1];
        tE = t;
        primE = __float_as_uint(e0.w);
        matE = (int32_t)__float_as_uint(e1.w);
        qstage = 2;
      } else if (!hitf) {
        alive = nee_after();
      } else {  // stale hit flag: a miss clears it, any other hit keeps it
        tE = kTmax;
        primE = 0xFFFFFFFFu;
        matE = -1;
        qstage = 2;
      }
      if (!alive) break;
    }
    if (!alive) continue;
    break;
  }
  ws.O[slot] = make_float4(O.x, O.y, O.z, tE);
  ws.D[slot] = make_float4(D.x, D.y, D.z, __uint_as_float(primE));
  ws.col[slot] = make_float4(color.x, color.y, color.z, rough);
  ws.att[slot] = make_float4(att.x, att.y, att.z, __uint_as_float((uint32_t)matE));
  ws.N[slot] = make_float4(N.x, N.y, N.z, __uint_as_float((uint32_t)light_mat));
  ws.Din[slot] = make_float4(Din.x, Din.y, Din.z, __uint_as_float(seed));
  ws.acc[slot] = make_float4(acc.x, acc.y, acc.z, __uint_as_float(s));
  ws.st[slot] = make_uint4(b | (tries << 8) | (qstage << 16) | ((hitf ? 1u : 0u) << 20), widx, scene, 0u);
  req = slot | (qstage == 2 ? kWfAny : 0u);
  return true;
}

// Appends the requests of the lanes with `need` to list `nxt`: a wave ballot
// and prefix sum place them in an LDS row, one atomic reserves the wave's
// range, and the row is written out in one coalesced store.
__device__ __forceinline__ void wf_enqueue(const WfState& ws, uint32_t nxt, bool need, uint32_t req,
                                           uint32_t* row) {
  const uint64_t m = __ballot(need);
  if (m == 0) return;
  const uint32_t lane = __lane_id();
  const uint32_t cnt = (uint32_t)__popcll(m);
  if (need) row[__popcll(m & ((1ull << lane) - 1ull))] = req;
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(&ws.cnt[nxt], cnt);
  base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
  __syncthreads();
  if (lane < cnt) ws.q[nxt][base + lane] = row[lane];
  __syncthreads();
}

__global__ __launch_bounds__(kWfBlock, 4) void k_wf_shade(const BatchScene* __restrict__ scenes, int32_t n_scenes,
                                                          uint32_t total, WfState ws, uint32_t cur) {
  __shared__ uint32_t stk[(kStackLds + 1) * kWfBlock];
  uint32_t* row = stk + kStackLds * kWfBlock;
  const uint32_t n = ws.cnt[cur];
  for (uint32_t base = blockIdx.x * kWfBlock; base < n; base += gridDim.x * kWfBlock) {
    const uint32_t i = base + threadIdx.x;
    uint32_t req = 0;
    bool need = false;
    if (i < n) need = wf_advance(scenes, n_scenes, total, ws, ws.q[cur][i] & ~kWfAny, false, stk + threadIdx.x, req);
    wf_enqueue(ws, cur ^ 1u, need, req, row);
  }
}

__global__ __launch_bounds__(kWfBlock, 4) void k_wf_start(const BatchScene* __restrict__ scenes, int32_t n_scenes,
                                                          uint32_t total, WfState ws) {
  __shared__ uint32_t stk[(kStackLds + 1) * kWfBlock];
  uint32_t* row = stk + kStackLds * kWfBlock;
  const uint32_t i = blockIdx.x * kWfBlock + threadIdx.x;
  uint32_t req = 0;
  bool need = false;
  if (i < ws.slots) need = wf_advance(scenes, n_scenes, total, ws, i, true, stk + threadIdx.x, req);
  wf_enqueue(ws, 0u, need, req, row);
}

// Zeroes the list lengths and work counter before k_wf_start (a kernel, so
// the launch sequence can be captured into a graph without memset nodes).
__global__ void k_wf_reset(uint32_t* cnt) {
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// LBVH build (Karras, HPG 2012).
// ---------------------------------------------------------------------------
constexpr int kBuildBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kBuildBlock * kSortItems;  // 4096 keys per tile

// Triangle k of a build: global triangle gid[k] (k itself when gid is null).
__device__ __forceinline__ uint32_t global_id(const uint32_t* __restrict__ gid, uint32_t k) {
  return gid ? gid[k] : k;
}

__device__ __forceinline__ float centroid(const float* __restrict__ v, size_t g, int a) {
  const float* t = v + 9 * g;
  return (t[a] + t[3 + a] + t[6 + a]) / 3.0f;
}

// Per-block min/max of triangle centroids (6 floats per block).
__global__ __launch_bounds__(kBuildBlock) void k_centroid_partial(const float* __restrict__ v,
                                                                  const uint32_t* __restrict__ gid,
                                                                  int32_t T, float* partial) {
  __shared__ float s[6][kBuildBlock];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int32_t k = blockIdx.x * kBuildBlock + threadIdx.x; k < T; k += gridDim.x * kBuildBlock) {
    const uint32_t g = global_id(gid, (uint32_t)k);
    for (int a = 0; a < 3; ++a) {
      const float c = centroid(v, (size_t)g, a);
      lo[a] = fminf(lo[a], c);
      hi[a] = fmaxf(hi[a], c);
    }
  }
  for (int a = 0; a < 3; ++a) {
    s[a][threadIdx.x] = lo[a];
    s[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = kBuildBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + w]);
        s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(kBuildBlock) void k_centroid_final(const float* partial, int32_t nb,
                                                                float* bounds) {
  __shared__ float s[6][kBuildBlock];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int32_t i = threadIdx.x; i < nb; i += kBuildBlock)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], partial[6 * i + a]);
      hi[a] = fmaxf(hi[a], partial[6 * i + 3 + a]);
    }
  for (int a = 0; a < 3; ++a) {
    s[a][threadIdx.x] = lo[a];
    s[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = kBuildBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + w]);
        s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) bounds[threadIdx.x] = s[threadIdx.x][0];
}

// 10-bit spread by multiply-and-mask (every third bit).
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
__device__ __forceinline__ uint32_t quant10(float u) {
  float q = u * 1024.0f;
  q = fminf(fmaxf(q, 0.0f), 1023.0f);
  return (uint32_t)q;
}

__global__ __launch_bounds__(kBuildBlock) void k_morton(const float* __restrict__ v,
                                                        const uint32_t* __restrict__ gid, int32_t T,
                                                        const float* __restrict__ bounds,
                                                        uint32_t* keys, uint32_t* vals) {
  const int32_t k = blockIdx.x * kBuildBlock + threadIdx.x;
  if (k >= T) return;
  const uint32_t g = global_id(gid, (uint32_t)k);
  uint32_t q[3];
  for (int a = 0; a < 3; ++a) {
    const float lo = bounds[a], hi = bounds[3 + a];
    const float ext = hi - lo;
    const float c = centroid(v, (size_t)g, a);
    const float u = ext > 0.0f ? (c - lo) / ext : 0.5f;
    q[a] = quant10(u);
  }
  keys[k] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  vals[k] = (uint32_t)k;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exscan256(uint32_t v, uint32_t* s_tmp /*>=4*/,
                                                    uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_tmp[wave] = x;
  __syncthreads();
  uint32_t wbase = 0;
  for (int w = 0; w < wave; ++w) wbase += s_tmp[w];
  total = s_tmp[0] + s_tmp[1] + s_tmp[2] + s_tmp[3];
  __syncthreads();
  return wbase + x - v;
}

// Order-keeping compaction of per-slot counts (the query flag of each
// sample, the training rows of each sample): chunks of kScanItems slots, 16
// per lane of a 256-lane workgroup; k_scan_chunks sums each chunk,
// k_scan_top turns the sums into chunk offsets, and a compaction kernel
// scans its chunk again and writes each slot's entries at their offsets.
constexpr uint32_t kScanLane = 16;
constexpr uint32_t kScanItems = 256 * kScanLane;

// The 16 counts of this lane in chunk blockIdx.x (0 past n); i0 = the first's
// slot.  One 16-byte load inside the array (i0 is a multiple of 16).
__device__ __forceinline__ uint32_t chunk_counts(const uint8_t* __restrict__ cnt, size_t n,
                                                 uint32_t c[kScanLane], size_t& i0) {
  i0 = (size_t)blockIdx.x * kScanItems + kScanLane * threadIdx.x;
  if (i0 + kScanLane <= n) {
    const uint4 v = *reinterpret_cast<const uint4*>(cnt + i0);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) c[k] = (w[k >> 2] >> (8 * (k & 3))) & 255u;
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) c[k] = i0 + k < n ? cnt[i0 + k] : 0u;
  }
  uint32_t t = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) t += c[k];
  return t;
}

__global__ __launch_bounds__(256) void k_scan_chunks(const uint8_t* __restrict__ cnt, size_t n,
                                                     uint32_t* __restrict__ sums) {
  __shared__ uint32_t s_tmp[4];
  uint32_t c[kScanLane], total;
  size_t i0;
  (void)block_exscan256(chunk_counts(cnt, n, c, i0), s_tmp, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive prefix sums of the nb chunk sums, in place (one workgroup, 16
// consecutive sums per lane); *total = their sum.
__global__ __launch_bounds__(256) void k_scan_top(uint32_t* __restrict__ sums, uint32_t nb,
                                                  uint32_t* __restrict__ total) {
  __shared__ uint32_t s_tmp[4];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += kScanItems) {
    const uint32_t i = b0 + kScanLane * threadIdx.x;
    uint32_t v[kScanLane], t = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      v[k] = i + k < nb ? sums[i + k] : 0u;
      t += v[k];
    }
    uint32_t bt;
    uint32_t e = carry + block_exscan256(t, s_tmp, bt);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (i + k < nb) sums[i + k] = e;
      e += v[k];
    }
    carry += bt;
  }
  if (threadIdx.x == 0) *total = carry;
}

// The queries of a cached frame in slot order (pixel, then sample): each
// flagged slot's features (from the hit point, incoming direction, shading
// normal and material written by the path loop), throughput and pixel item
// (slot / spp), at its place among the flagged slots.  The chunk's flagged
// slots are listed in LDS first, so that every lane then works on one.
__global__ __launch_bounds__(256) void k_compact_queries(
    const uint8_t* __restrict__ has, const float4* __restrict__ rec, size_t n,
    const uint32_t* __restrict__ offs, uint32_t spp, F3 bmin, F3 binv, const DevMat* __restrict__ mats,
    float* __restrict__ feat, float* __restrict__ att, int32_t* __restrict__ item) {
  __shared__ uint32_t s_tmp[4];
  __shared__ uint16_t s_list[kScanItems];
  uint32_t c[kScanLane], total;
  size_t i0;
  uint32_t lp = block_exscan256(chunk_counts(has, n, c, i0), s_tmp, total);
#pragma unroll
  for (int k = 0; k < 16; ++k)
    if (c[k]) s_list[lp++] = (uint16_t)(kScanLane * threadIdx.x + k);
  __syncthreads();
  const uint32_t base = offs[blockIdx.x];
  const size_t chunk0 = (size_t)blockIdx.x * kScanItems;
  for (uint32_t j = threadIdx.x; j < total; j += 256) {
    const size_t slot = chunk0 + s_list[j];
    const uint32_t pos = base + j;
    const float4 P = rec[4 * slot], D = rec[4 * slot + 1], N = rec[4 * slot + 2], t = rec[4 * slot + 3];
    const DevMat m = mats[__float_as_uint(P.w)];
    cache_features(bmin, binv, mk(P.x, P.y, P.z), mk(D.x, D.y, D.z), mk(N.x, N.y, N.z), m,
                   feat + 12 * (size_t)pos);
    att[3 * (size_t)pos] = t.x;
    att[3 * (size_t)pos + 1] = t.y;
    att[3 * (size_t)pos + 2] = t.z;
    item[pos] = (int32_t)(slot / spp);
  }
}

// The training rows in slot order (pixel, sample, vertex): the rows of
// sample p (cnt[p] of them, at row slots p * slots + v) follow those of the
// samples before it; rows at or past cap are dropped.  The chunk's samples
// with rows are listed in LDS with their first row (<< 7 | count), so that
// every lane then copies the rows of one.
__global__ __launch_bounds__(256) void k_compact_rows(const uint8_t* __restrict__ cnt,
                                                      const float* __restrict__ row, size_t n,
                                                      uint32_t slots, const uint32_t* __restrict__ offs,
                                                      uint32_t cap, float* __restrict__ feat,
                                                      float* __restrict__ tgt) {
  __shared__ uint32_t s_tmp[4];
  __shared__ uint16_t s_list[kScanItems];
  __shared__ uint32_t s_row[kScanItems];
  uint32_t c[kScanLane], rows_total, listed;
  size_t i0;
  uint32_t r0 = block_exscan256(chunk_counts(cnt, n, c, i0), s_tmp, rows_total);
  uint32_t nz = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) nz += c[k] != 0u;
  uint32_t lp = block_exscan256(nz, s_tmp, listed);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (c[k]) {
      s_list[lp] = (uint16_t)(kScanLane * threadIdx.x + k);
      s_row[lp] = (r0 << 7) | c[k];
      ++lp;
    }
    r0 += c[k];
  }
  __syncthreads();
  const uint32_t base = offs[blockIdx.x];
  const size_t chunk0 = (size_t)blockIdx.x * kScanItems;
  for (uint32_t j = threadIdx.x; j < listed; j += 256) {
    const size_t p = chunk0 + s_list[j];
    const uint32_t nr = s_row[j] & 127u;
    uint32_t pos = base + (s_row[j] >> 7);
    for (uint32_t v = 0; v < nr && pos < cap; ++v, ++pos) {
      const float* r = row + 15 * (p * slots + v);
      float* f = feat + 12 * (size_t)pos;
      for (int q = 0; q < 12; ++q) f[q] = r[q];
      tgt[3 * (size_t)pos] = r[12];
      tgt[3 * (size_t)pos + 1] = r[13];
      tgt[3 * (size_t)pos + 2] = r[14];
    }
  }
}

__global__ __launch_bounds__(kBuildBlock) void k_sort_hist(const uint32_t* __restrict__ keys,
                                                           int32_t n, int shift, uint32_t* hist,
                                                           int32_t n_tiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t i = base + r * kBuildBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// One block per digit: exclusive scan of that digit's per-tile counts.
__global__ __launch_bounds__(kBuildBlock) void k_sort_scan(uint32_t* hist, int32_t n_tiles,
                                                           uint32_t* totals) {
  __shared__ uint32_t s_tmp[4];
  uint32_t* row = hist + (size_t)blockIdx.x * n_tiles;
  uint32_t running = 0;
  for (int32_t start = 0; start < n_tiles; start += kBuildBlock) {
    const int32_t i = start + threadIdx.x;
    const uint32_t v = i < n_tiles ? row[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exscan256(v, s_tmp, tot);
    if (i < n_tiles) row[i] = running + ex;
    running += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = running;
}

// Stable scatter: wave64 match via 8 ballots, ranks across the 4 waves of a
// round and across the kSortItems rounds of a tile in LDS.
__global__ __launch_bounds__(kBuildBlock) void k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* kout,
    uint32_t* vout, int32_t n, int shift, const uint32_t* __restrict__ hist,
    const uint32_t* __restrict__ totals, int32_t n_tiles) {
  __shared__ uint32_t s_tmp[4];
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_run[256];
  __shared__ uint32_t s_wcnt[4][256];
  __shared__ uint32_t s_woff[4][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t tot;
  const uint32_t dbase = block_exscan256(totals[tid], s_tmp, tot);
  s_base[tid] = dbase + hist[(size_t)tid * n_tiles + blockIdx.x];
  s_run[tid] = 0;
  for (int w = 0; w < 4; ++w) s_wcnt[w][tid] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t i = base + r * kBuildBlock + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool set = (d >> bit) & 1u;
      const uint64_t bm = __ballot(set);
      peers &= set ? bm : ~bm;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) s_wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    uint32_t run = s_run[tid];
    for (int w = 0; w < 4; ++w) {
      const uint32_t c = s_wcnt[w][tid];
      s_woff[w][tid] = run;
      run += c;
      s_wcnt[w][tid] = 0;
    }
    s_run[tid] = run;
    __syncthreads();
    if (valid) {
      const uint32_t pos = s_base[d] + s_woff[wave][d] + rank;
      if (pos >= (uint32_t)n) continue;  // unreachable when the histogram is consistent
      kout[pos] = key;
      vout[pos] = val;
    }
  }
}

// Sorted position k: local index order[k] -> global triangle p.  Writes the
// triangle (and normal) records and replaces order[k] by p in place.
__global__ __launch_bounds__(kBuildBlock) void k_gather(const float* __restrict__ v,
                                                        const float* __restrict__ nrm,
                                                        const int32_t* __restrict__ mat,
                                                        const uint32_t* __restrict__ gid,
                                                        uint32_t* order, int32_t T, float4* tris,
                                                        float4* tn) {
  const int32_t k = blockIdx.x * kBuildBlock + threadIdx.x;
  if (k >= T) return;
  const uint32_t local = order[k];
  if (local >= (uint32_t)T) return;
  const uint32_t p = global_id(gid, local);
  order[k] = p;
  const float* a = v + 9 * (size_t)p;
  tris[3 * (size_t)k + 0] = make_float4(a[0], a[1], a[2], __uint_as_float(p));
  tris[3 * (size_t)k + 1] = make_float4(a[3], a[4], a[5], __uint_as_float((uint32_t)mat[p]));
  tris[3 * (size_t)k + 2] = make_float4(a[6], a[7], a[8], 0.0f);
  if (tn) {
    const float* b = nrm + 9 * (size_t)p;
    tn[3 * (size_t)k + 0] = make_float4(b[0], b[1], b[2], 0.0f);
    tn[3 * (size_t)k + 1] = make_float4(b[3], b[4], b[5], 0.0f);
    tn[3 * (size_t)k + 2] = make_float4(b[6], b[7], b[8], 0.0f);
  }
}

__device__ __forceinline__ int delta(const uint32_t* __restrict__ keys, int64_t n, int64_t i,
                                     int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = keys[i], b = keys[j];
  if (a != b) return __clz((int)(a ^ b));
  return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

__global__ __launch_bounds__(kBuildBlock) void k_karras(const uint32_t* __restrict__ keys,
                                                        int32_t n, float4* nodes, int32_t* parent,
                                                        int32_t* leaf_parent, uint2* range) {
  const int64_t i = (int64_t)blockIdx.x * kBuildBlock + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int64_t lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  range[i] = make_uint2((uint32_t)(i < j ? i : j), (uint32_t)(i > j ? i : j));  // the node's leaves
  const int dnode = delta(keys, n, i, j);
  int64_t s = 0, t = 0, div = 2;
  do {
    t = (l + div - 1) / div;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    div *= 2;
  } while (t > 1);
  const int64_t g = i + s * d + (d < 0 ? -1 : 0);
  if (g < 0 || g + 1 >= n) return;  // unreachable for sorted keys
  const bool left_leaf = (i < j ? i : j) == g;
  const bool right_leaf = (i > j ? i : j) == g + 1;
  const uint32_t left = left_leaf ? (kLeaf | (uint32_t)g) : (uint32_t)g;
  const uint32_t right = right_leaf ? (kLeaf | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
  nodes[4 * i + 3] = make_float4(__uint_as_float(left), __uint_as_float(right), 0.0f, 0.0f);
  if (left_leaf)
    leaf_parent[g] = (int32_t)i;
  else
    parent[g] = (int32_t)i;
  if (right_leaf)
    leaf_parent[g + 1] = (int32_t)i;
  else
    parent[g + 1] = (int32_t)i;
}

// Triangle bounds widened outward by one ulp: an origin exactly on a face
// plane of the tight box with a zero direction component would make the
// slab test evaluate 0 * inf = NaN; inside the widened box it cannot.
__device__ __forceinline__ void leaf_box(const float4* __restrict__ tris, int32_t k, float* b) {
  const float4 a = tris[3 * (size_t)k], c = tris[3 * (size_t)k + 1], e = tris[3 * (size_t)k + 2];
  b[0] = nextafterf(fminf(fminf(a.x, c.x), e.x), -INFINITY);
  b[1] = nextafterf(fminf(fminf(a.y, c.y), e.y), -INFINITY);
  b[2] = nextafterf(fminf(fminf(a.z, c.z), e.z), -INFINITY);
  b[3] = nextafterf(fmaxf(fmaxf(a.x, c.x), e.x), INFINITY);
  b[4] = nextafterf(fmaxf(fmaxf(a.y, c.y), e.y), INFINITY);
  b[5] = nextafterf(fmaxf(fmaxf(a.z, c.z), e.z), INFINITY);
}

__device__ __forceinline__ void store_slot(float4* nodes, int32_t node, int slot, const float* b) {
  float* f = reinterpret_cast<float*>(nodes) + 16 * (size_t)node + 6 * slot;
  for (int c = 0; c < 6; ++c) f[c] = b[c];
}

// Refit by levels of a union table, without atomics: node i of a Karras
// tree covers the sorted leaves range[i] = [lo, hi], so its box is the
// union of their boxes.  Level k of the table holds, for every leaf i, the
// union of the 2^k boxes from i on (clamped at the last leaf); a node with
// L = hi - lo + 1 leaves, k = floor(log2 L), takes table_k[lo] | table_k[hi -
// 2^k + 1] (two windows that cover the range).  Each level is one streaming
// pass; the kernel boundary orders the levels.  min / max are exact, so the
// boxes equal a bottom-up union's bit for bit.
__device__ __forceinline__ void box_union(const float4* __restrict__ tab, uint32_t a, uint32_t c, float* b) {
  const float4 l0 = tab[2 * (size_t)a], h0 = tab[2 * (size_t)a + 1];
  const float4 l1 = tab[2 * (size_t)c], h1 = tab[2 * (size_t)c + 1];
  b[0] = fminf(l0.x, l1.x);
  b[1] = fminf(l0.y, l1.y);
  b[2] = fminf(l0.z, l1.z);
  b[3] = fmaxf(h0.x, h1.x);
  b[4] = fmaxf(h0.y, h1.y);
  b[5] = fmaxf(h0.z, h1.z);
}

// Level 0: the leaf boxes, also stored in their parents' child slots.
__global__ __launch_bounds__(kBuildBlock) void k_rf_leaves(const float4* __restrict__ tris, int32_t n,
                                                           float4* nodes,
                                                           const int32_t* __restrict__ leaf_parent,
                                                           float4* __restrict__ tab) {
  const int32_t k = blockIdx.x * kBuildBlock + threadIdx.x;
  if (k >= n) return;
  float b[6];
  leaf_box(tris, k, b);
  tab[2 * (size_t)k] = make_float4(b[0], b[1], b[2], 0.0f);
  tab[2 * (size_t)k + 1] = make_float4(b[3], b[4], b[5], 0.0f);
  const int32_t p = leaf_parent[k];
  if (p < 0 || p >= n - 1) return;
  const uint32_t r0 = __float_as_uint(nodes[4 * (size_t)p + 3].x);
  store_slot(nodes, p, r0 == (kLeaf | (uint32_t)k) ? 0 : 1, b);
}

// Level k from level k - 1 (when `next`), and the boxes of the internal
// nodes of level k - 1 into their parents' slots (the root has none).
__global__ __launch_bounds__(kBuildBlock) void k_rf_level(int32_t n, int32_t k, const float4* __restrict__ prev,
                                                          float4* __restrict__ next,
                                                          const uint2* __restrict__ range,
                                                          const int32_t* __restrict__ parent, float4* nodes) {
  const int32_t i = blockIdx.x * kBuildBlock + threadIdx.x;
  if (i >= n) return;
  if (next) {
    float b[6];
    box_union(prev, (uint32_t)i, (uint32_t)min(i + (1 << (k - 1)), n - 1), b);
    next[2 * (size_t)i] = make_float4(b[0], b[1], b[2], 0.0f);
    next[2 * (size_t)i + 1] = make_float4(b[3], b[4], b[5], 0.0f);
  }
  if (i == 0 || i >= n - 1) return;
  const uint2 r = range[i];
  const int32_t lev = 31 - __clz((int)(r.y - r.x + 1));
  if (lev != k - 1) return;
  float b[6];
  box_union(prev, r.x, r.y - (1u << lev) + 1u, b);
  const int32_t p = parent[i];
  if (p < 0 || p >= n - 1) return;  // unreachable for a valid tree
  const uint32_t r0 = __float_as_uint(nodes[4 * (size_t)p + 3].x);
  store_slot(nodes, p, r0 == (uint32_t)i ? 0 : 1, b);
}

// New positions (and normals) for sorted triangle k = global prim[k]; the
// ids in the .w lanes are kept.
__global__ __launch_bounds__(kBuildBlock) void k_regather(const float* __restrict__ v,
                                                          const float* __restrict__ nrm,
                                                          const uint32_t* __restrict__ prim,
                                                          int32_t T, float4* tris, float4* tn) {
  const int32_t k = blockIdx.x * kBuildBlock + threadIdx.x;
  if (k >= T) return;
  const uint32_t p = prim[k];
  const float* a = v + 9 * (size_t)p;
  for (int j = 0; j < 3; ++j) {
    float4 r = tris[3 * (size_t)k + j];
    r.x = a[3 * j];
    r.y = a[3 * j + 1];
    r.z = a[3 * j + 2];
    tris[3 * (size_t)k + j] = r;
  }
  if (tn && nrm) {
    const float* b = nrm + 9 * (size_t)p;
    for (int j = 0; j < 3; ++j)
      tn[3 * (size_t)k + j] = make_float4(b[3 * j], b[3 * j + 1], b[3 * j + 2], 0.0f);
  }
}

// T == 1: a root with the single leaf as child 0 and an empty child 1.
__global__ void k_single_root(const float4* __restrict__ tris, float4* nodes) {
  float b[6];
  leaf_box(tris, 0, b);
  nodes[0] = make_float4(b[0], b[1], b[2], b[3]);
  nodes[1] = make_float4(b[4], b[5], 0.0f, 0.0f);
  nodes[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  nodes[3] = make_float4(__uint_as_float(kLeaf), __uint_as_float(kEmpty), 0.0f, 0.0f);
}

// Child ch (0, 1) of the BVH2 node at np: its box and ref.
__device__ __forceinline__ uint32_t node_child(const float4* np, int ch, float lo[3], float hi[3]) {
  const float4 a = np[0], b = np[1], c = np[2], d = np[3];
  if (ch == 0) {
    lo[0] = a.x, lo[1] = a.y, lo[2] = a.z, hi[0] = a.w, hi[1] = b.x, hi[2] = b.y;
    return __float_as_uint(d.x);
  }
  lo[0] = b.z, lo[1] = b.w, lo[2] = c.x, hi[0] = c.y, hi[1] = c.z, hi[2] = c.w;
  return __float_as_uint(d.y);
}

// The 4-wide form of the tree: for every BVH2 inner node n, 128 bytes at
// wide[8n..8n+8) holding n's children, each inner child replaced by its own
// two children, so that a walk from the root visits the inner nodes of even
// depth only (an inner slot's ref is the grandchild's own index).  Structure
// of arrays over the four slots: lo.x, lo.y, lo.z, hi.x, hi.y, hi.z (one
// float4 each), the refs (kEmpty: no child; bit 31: leaf), one float4 unused.
__global__ __launch_bounds__(kBuildBlock) void k_wide(const float4* __restrict__ nodes, int32_t nI,
                                                      float4* __restrict__ wide) {
  const int32_t n = blockIdx.x * kBuildBlock + threadIdx.x;
  if (n >= nI) return;
  float lo[3][4] = {}, hi[3][4] = {};
  uint32_t ref[4] = {kEmpty, kEmpty, kEmpty, kEmpty};
  int j = 0;
  const float4* np = nodes + 4 * (size_t)n;
  for (int ch = 0; ch < 2; ++ch) {
    float l[3], h[3];
    const uint32_t r = node_child(np, ch, l, h);
    if (r == kEmpty) continue;
    if (r & kLeaf) {
      for (int a = 0; a < 3; ++a) lo[a][j] = l[a], hi[a][j] = h[a];
      ref[j++] = r;
      continue;
    }
    const float4* gp = nodes + 4 * (size_t)r;
    for (int gc = 0; gc < 2; ++gc) {
      const uint32_t g = node_child(gp, gc, l, h);
      if (g == kEmpty) continue;
      for (int a = 0; a < 3; ++a) lo[a][j] = l[a], hi[a][j] = h[a];
      ref[j++] = g;
    }
  }
  float4* w = wide + 8 * (size_t)n;
  for (int a = 0; a < 3; ++a) {
    w[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
    w[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
  }
  w[6] = make_float4(__uint_as_float(ref[0]), __uint_as_float(ref[1]), __uint_as_float(ref[2]),
                     __uint_as_float(ref[3]));
  w[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
// A device LBVH over T triangles of the scene (see the layout at the top).
struct DevBvh {
  int32_t T = 0, nI = 0;
  float4* nodes = nullptr;
  float4* wide = nullptr;  // the 4-wide form (k_wide), 128 bytes per inner node; null: none
  float4* tris = nullptr;
  uint32_t* morton = nullptr;
  uint32_t* prim = nullptr;       // global triangle ids, Morton order
  int32_t* parent = nullptr;      // parent of each internal node (kept for refits)
  int32_t* leaf_parent = nullptr; // parent of each leaf
  uint2* range = nullptr;         // sorted leaves [lo, hi] of each internal node (kept for refits)
  uint64_t bytes = 0;
};

// Device buffer that only grows: reallocated when a call needs more.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t want) {
    if (p && want <= cap) return hipSuccess;
    (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(want, 1) * sizeof(T));
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Device memory of lisa_update_vertices_noalloc (the refit for callers that
// cannot allocate): the two union tables of the level refit, 2 float4 per
// triangle each, for scenes of at most kNoAllocMaxTris triangles (64 MB per
// device, held until the process ends).  Host code never uses the array's
// name: the address on each device comes from hipGetSymbolAddress, taken
// once per device in lisa_create with that device current.
constexpr int32_t kNoAllocMaxTris = 1 << 20;
static __device__ __align__(256) float4 scratch[4 * (size_t)kNoAllocMaxTris];
constexpr int kMaxDevices = 64;
static float4* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Device buffers of the wavefront form: the state of `slots` paths in
// flight (WfState), the two request lists, the counters, and (for a single
// render) the one-entry scene table.  Grown on demand, one allocation.
struct WfBuffers {
  void* mem = nullptr;
  size_t bytes = 0;
  BatchScene* d_scenes = nullptr;
  int32_t scene_cap = 0;
  WfState ws{};
  hipError_t ensure(uint32_t slots) {
    const size_t P = ((size_t)slots + 63) / 64 * 64;
    const size_t want = P * (7 * 16 + 16 + 8 + 4 + 4) + 256;
    if (!mem || want > bytes) {
      (void)hipFree(mem);
      mem = nullptr;
      bytes = 0;
      const hipError_t e = hipMalloc(&mem, want);
      if (e != hipSuccess) return e;
      bytes = want;
    }
    char* c = static_cast<char*>(mem);
    float4** f4[7] = {&ws.O, &ws.D, &ws.col, &ws.att, &ws.N, &ws.Din, &ws.acc};
    for (float4** f : f4) {
      *f = reinterpret_cast<float4*>(c);
      c += 16 * P;
    }
    ws.st = reinterpret_cast<uint4*>(c);
    c += 16 * P;
    ws.res = reinterpret_cast<uint2*>(c);
    c += 8 * P;
    ws.q[0] = reinterpret_cast<uint32_t*>(c);
    c += 4 * P;
    ws.q[1] = reinterpret_cast<uint32_t*>(c);
    c += 4 * P;
    ws.cnt = reinterpret_cast<uint32_t*>(c);
    ws.slots = slots;
    return hipSuccess;
  }
  hipError_t ensure_scenes(int32_t n) {
    if (d_scenes && n <= scene_cap) return hipSuccess;
    (void)hipFree(d_scenes);
    d_scenes = nullptr;
    scene_cap = 0;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_scenes), sizeof(BatchScene) * (size_t)n);
    if (e == hipSuccess) scene_cap = n;
    return e;
  }
  void release() {
    (void)hipFree(mem);
    (void)hipFree(d_scenes);
    mem = nullptr;
    d_scenes = nullptr;
    bytes = 0;
    scene_cap = 0;
  }
};

// The rounds of the wavefront form on `st`: reset + start, then rounds of
// (k_wf_trace, k_wf_shade) in blocks of kWfRoundsPerCheck, after each block
// reading the request count (the loop ends when no path needs a traversal).
constexpr int kWfRoundsPerCheck = 16;
constexpr int kWfWavesPerCu = 32;   // k_wf_trace: 8 waves per SIMD
constexpr int kWfShadePerCu = 16;   // k_wf_shade: 4 waves per SIMD
hipError_t wf_enqueue_start(hipStream_t st, const BatchScene* d_scenes, int32_t n, uint32_t total,
                            const WfState& ws) {
  k_wf_reset<<<1, 64, 0, st>>>(ws.cnt);
  k_wf_start<<<(ws.slots + kWfBlock - 1) / kWfBlock, kWfBlock, 0, st>>>(d_scenes, n, total, ws);
  return hipGetLastError();
}
hipError_t wf_enqueue_rounds(hipStream_t st, const BatchScene* d_scenes, int32_t n, uint32_t total,
                             const WfState& ws, int32_t trace_blocks, int32_t shade_blocks) {
  uint32_t cur = 0;
  for (int r = 0; r < kWfRoundsPerCheck; ++r) {  // an even count: list 0 is the input again
    k_wf_trace<<<trace_blocks, kWfBlock, 0, st>>>(d_scenes, n, ws, cur);
    k_wf_shade<<<shade_blocks, kWfBlock, 0, st>>>(d_scenes, n, total, ws, cur);
    cur ^= 1u;
  }
  return hipGetLastError();
}
// Reads the pending request count after a block of rounds (list 0).
hipError_t wf_pending(hipStream_t st, const WfState& ws, uint32_t* pending) {
  hipError_t e = hipMemcpyAsync(pending, ws.cnt, 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

// lisa_render_progressive: n launches of the plain megakernel, one per
// subframe, chained in one instantiated hipGraph; before each launch the
// nodes' arguments (subframe index, buffers) are set on the executable graph,
// so one graph serves every run of n subframes at one image size.
// Triangle count from which a scene's tree also gets its 4-wide form
// (closest_wide): 2^20 triangles (64 MB of nodes and 48 MB of triangle
// records) no longer stay in the caches, and the fewer, wider visits pay;
// below it the binary walk is the faster one.
constexpr int64_t kWideMinTris = 1 << 20;

struct ProgGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipGraphNode_t> nodes;
  uint32_t grid = 0;
  bool f64 = false;
  const void* fn = nullptr;  // the kernel of the nodes
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    nodes.clear();
    grid = 0;
    fn = nullptr;
  }
};

struct lisa_ctx {
  int32_t device = 0;
  int32_t cus = 256;  // compute units (persistent grids)
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // cached frames: the training paths beside the frame's (made on first use)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  int32_t T = 0, M = 0;
  DevBvh all, emit;  // over every triangle / over the emitting triangles
  float4* d_nrm = nullptr;
  DevMat* d_mats = nullptr;
  bool has_scene = false;
  float4* d_accum = nullptr;
  size_t accum_px = 0;
  unsigned long long* d_counters = nullptr;
  uint64_t counters[LISA_NCOUNTERS] = {0, 0, 0, 0, 0, 0, 0, 0};
  float render_ms = 0.0f, bvh_ms = 0.0f;
  uint64_t scene_bytes = 0, accum_bytes = 0;
  float bmin[3] = {0, 0, 0}, bmax[3] = {1, 1, 1};  // scene box (cache features)
  float ecull_c[3] = {0, 0, 0}, ecull_h[3] = {INFINITY, INFINITY, INFINITY};  // never culls
  // radiance-cache frame scratch (lisa_render_cached), grown on demand
  DevBuf<float4> c_frame, c_tacc;
  DevBuf<float> c_qfeat, c_qatt, c_qout, c_trec, c_tfeat, c_ttgt;
  DevBuf<int32_t> c_qitem, c_tpix;
  DevBuf<uint32_t> c_qcount;
  // query and training-row slots before compaction (one per sample), the
  // flags / row counts of those slots and the compaction's chunk offsets
  DevBuf<float4> c_qrec;
  DevBuf<float> c_trow;
  DevBuf<uint8_t> c_qhas, c_tn;
  DevBuf<uint32_t> c_qscan, c_tscan;
  DevBuf<float> c_ecdf;     // area prefix sums of the emitting triangles (LISA_CACHE_EMISSION)
  float emit_area = 0.0f;
  // the plain tracer judging cached frames (judge_pixels > 0): its pixels,
  // per pixel-channel sums and squares of its samples and of the cached
  // frames' values there, samples per pixel and frames judged, and the frame
  // from which plain frames replace cached ones (-1: none)
  std::vector<int32_t> j_pix;
  std::vector<double> j_sum, j_sq, j_csum, j_csq;
  uint32_t j_n = 0, j_f = 0;
  int64_t j_from = -1;
  // the cached sequence in the accum buffer: started by a successful cached
  // frame 0 (false after any other write of the buffer), with j_nj judge
  // pixels; later frames must continue it with the same count
  bool j_seq = false;
  uint32_t j_nj = 0;
  // bumped whenever a device buffer a batch captured may have moved or its
  // captured values changed (scene, refit, accum buffer)
  uint64_t gen = 0;
  DevBuf<int32_t> c_jpix;
  DevBuf<float4> c_jout, c_jacc;
  DevBuf<float> c_jrgb;
  WfBuffers wf;  // the wavefront form of lisa_render (LISA_FLAG_WAVEFRONT)
  ProgGraph pg;  // lisa_render_progressive's chain of subframe launches
  // scenes of at least this many triangles get the 4-wide form of their tree
  // and the path kernels walk it (lisa_set_wide_threshold)
  int64_t wide_min_tris = kWideMinTris;
  // FP64 progressive sums (LISA_FLAG_FP64_ACCUM full renders): r, g, b sums
  // and the sample count per pixel; valid while they belong to the accum
  // buffer's current sequence (an FP64 subframe 0, or lisa_accum64_load)
  DevBuf<double4> c_sum64;
  bool sum64_valid = false;
};

struct lisa_batch {
  int32_t device = 0, n = 0, cus = 256;
  std::vector<lisa_ctx*> ctxs;
  std::vector<lisa_render_params> params;
  std::vector<size_t> accum_px;
  std::vector<uint64_t> gens;  // each context's generation when the batch was made
  uint32_t total = 0;
  hipStream_t stream = nullptr;
  BatchScene* d_scenes = nullptr;
  uint32_t* d_counter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // form 2: the start and a block of rounds, captured separately (the host
  // replays the rounds until no path needs a traversal)
  hipGraph_t graph2 = nullptr;
  hipGraphExec_t exec2 = nullptr;
  WfBuffers wf;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  lisa_batch_config cfg{0, 1, 0, 1};
  float last_ms = 0.0f;
};

namespace {

lisa_status fail(lisa_ctx* ctx, lisa_status code, const std::string& msg) {
  if (ctx)
    ctx->err = msg;
  else
    lisa::set_global_error(msg);
  return code;
}

#define HIP_TRY(ctx, expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail((ctx), LISA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

void free_bvh(DevBvh& b) {
  (void)hipFree(b.nodes);
  (void)hipFree(b.wide);
  (void)hipFree(b.tris);
  (void)hipFree(b.morton);
  (void)hipFree(b.prim);
  (void)hipFree(b.parent);
  (void)hipFree(b.leaf_parent);
  (void)hipFree(b.range);
  b = DevBvh();
}

void free_scene(lisa_ctx* c) {
  c->gen++;
  c->j_seq = false;
  free_bvh(c->all);
  free_bvh(c->emit);
  (void)hipFree(c->d_nrm);
  (void)hipFree(c->d_mats);
  c->d_nrm = nullptr;
  c->d_mats = nullptr;
  c->T = c->M = 0;
  c->has_scene = false;
  c->scene_bytes = 0;
}

SceneView view(const lisa_ctx* c) {
  return SceneView{BvhView{c->all.nodes, c->all.tris, c->all.T, c->all.wide},
                   BvhView{c->emit.nodes, c->emit.tris, c->emit.T, nullptr},
                   c->d_nrm,
                   c->d_mats,
                   make_float4(c->ecull_c[0], c->ecull_c[1], c->ecull_c[2], 0.0f),
                   make_float4(c->ecull_h[0], c->ecull_h[1], c->ecull_h[2], 0.0f)};
}

// Device scratch freed at scope exit.
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T** p, size_t count) {
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// The context's accum buffer for npx pixels (reallocated: a new generation).
lisa_status ensure_accum(lisa_ctx* ctx, size_t npx) {
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  (void)hipFree(ctx->d_accum);
  ctx->d_accum = nullptr;
  ctx->accum_px = 0;
  ctx->gen++;
  ctx->j_seq = false;
  ctx->sum64_valid = false;
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->d_accum), npx * sizeof(float4)));
  ctx->accum_px = npx;
  ctx->accum_bytes = npx * sizeof(float4);
  return LISA_OK;
}

lisa_status check_params(lisa_ctx* ctx, const lisa_render_params* p) {
  if (!p) return fail(ctx, LISA_ERR_INVALID_ARG, "render params is NULL");
  if ((p->flags & LISA_FLAG_PERSISTENT) && (p->flags & (LISA_FLAG_COUNT | LISA_FLAG_FP64_ACCUM)))
    return fail(ctx, LISA_ERR_UNSUPPORTED, "the persistent launch has no counting or FP64 form");
  if ((p->flags & LISA_FLAG_WAVEFRONT) &&
      (p->flags & (LISA_FLAG_COUNT | LISA_FLAG_FP64_ACCUM | LISA_FLAG_PERSISTENT)))
    return fail(ctx, LISA_ERR_UNSUPPORTED, "the wavefront form has no counting, FP64 or persistent variant");
  if (p->width == 0 || p->height == 0) return fail(ctx, LISA_ERR_INVALID_ARG, "width/height must be > 0");
  if ((uint64_t)p->width * p->height > (1ull << 31))
    return fail(ctx, LISA_ERR_INVALID_ARG, "image larger than 2^31 pixels");
  if (p->spp == 0) return fail(ctx, LISA_ERR_INVALID_ARG, "spp must be > 0");
  if (p->mode != LISA_MODE_PLAIN) return fail(ctx, LISA_ERR_UNSUPPORTED, "unknown render mode");
  if (!ctx->has_scene) return fail(ctx, LISA_ERR_NO_SCENE, "lisa_set_scene has not been called");
  return LISA_OK;
}

PathArgs make_args(const lisa_render_params* p) {
  PathArgs a;
  std::memset(&a, 0, sizeof(a));
  a.width = p->width;
  a.height = p->height;
  a.spp = p->spp;
  a.bounces = p->bounces;
  a.subframe = p->subframe;
  a.eye = mk(p->eye[0], p->eye[1], p->eye[2]);
  a.U = mk(p->U[0], p->U[1], p->U[2]);
  a.V = mk(p->V[0], p->V[1], p->V[2]);
  a.W = mk(p->W[0], p->W[1], p->W[2]);
  a.fp64 = (p->flags & LISA_FLAG_FP64_ACCUM) ? 1 : 0;
  return a;
}

// Every child box of b (T >= 2) from its triangles: the leaf level, then one
// k_rf_level pass per level of the union table (scratch: two tables of T
// entries, 32 bytes each).  Enqueued on ctx->stream; waits before returning
// (the scratch is freed).
// The same passes on caller-provided tables (2 * T float4 each); no allocation, no wait.
lisa_status refit_levels_on(lisa_ctx* ctx, DevBvh& b, float4* ta, float4* tb) {
  hipStream_t st = ctx->stream;
  const int32_t n = b.T;
  k_rf_leaves<<<grid_for(n, kBuildBlock), kBuildBlock, 0, st>>>(b.tris, n, b.nodes, b.leaf_parent, ta);
  const int32_t K = 31 - __builtin_clz((unsigned)n);  // the root's level
  for (int32_t k = 1; k <= K + 1; ++k) {
    k_rf_level<<<grid_for(n, kBuildBlock), kBuildBlock, 0, st>>>(n, k, ta, k <= K ? tb : nullptr, b.range,
                                                               b.parent, b.nodes);
    std::swap(ta, tb);
  }
  HIP_TRY(ctx, hipGetLastError());
  return LISA_OK;
}
lisa_status refit_levels(lisa_ctx* ctx, DevBvh& b) {
  Scratch S;
  float4 *ta = nullptr, *tb = nullptr;
  HIP_TRY(ctx, S.alloc(&ta, 2 * (size_t)b.T));
  HIP_TRY(ctx, S.alloc(&tb, 2 * (size_t)b.T));
  lisa_status rc = refit_levels_on(ctx, b, ta, tb);
  if (rc != LISA_OK) return rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the tables are freed on return
  return LISA_OK;
}

// The 4-wide form of b from its nodes (when b has one), on stream st.
lisa_status wide_nodes(lisa_ctx* ctx, DevBvh& b, hipStream_t st) {
  if (!b.wide || b.T == 0) return LISA_OK;
  k_wide<<<grid_for(b.nI, kBuildBlock), kBuildBlock, 0, st>>>(b.nodes, b.nI, b.wide);
  HIP_TRY(ctx, hipGetLastError());
  return LISA_OK;
}

// LBVH over the n triangles gid[0..n) (all T when gid is null) of the soup
// already on the device: centroid bounds -> 30-bit Morton codes -> 4-pass LDS
// radix sort -> Morton-ordered triangle records -> Karras hierarchy -> refit.
// Enqueued on ctx->stream; `nrm_out` (optional) receives the normal records.
lisa_status build_bvh(lisa_ctx* ctx, DevBvh& b, const float* d_v, const float* d_n,
                      const int32_t* d_mat, const uint32_t* d_gid, int32_t n, float4* nrm_out,
                      bool wide = false) {
  hipStream_t st = ctx->stream;
  b.T = n;
  b.nI = n > 1 ? n - 1 : 1;
  const size_t nodes_b = 64 * (size_t)b.nI, tris_b = 48 * (size_t)std::max(n, 1);
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&b.nodes), nodes_b));
  if (wide) {
    HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&b.wide), 2 * nodes_b));
    b.bytes += 2 * nodes_b;
  }
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&b.tris), tris_b));
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&b.morton), 4 * (size_t)std::max(n, 1)));
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&b.prim), 4 * (size_t)std::max(n, 1)));
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&b.parent), 4 * (size_t)std::max(n, 1)));
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&b.leaf_parent), 4 * (size_t)std::max(n, 1)));
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&b.range), 8 * (size_t)std::max(n, 1)));
  b.bytes += nodes_b + tris_b + 16 * (size_t)std::max(n, 1);
  HIP_TRY(ctx, hipMemsetAsync(b.nodes, 0, nodes_b, st));
  if (n == 0) return LISA_OK;
  Scratch S;
  float *d_part = nullptr, *d_bounds = nullptr;
  int32_t *d_parent = b.parent, *d_leafp = b.leaf_parent;
  uint32_t *d_k2 = nullptr, *d_v2 = nullptr, *d_hist = nullptr, *d_tot = nullptr;
  const int32_t n_tiles = (n + kSortTile - 1) / kSortTile;
  const int32_t n_part = std::min<int32_t>(1024, (n + kBuildBlock - 1) / kBuildBlock);
  HIP_TRY(ctx, S.alloc(&d_part, 6 * (size_t)n_part));
  HIP_TRY(ctx, S.alloc(&d_bounds, 6));
  HIP_TRY(ctx, S.alloc(&d_k2, (size_t)n));
  HIP_TRY(ctx, S.alloc(&d_v2, (size_t)n));
  HIP_TRY(ctx, S.alloc(&d_hist, 256 * (size_t)n_tiles));
  HIP_TRY(ctx, S.alloc(&d_tot, 256));
  k_centroid_partial<<<n_part, kBuildBlock, 0, st>>>(d_v, d_gid, n, d_part);
  k_centroid_final<<<1, kBuildBlock, 0, st>>>(d_part, n_part, d_bounds);
  k_morton<<<grid_for(n, kBuildBlock), kBuildBlock, 0, st>>>(d_v, d_gid, n, d_bounds, b.morton, b.prim);
  uint32_t *ka = b.morton, *va = b.prim, *kb = d_k2, *vb = d_v2;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 8 * pass;
    k_sort_hist<<<n_tiles, kBuildBlock, 0, st>>>(ka, n, shift, d_hist, n_tiles);
    k_sort_scan<<<256, kBuildBlock, 0, st>>>(d_hist, n_tiles, d_tot);
    k_sort_scatter<<<n_tiles, kBuildBlock, 0, st>>>(ka, va, kb, vb, n, shift, d_hist, d_tot, n_tiles);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  // four passes: the sorted data is back in b.morton / b.prim
  k_gather<<<grid_for(n, kBuildBlock), kBuildBlock, 0, st>>>(d_v, d_n, d_mat, d_gid, b.prim, n,
                                                             b.tris, nrm_out);
  if (n == 1) {
    k_single_root<<<1, 1, 0, st>>>(b.tris, b.nodes);
  } else {
    HIP_TRY(ctx, hipMemsetAsync(d_parent, 0xFF, sizeof(int32_t) * (size_t)n, st));
    HIP_TRY(ctx, hipMemsetAsync(d_leafp, 0xFF, sizeof(int32_t) * (size_t)n, st));
    k_karras<<<grid_for(n - 1, kBuildBlock), kBuildBlock, 0, st>>>(b.morton, n, b.nodes, d_parent,
                                                                  d_leafp, b.range);
    lisa_status rc = refit_levels(ctx, b);
    if (rc != LISA_OK) return rc;
  }
  lisa_status rw = wide_nodes(ctx, b, st);
  if (rw != LISA_OK) return rw;
  HIP_TRY(ctx, hipGetLastError());
  // scratch is freed on return: wait for the kernels that use it
  HIP_TRY(ctx, hipStreamSynchronize(st));
  return LISA_OK;
}

// Rewrites b's triangle records from the soup d_v (Morton order kept) and
// refits every box bottom-up with the stored parent links.
lisa_status refit_bvh(lisa_ctx* ctx, DevBvh& b, const float* d_v, const float* d_n, float4* nrm) {
  if (b.T == 0) return LISA_OK;
  hipStream_t st = ctx->stream;
  k_regather<<<grid_for(b.T, kBuildBlock), kBuildBlock, 0, st>>>(d_v, d_n, b.prim, b.T, b.tris, nrm);
  if (b.T == 1) {
    k_single_root<<<1, 1, 0, st>>>(b.tris, b.nodes);
  } else {
    lisa_status rc = refit_levels(ctx, b);
    if (rc != LISA_OK) return rc;
  }
  lisa_status rw = wide_nodes(ctx, b, st);
  if (rw != LISA_OK) return rw;
  HIP_TRY(ctx, hipStreamSynchronize(st));
  HIP_TRY(ctx, hipGetLastError());
  return LISA_OK;
}

}  // namespace

// Persistent launch of the path loop in mode kMode over `a`'s pixel list or
// full image: 20 workgroups of 64 per CU (the occupancy of the path kernels),
// work handed out by counter slot `slot` of the context's scratch.
constexpr int kPersistentBlocksPerCu = 20;
template <int kMode>
lisa_status launch_persistent(lisa_ctx* ctx, PathArgs a, int slot) {
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, ctx->c_qcount.ensure(64));
  uint32_t* counter = ctx->c_qcount.p + slot;
  HIP_TRY(ctx, hipMemsetAsync(counter, 0, 4, st));
  uint32_t total;
  if (a.pixels) {
    total = (uint32_t)a.n_items;
  } else {
    a.tiles_x = (a.width + 7) / 8;
    total = a.tiles_x * ((a.height + 7) / 8) * 64u;
  }
  const unsigned blocks = (unsigned)(ctx->cus * kPersistentBlocksPerCu);
  if (ctx->all.wide)
    k_path_pq<kMode, true><<<blocks, kPathBlock, 0, st>>>(view(ctx), a, counter, total, 1u);
  else
    k_path_pq<kMode><<<blocks, kPathBlock, 0, st>>>(view(ctx), a, counter, total, 1u);
  HIP_TRY(ctx, hipGetLastError());
  return LISA_OK;
}

// The plain render `a` (full image) through the wavefront form: the work
// list is the image's 8x8 tiles; `a.tiles_x` counts 8-pixel tiles here.
lisa_status render_wavefront(lisa_ctx* ctx, PathArgs a) {
  hipStream_t st = ctx->stream;
  a.tiles_x = (a.width + 7) / 8;
  const uint32_t total = a.tiles_x * ((a.height + 7) / 8) * 64u;
  const uint32_t slots = std::min<uint32_t>(total, (uint32_t)ctx->cus * kWfWavesPerCu * kWfBlock * 2u);
  HIP_TRY(ctx, ctx->wf.ensure(slots));
  HIP_TRY(ctx, ctx->wf.ensure_scenes(1));
  BatchScene hs;
  hs.sc = view(ctx);
  hs.p = a;
  hs.first = 0;
  HIP_TRY(ctx, hipMemcpyAsync(ctx->wf.d_scenes, &hs, sizeof(hs), hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, wf_enqueue_start(st, ctx->wf.d_scenes, 1, total, ctx->wf.ws));
  // every path ends: at most spp * (bounces * (kNeeTries + 1) + 1) + 1
  // traversals per pixel, pixels in turn on at most `total` slots
  const uint64_t max_rounds = (uint64_t)a.spp * ((uint64_t)a.bounces * (kNeeTries + 1) + 1) *
                              (((uint64_t)total + slots - 1) / slots) + 2 * kWfRoundsPerCheck;
  for (uint64_t r = 0;; r += kWfRoundsPerCheck) {
    if (r > max_rounds) return fail(ctx, LISA_ERR_HIP, "wavefront form: paths did not finish");
    HIP_TRY(ctx, wf_enqueue_rounds(st, ctx->wf.d_scenes, 1, total, ctx->wf.ws, ctx->cus * kWfWavesPerCu,
                                   ctx->cus * kWfShadePerCu));
    uint32_t pending = 0;
    HIP_TRY(ctx, wf_pending(st, ctx->wf.ws, &pending));
    if (pending == 0) break;
  }
  return LISA_OK;
}

// Area prefix sums of the emitting triangles in emitter-tree order, in a
// fixed FP32 summation order: chunks of kCdfChunk triangles are summed
// sequentially from 0 (p_k, one work-item per chunk), the chunk totals
// sequentially into bases b_c (one work-item: T / 64 additions), and
// cdf[k] = b_c + p_k.  With one chunk these are the plain sequential sums.
constexpr int32_t kCdfChunk = 64;
__global__ __launch_bounds__(256) void k_emit_chunks(const float4* __restrict__ tris, int32_t T,
                                                     float* cdf, float* chunk_tot) {
  const int32_t c = (int32_t)(blockIdx.x * 256 + threadIdx.x);
  const int32_t k0 = c * kCdfChunk;
  if (k0 >= T) return;
  const int32_t k1 = min(k0 + kCdfChunk, T);
  float acc = 0.0f;
  for (int32_t k = k0; k < k1; ++k) {
    const F3 v0 = xyz(tris[3 * (size_t)k]), v1 = xyz(tris[3 * (size_t)k + 1]), v2 = xyz(tris[3 * (size_t)k + 2]);
    const F3 e1 = v1 - v0, e2 = v2 - v0;
    const F3 x = mk(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    acc = acc + 0.5f * sqrtf(dot(x, x));
    cdf[k] = acc;
  }
  chunk_tot[c] = acc;
}
__global__ void k_emit_bases(float* chunk_tot, int32_t n_chunks) {
  float base = 0.0f;
  for (int32_t c = 0; c < n_chunks; ++c) {
    const float t = chunk_tot[c];
    chunk_tot[c] = base;
    base = base + t;
  }
}
__global__ __launch_bounds__(256) void k_emit_add(float* cdf, int32_t T, const float* __restrict__ base) {
  const int32_t k = (int32_t)(blockIdx.x * 256 + threadIdx.x);
  if (k < T) cdf[k] = base[k / kCdfChunk] + cdf[k];
}

// Union of the child boxes of a root node (its 16 words).
void root_box(const float* rb, float lo[3], float hi[3]) {
  uint32_t ref[2];
  std::memcpy(ref, rb + 12, 8);
  for (int a = 0; a < 3; ++a) {
    lo[a] = INFINITY;
    hi[a] = -INFINITY;
  }
  for (int ch = 0; ch < 2; ++ch) {
    if (ref[ch] == kEmpty) continue;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], rb[6 * ch + a]);
      hi[a] = std::max(hi[a], rb[6 * ch + 3 + a]);
    }
  }
}

// The scene box (root of `all`) and the emitter cull box (emit_culled): the
// emitter root box inflated by S * 2^-11 per side, S the largest |coordinate|
// of the scene box.
lisa_status read_scene_box(lisa_ctx* ctx, bool may_alloc = true) {
  const int64_t T = ctx->all.T;
  for (int a = 0; a < 3; ++a) {
    ctx->ecull_c[a] = 0.0f;
    ctx->ecull_h[a] = INFINITY;
  }
  if (T <= 0) return LISA_OK;
  float rb[16];
  HIP_TRY(ctx, hipMemcpy(rb, ctx->all.nodes, 64, hipMemcpyDeviceToHost));
  root_box(rb, ctx->bmin, ctx->bmax);
  ctx->emit_area = 0.0f;  // no light samples unless the emitters below have area
  if (ctx->emit.T <= 0) return LISA_OK;
  float S = 0.0f;
  for (int a = 0; a < 3; ++a) S = std::max(S, std::max(std::fabs(ctx->bmin[a]), std::fabs(ctx->bmax[a])));
  const float m2 = S * 0x1p-11f;
  float lo[3], hi[3];
  HIP_TRY(ctx, hipMemcpy(rb, ctx->emit.nodes, 64, hipMemcpyDeviceToHost));
  root_box(rb, lo, hi);
  // light sampling of LISA_CACHE_EMISSION: area prefix sums of the emitters
  const int32_t TE = ctx->emit.T, n_chunks = (TE + kCdfChunk - 1) / kCdfChunk;
  if (may_alloc)
    HIP_TRY(ctx, ctx->c_ecdf.ensure((size_t)TE + (size_t)n_chunks));
  else if (ctx->c_ecdf.cap < (size_t)TE + (size_t)n_chunks)  // sized by lisa_set_scene for these emitters
    return fail(ctx, LISA_ERR_ALLOC, "emitter prefix-sum buffer smaller than the scene's emitters");
  float* chunk_tot = ctx->c_ecdf.p + TE;
  k_emit_chunks<<<grid_for(n_chunks, 256), 256, 0, ctx->stream>>>(ctx->emit.tris, TE, ctx->c_ecdf.p, chunk_tot);
  k_emit_bases<<<1, 1, 0, ctx->stream>>>(chunk_tot, n_chunks);
  k_emit_add<<<grid_for(TE, 256), 256, 0, ctx->stream>>>(ctx->c_ecdf.p, TE, chunk_tot);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipMemcpyAsync(&ctx->emit_area, ctx->c_ecdf.p + (TE - 1), 4, hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int a = 0; a < 3; ++a) {
    ctx->ecull_c[a] = 0.5f * (lo[a] + hi[a]);
    ctx->ecull_h[a] = 0.5f * (hi[a] - lo[a]) + m2;
  }
  return LISA_OK;
}

// Launch arguments of a cache-mode path loop: the plain ones plus the
// position normalisation of the scene box (union of the root's child boxes).
PathArgs cache_args(const lisa_ctx* ctx, const lisa_render_params* p) {
  PathArgs a = make_args(p);
  a.bmin = mk(ctx->bmin[0], ctx->bmin[1], ctx->bmin[2]);
  a.binv = mk(1.0f / std::max(ctx->bmax[0] - ctx->bmin[0], 1e-20f),
              1.0f / std::max(ctx->bmax[1] - ctx->bmin[1], 1e-20f),
              1.0f / std::max(ctx->bmax[2] - ctx->bmin[2], 1e-20f));
  a.fp64 = 0;
  a.emit_expect = 0;
  a.nee_loops = 1;
  a.nee_loops_deep = 1;
  a.emit_cdf = ctx->c_ecdf.p;
  a.emit_area = ctx->emit_area;
  a.emit_area_below = std::nextafter(ctx->emit_area, 0.0f);
  return a;
}

// Order-keeping compaction of n per-slot counts on stream st: chunk offsets
// into `offs`, the total into *d_total (device).
lisa_status scan_counts(lisa_ctx* ctx, const uint8_t* cnt, size_t n, DevBuf<uint32_t>& offs,
                        uint32_t* d_total, hipStream_t st) {
  const size_t nb = (n + kScanItems - 1) / kScanItems;
  HIP_TRY(ctx, offs.ensure(nb));
  if (nb > 0) k_scan_chunks<<<(unsigned)nb, 256, 0, st>>>(cnt, n, offs.p);
  k_scan_top<<<1, 256, 0, st>>>(offs.p, (uint32_t)nb, d_total);
  HIP_TRY(ctx, hipGetLastError());
  return LISA_OK;
}

// Training paths (path loop kMode 2) on n listed device pixels with a's spp
// and bounces: rows (features, suffix targets) of every surface hit, or of the
// hits of radiance ray number `depth` > 0 only, into ctx->c_tfeat / c_ttgt in
// slot order (pixel, sample, vertex), at most `cap`; *rows = rows produced
// (may exceed cap).
lisa_status trace_train_rows(lisa_ctx* ctx, PathArgs a, const int32_t* d_pix, size_t n, size_t cap,
                             uint32_t depth, uint32_t* rows, hipStream_t st) {
  // vertex slots per path: one hit per sample at a fixed depth, else up to one per bounce
  const size_t slots = depth > 0 ? 1 : a.bounces;
  const size_t paths = n * a.spp;
  HIP_TRY(ctx, ctx->c_trec.ensure(n * slots * 18));
  HIP_TRY(ctx, ctx->c_trow.ensure(paths * slots * 15));
  HIP_TRY(ctx, ctx->c_tn.ensure(paths));
  HIP_TRY(ctx, ctx->c_tfeat.ensure(cap * 12));
  HIP_TRY(ctx, ctx->c_ttgt.ensure(cap * 3));
  HIP_TRY(ctx, ctx->c_tacc.ensure(n));
  HIP_TRY(ctx, ctx->c_qcount.ensure(64));
  HIP_TRY(ctx, hipMemsetAsync(ctx->c_tn.p, 0, paths, st));
  a.pixels = d_pix;
  a.n_items = (int32_t)n;
  a.accum = ctx->c_tacc.p;
  a.accum_prev = nullptr;
  a.t_rec = ctx->c_trec.p;
  a.t_slots = (uint32_t)slots;
  a.t_row = ctx->c_trow.p;
  a.t_n = ctx->c_tn.p;
  a.t_depth = depth;
  if (ctx->all.wide)
    k_path_mode<2, true><<<grid_for((int64_t)n, kPathBlock), kPathBlock, 0, st>>>(view(ctx), a);
  else
    k_path_mode<2><<<grid_for((int64_t)n, kPathBlock), kPathBlock, 0, st>>>(view(ctx), a);
  HIP_TRY(ctx, hipGetLastError());
  lisa_status rc = scan_counts(ctx, ctx->c_tn.p, paths, ctx->c_tscan, ctx->c_qcount.p + 1, st);
  if (rc != LISA_OK) return rc;
  if (paths > 0) {
    k_compact_rows<<<(unsigned)((paths + kScanItems - 1) / kScanItems), 256, 0, st>>>(
        ctx->c_tn.p, ctx->c_trow.p, paths, (uint32_t)slots, ctx->c_tscan.p, (uint32_t)cap, ctx->c_tfeat.p,
        ctx->c_ttgt.p);
    HIP_TRY(ctx, hipGetLastError());
  }
  HIP_TRY(ctx, hipMemcpyAsync(rows, ctx->c_qcount.p + 1, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  return LISA_OK;
}

// Prefix paths (path loop kMode 1, or 3 with a.emit_expect): the full image in tiles (d_pix null) or
// n listed device pixels; radiance gathered before the cache into `frame`
// (per pixel mean), a query slot per sample.  The launch returns without
// waiting; trace_queries_finish compacts the queries into ctx->c_q* in slot
// order (pixel, then sample) with their features, and waits for the count
// (*nq).
lisa_status trace_queries_launch(lisa_ctx* ctx, PathArgs& a, const int32_t* d_pix, size_t n, uint32_t prefix,
                                 float4* frame, bool persistent, bool full = false) {
  hipStream_t st = ctx->stream;
  const size_t slots = full ? 0 : n * a.spp;
  HIP_TRY(ctx, ctx->c_qrec.ensure(slots * 4));
  HIP_TRY(ctx, ctx->c_qhas.ensure(slots));
  HIP_TRY(ctx, ctx->c_qfeat.ensure(slots * 12));
  HIP_TRY(ctx, ctx->c_qatt.ensure(slots * 3));
  HIP_TRY(ctx, ctx->c_qout.ensure(slots * 3));
  HIP_TRY(ctx, ctx->c_qitem.ensure(slots));
  HIP_TRY(ctx, ctx->c_qcount.ensure(64));
  if (slots > 0) HIP_TRY(ctx, hipMemsetAsync(ctx->c_qhas.p, 0, slots, st));
  a.pixels = d_pix;
  a.n_items = (int32_t)n;
  a.accum = frame;
  a.accum_prev = nullptr;
  // full paths (LISA_CACHE_FULL_PATHS): a.bounces radiance rays and no
  // query (radiance ray numbers start at 1)
  if (!full) a.bounces = prefix + 1;
  a.cache_at = full ? 0u : prefix + 1;
  a.q_rec = ctx->c_qrec.p;
  a.q_has = ctx->c_qhas.p;
  if (persistent) {
    lisa_status rc = a.emit_expect ? launch_persistent<3>(ctx, a, 2) : launch_persistent<1>(ctx, a, 2);
    if (rc != LISA_OK) return rc;
  } else {
    unsigned blocks;
    if (d_pix) {
      blocks = grid_for((int64_t)n, kPathBlock);
    } else {
      a.tiles_x = (a.width + kTileW - 1) / kTileW;
      blocks = a.tiles_x * ((a.height + kTileH - 1) / kTileH);
    }
    const bool wide = ctx->all.wide != nullptr;
    if (a.emit_expect && wide)
      k_path_mode<3, true><<<blocks, kPathBlock, 0, st>>>(view(ctx), a);
    else if (a.emit_expect)
      k_path_mode<3><<<blocks, kPathBlock, 0, st>>>(view(ctx), a);
    else if (wide)
      k_path_mode<1, true><<<blocks, kPathBlock, 0, st>>>(view(ctx), a);
    else
      k_path_mode<1><<<blocks, kPathBlock, 0, st>>>(view(ctx), a);
    HIP_TRY(ctx, hipGetLastError());
  }
  return LISA_OK;
}

lisa_status trace_queries_finish(lisa_ctx* ctx, const PathArgs& a, uint32_t* nq) {
  hipStream_t st = ctx->stream;
  const size_t slots = a.cache_at == 0 ? 0 : (size_t)a.n_items * a.spp;
  if (slots > 0) {
    lisa_status rc = scan_counts(ctx, ctx->c_qhas.p, slots, ctx->c_qscan, ctx->c_qcount.p, st);
    if (rc != LISA_OK) return rc;
    k_compact_queries<<<(unsigned)((slots + kScanItems - 1) / kScanItems), 256, 0, st>>>(
        ctx->c_qhas.p, ctx->c_qrec.p, slots, ctx->c_qscan.p, a.spp, a.bmin, a.binv, ctx->d_mats,
        ctx->c_qfeat.p, ctx->c_qatt.p, ctx->c_qitem.p);
    HIP_TRY(ctx, hipGetLastError());
  } else {
    HIP_TRY(ctx, hipMemsetAsync(ctx->c_qcount.p, 0, 4, st));
  }
  HIP_TRY(ctx, hipMemcpyAsync(nq, ctx->c_qcount.p, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));  // the network reads them on the cache's stream
  return LISA_OK;
}

lisa_status trace_queries(lisa_ctx* ctx, PathArgs a, const int32_t* d_pix, size_t n, uint32_t prefix,
                          float4* frame, uint32_t* nq, bool persistent = false, bool full = false) {
  lisa_status rc = trace_queries_launch(ctx, a, d_pix, n, prefix, frame, persistent, full);
  if (rc != LISA_OK) return rc;
  return trace_queries_finish(ctx, a, nq);
}

extern "C" {

lisa_status lisa_update_vertices(lisa_ctx* ctx, const float* verts, const float* normals) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!ctx->has_scene) return fail(ctx, LISA_ERR_NO_SCENE, "lisa_set_scene has not been called");
  const int32_t T = ctx->T;
  if (T == 0) return LISA_OK;
  if (!verts) return fail(ctx, LISA_ERR_INVALID_ARG, "verts is NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  Scratch S;
  float *d_v = nullptr, *d_n = nullptr;
  HIP_TRY(ctx, S.alloc(&d_v, 9 * (size_t)T));
  HIP_TRY(ctx, hipMemcpy(d_v, verts, 36 * (size_t)T, hipMemcpyHostToDevice));
  if (normals) {
    HIP_TRY(ctx, S.alloc(&d_n, 9 * (size_t)T));
    HIP_TRY(ctx, hipMemcpy(d_n, normals, 36 * (size_t)T, hipMemcpyHostToDevice));
  }
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  lisa_status rc = refit_bvh(ctx, ctx->all, d_v, d_n, ctx->d_nrm);
  if (rc == LISA_OK) rc = refit_bvh(ctx, ctx->emit, d_v, nullptr, nullptr);
  if (rc != LISA_OK) return rc;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  HIP_TRY(ctx, hipEventElapsedTime(&ctx->bvh_ms, ctx->ev0, ctx->ev1));
  ctx->gen++;  // the cull box a batch captured may have changed
  return read_scene_box(ctx);
}

lisa_status lisa_update_vertices_noalloc(lisa_ctx* ctx, const float* d_verts, const float* d_normals) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!ctx->has_scene) return fail(ctx, LISA_ERR_NO_SCENE, "lisa_set_scene has not been called");
  const int32_t T = ctx->T;
  if (T == 0) return LISA_OK;
  if (!d_verts) return fail(ctx, LISA_ERR_INVALID_ARG, "d_verts is NULL");
  if (T > kNoAllocMaxTris)
    return fail(ctx, LISA_ERR_INVALID_ARG,
                "lisa_update_vertices_noalloc takes scenes of at most 1048576 triangles (this one has " +
                    std::to_string(T) + "); use lisa_update_vertices");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  float4* const ptr = g_scratch_ptr[ctx->device];  // set by lisa_create on this device
  if (!ptr) return fail(ctx, LISA_ERR_HIP, "no refit tables on this device");
  float4* const ta = ptr;
  float4* const tb = ptr + 2 * (size_t)kNoAllocMaxTris;
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  DevBvh* trees[2] = {&ctx->all, &ctx->emit};
  for (DevBvh* b : trees) {
    if (b->T == 0) continue;
    const bool all = b == &ctx->all;
    k_regather<<<grid_for(b->T, kBuildBlock), kBuildBlock, 0, st>>>(d_verts, all ? d_normals : nullptr, b->prim,
                                                                    b->T, b->tris, all ? ctx->d_nrm : nullptr);
    if (b->T == 1) {
      k_single_root<<<1, 1, 0, st>>>(b->tris, b->nodes);
    } else {
      lisa_status rc = refit_levels_on(ctx, *b, ta, tb);
      if (rc != LISA_OK) return rc;
    }
    lisa_status rw = wide_nodes(ctx, *b, st);
    if (rw != LISA_OK) return rw;
  }
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  HIP_TRY(ctx, hipEventElapsedTime(&ctx->bvh_ms, ctx->ev0, ctx->ev1));
  ctx->gen++;  // the cull box a batch captured may have changed
  return read_scene_box(ctx, false);
}

lisa_status lisa_create(lisa_ctx** out, int32_t device) {
  if (!out) return fail(nullptr, LISA_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fail(nullptr, LISA_ERR_HIP, std::string("no HIP device: ") + hipGetErrorString(e));
  if (device < 0 || device >= n) return fail(nullptr, LISA_ERR_INVALID_ARG, "device index out of range");
  lisa_ctx* c = new lisa_ctx;
  c->device = device;
  if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreate(&c->stream)) != hipSuccess ||
      (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->d_counters),
                     LISA_NCOUNTERS * sizeof(unsigned long long))) != hipSuccess ||
      (e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device)) !=
          hipSuccess) {
    const std::string msg = std::string("lisa_create: ") + hipGetErrorString(e);
    lisa_destroy(c);
    return fail(nullptr, LISA_ERR_HIP, msg);
  }
  // this device's copy of the allocation-free refit's tables (once per device)
  if (device >= kMaxDevices) {
    lisa_destroy(c);
    return fail(nullptr, LISA_ERR_UNSUPPORTED, "device index >= 64");
  }
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void* ptr = nullptr;
      if ((e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch))) != hipSuccess || !ptr) {
        const std::string msg = std::string("lisa_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        lisa_destroy(c);
        return fail(nullptr, LISA_ERR_HIP, msg);
      }
      g_scratch_ptr[device] = static_cast<float4*>(ptr);
    }
  }
  *out = c;
  return LISA_OK;
}

void lisa_destroy(lisa_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->c_frame.release();
  c->c_tacc.release();
  c->c_qfeat.release();
  c->c_qatt.release();
  c->c_qout.release();
  c->c_trec.release();
  c->c_tfeat.release();
  c->c_ttgt.release();
  c->c_qitem.release();
  c->c_tpix.release();
  c->c_qcount.release();
  c->c_qrec.release();
  c->c_trow.release();
  c->c_qhas.release();
  c->c_tn.release();
  c->c_qscan.release();
  c->c_tscan.release();
  c->c_ecdf.release();
  c->c_jpix.release();
  c->c_jout.release();
  c->c_jacc.release();
  c->c_jrgb.release();
  c->wf.release();
  c->c_sum64.release();
  c->pg.release();
  free_scene(c);
  (void)hipFree(c->d_accum);
  (void)hipFree(c->d_counters);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  delete c;
}

const char* lisa_last_error(const lisa_ctx* c) { return c ? c->err.c_str() : lisa::global_error(); }

lisa_status lisa_set_scene(lisa_ctx* ctx, const float* verts, const float* normals,
                           const int32_t* mat_idx, int32_t T, const lisa_material* mats,
                           int32_t M) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (T < 0 || M < 0) return fail(ctx, LISA_ERR_INVALID_ARG, "negative T or M");
  if (T > 0 && (!verts || !normals || !mat_idx))
    return fail(ctx, LISA_ERR_INVALID_ARG, "verts/normals/mat_idx NULL with T > 0");
  if (T > 0 && (M <= 0 || !mats)) return fail(ctx, LISA_ERR_INVALID_ARG, "no materials");
  if (T >= (1 << 30)) return fail(ctx, LISA_ERR_INVALID_ARG, "too many triangles (>= 2^30)");
  for (int32_t k = 0; k < T; ++k)
    if (mat_idx[k] < 0 || mat_idx[k] >= M)
      return fail(ctx, LISA_ERR_INVALID_ARG, "material index out of range at triangle " + std::to_string(k));
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  free_scene(ctx);
  ctx->T = T;
  ctx->M = M;
  ctx->bvh_ms = 0.0f;
  std::vector<DevMat> hm(std::max(M, 1));
  for (int32_t i = 0; i < M; ++i) {
    const lisa_material& m = mats[i];
    hm[i].diffuse_alpha = make_float4(m.diffuse[0], m.diffuse[1], m.diffuse[2], m.alpha);
    hm[i].emission_emit = make_float4(m.emission[0], m.emission[1], m.emission[2], m.emit ? 1.0f : 0.0f);
    hm[i].rough_n = make_float4(m.roughness, m.n, 0.0f, 0.0f);
  }
  std::vector<uint32_t> emit_ids;
  for (int32_t k = 0; k < T; ++k)
    if (mats[mat_idx[k]].emit) emit_ids.push_back((uint32_t)k);
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->d_mats), sizeof(DevMat) * hm.size()));
  HIP_TRY(ctx, hipMemcpy(ctx->d_mats, hm.data(), sizeof(DevMat) * hm.size(), hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->d_nrm), 48 * (size_t)std::max(T, 1)));
  ctx->scene_bytes = sizeof(DevMat) * hm.size() + 48 * (size_t)std::max(T, 1);
  Scratch S;
  float *d_v = nullptr, *d_n = nullptr;
  int32_t* d_mat = nullptr;
  uint32_t* d_gid = nullptr;
  if (T > 0) {
    HIP_TRY(ctx, S.alloc(&d_v, 9 * (size_t)T));
    HIP_TRY(ctx, S.alloc(&d_n, 9 * (size_t)T));
    HIP_TRY(ctx, S.alloc(&d_mat, (size_t)T));
    HIP_TRY(ctx, hipMemcpy(d_v, verts, 9 * sizeof(float) * (size_t)T, hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(d_n, normals, 9 * sizeof(float) * (size_t)T, hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(d_mat, mat_idx, sizeof(int32_t) * (size_t)T, hipMemcpyHostToDevice));
  }
  if (!emit_ids.empty()) {
    HIP_TRY(ctx, S.alloc(&d_gid, emit_ids.size()));
    HIP_TRY(ctx, hipMemcpy(d_gid, emit_ids.data(), 4 * emit_ids.size(), hipMemcpyHostToDevice));
  }
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  lisa_status rc = build_bvh(ctx, ctx->all, d_v, d_n, d_mat, nullptr, T, ctx->d_nrm,
                             (int64_t)T >= ctx->wide_min_tris);
  if (rc == LISA_OK)
    rc = build_bvh(ctx, ctx->emit, d_v, d_n, d_mat, d_gid, (int32_t)emit_ids.size(), nullptr);
  if (rc != LISA_OK) return rc;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  HIP_TRY(ctx, hipEventElapsedTime(&ctx->bvh_ms, ctx->ev0, ctx->ev1));
  ctx->scene_bytes += ctx->all.bytes + ctx->emit.bytes;
  rc = read_scene_box(ctx);
  if (rc != LISA_OK) return rc;
  ctx->has_scene = true;
  return LISA_OK;
}

lisa_status lisa_render(lisa_ctx* ctx, const lisa_render_params* p, float* accum_rgba_out,
                        uint8_t* rgba8_out) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  lisa_status rc = check_params(ctx, p);
  if (rc != LISA_OK) return rc;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t npx = (size_t)p->width * p->height;
  if (ctx->accum_px != npx) {
    if (p->subframe > 0)
      return fail(ctx, LISA_ERR_INVALID_ARG, "subframe > 0 but the accum buffer has another size");
    rc = ensure_accum(ctx, npx);
    if (rc != LISA_OK) return rc;
  }
  ctx->j_seq = false;  // the buffer now holds a plain render
  const bool f64 = (p->flags & LISA_FLAG_FP64_ACCUM) != 0;
  if (f64) {
    if (p->subframe > 0 && !ctx->sum64_valid)
      return fail(ctx, LISA_ERR_INVALID_ARG,
                  "FP64 subframe > 0 needs the FP64 sums of subframe 0 (or lisa_accum64_load)");
    HIP_TRY(ctx, ctx->c_sum64.ensure(npx));
  }
  ctx->sum64_valid = false;  // set again when an FP64 render succeeds
  PathArgs a = make_args(p);
  a.tiles_x = (p->width + kTileW - 1) / kTileW;
  const uint32_t tiles_y = (p->height + kTileH - 1) / kTileH;
  a.n_items = (int32_t)npx;
  a.accum = ctx->d_accum;
  a.accum_prev = ctx->d_accum;
  a.counters = ctx->d_counters;
  a.sum64 = f64 ? ctx->c_sum64.p : nullptr;
  const bool count = (p->flags & LISA_FLAG_COUNT) != 0;
  hipStream_t st = ctx->stream;
  if (count)
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_counters, 0, LISA_NCOUNTERS * sizeof(unsigned long long), st));
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  const unsigned grid = a.tiles_x * tiles_y;
  if (p->flags & LISA_FLAG_WAVEFRONT) {
    lisa_status rc2 = render_wavefront(ctx, a);
    if (rc2 != LISA_OK) return rc2;
  } else if (p->flags & LISA_FLAG_PERSISTENT) {
    lisa_status rc2 = launch_persistent<0>(ctx, a, 4);
    if (rc2 != LISA_OK) return rc2;
  } else if (count && a.fp64)
    k_path<false, true, true><<<grid, kPathBlock, 0, st>>>(view(ctx), a);
  else if (count)
    k_path<false, true><<<grid, kPathBlock, 0, st>>>(view(ctx), a);
  else if (a.fp64 && ctx->all.wide)
    k_path<false, false, true, true><<<grid, kPathBlock, 0, st>>>(view(ctx), a);
  else if (a.fp64)
    k_path<false, false, true><<<grid, kPathBlock, 0, st>>>(view(ctx), a);
  else if (ctx->all.wide)
    k_path<false, false, false, true><<<grid, kPathBlock, 0, st>>>(view(ctx), a);
  else
    k_path<false, false><<<grid, kPathBlock, 0, st>>>(view(ctx), a);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  if (accum_rgba_out || rgba8_out) {
    std::vector<float> tmp;
    float* dst = accum_rgba_out;
    if (!dst) {
      tmp.resize(npx * 4);
      dst = tmp.data();
    }
    HIP_TRY(ctx, hipMemcpyAsync(dst, ctx->d_accum, npx * sizeof(float4), hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    if (rgba8_out) lisa_srgb_quantize(dst, (int64_t)npx, rgba8_out);
  }
  HIP_TRY(ctx, hipStreamSynchronize(st));
  HIP_TRY(ctx, hipEventElapsedTime(&ctx->render_ms, ctx->ev0, ctx->ev1));
  ctx->sum64_valid = f64;  // a plain render leaves sums that no longer match the buffer
  if (count) {
    unsigned long long h[LISA_NCOUNTERS];
    HIP_TRY(ctx, hipMemcpy(h, ctx->d_counters, sizeof(h), hipMemcpyDeviceToHost));
    for (int i = 0; i < LISA_NCOUNTERS; ++i) ctx->counters[i] = h[i];
  }
  return LISA_OK;
}

lisa_status lisa_render_progressive(lisa_ctx* ctx, const lisa_render_params* p, int32_t n_subframes) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  lisa_status rc = check_params(ctx, p);
  if (rc != LISA_OK) return rc;
  if (n_subframes <= 0 || n_subframes > 4096)
    return fail(ctx, LISA_ERR_INVALID_ARG, "n_subframes must be in 1..4096");
  if ((uint64_t)p->subframe + (uint64_t)n_subframes > 0xFFFFFFFFull)
    return fail(ctx, LISA_ERR_INVALID_ARG, "subframe + n_subframes overflows the subframe index");
  if (p->flags & ~LISA_FLAG_FP64_ACCUM)
    return fail(ctx, LISA_ERR_UNSUPPORTED, "progressive graphs take no flag but LISA_FLAG_FP64_ACCUM");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t npx = (size_t)p->width * p->height;
  if (ctx->accum_px != npx) {
    if (p->subframe > 0)
      return fail(ctx, LISA_ERR_INVALID_ARG, "subframe > 0 but the accum buffer has another size");
    rc = ensure_accum(ctx, npx);
    if (rc != LISA_OK) return rc;
  }
  ctx->j_seq = false;  // the buffer now holds a plain render
  const bool f64 = (p->flags & LISA_FLAG_FP64_ACCUM) != 0;
  if (f64) {
    if (p->subframe > 0 && !ctx->sum64_valid)
      return fail(ctx, LISA_ERR_INVALID_ARG,
                  "FP64 subframe > 0 needs the FP64 sums of subframe 0 (or lisa_accum64_load)");
    HIP_TRY(ctx, ctx->c_sum64.ensure(npx));
  }
  ctx->sum64_valid = false;  // set again when an FP64 run succeeds
  PathArgs a = make_args(p);
  a.tiles_x = (p->width + kTileW - 1) / kTileW;
  const uint32_t grid = a.tiles_x * ((p->height + kTileH - 1) / kTileH);
  a.n_items = (int32_t)npx;
  a.accum = ctx->d_accum;
  a.accum_prev = ctx->d_accum;
  a.counters = ctx->d_counters;
  a.sum64 = f64 ? ctx->c_sum64.p : nullptr;
  SceneView sv = view(ctx);
  const bool wide = ctx->all.wide != nullptr;
  void* const fn = f64 ? (wide ? reinterpret_cast<void*>(&k_path<false, false, true, true>)
                               : reinterpret_cast<void*>(&k_path<false, false, true>))
                       : (wide ? reinterpret_cast<void*>(&k_path<false, false, false, true>)
                               : reinterpret_cast<void*>(&k_path<false, false, false>));
  auto node_params = [&](PathArgs& pa, void** args) {
    hipKernelNodeParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.func = fn;
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(kPathBlock);
    kp.sharedMemBytes = 0;
    args[0] = &sv;
    args[1] = &pa;
    kp.kernelParams = args;
    kp.extra = nullptr;
    return kp;
  };
  ProgGraph& g = ctx->pg;
  hipError_t e = hipSuccess;
  if (!g.exec || g.nodes.size() != (size_t)n_subframes || g.grid != grid || g.f64 != f64 || g.fn != fn) {
    g.release();
    e = hipGraphCreate(&g.graph, 0);
    for (int32_t i = 0; e == hipSuccess && i < n_subframes; ++i) {
      void* args[2];
      hipKernelNodeParams kp = node_params(a, args);
      hipGraphNode_t node = nullptr;
      const hipGraphNode_t prev = i > 0 ? g.nodes.back() : nullptr;
      e = hipGraphAddKernelNode(&node, g.graph, i > 0 ? &prev : nullptr, i > 0 ? 1 : 0, &kp);
      if (e == hipSuccess) g.nodes.push_back(node);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      g.release();
      return fail(ctx, LISA_ERR_HIP, std::string("progressive graph: ") + hipGetErrorString(e));
    }
    g.grid = grid;
    g.f64 = f64;
    g.fn = fn;
  }
  for (int32_t i = 0; e == hipSuccess && i < n_subframes; ++i) {  // subframe p->subframe + i
    PathArgs ai = a;
    ai.subframe = p->subframe + (uint32_t)i;
    void* args[2];
    hipKernelNodeParams kp = node_params(ai, args);
    e = hipGraphExecKernelNodeSetParams(g.exec, g.nodes[(size_t)i], &kp);
  }
  if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
  if (e == hipSuccess) e = hipGraphLaunch(g.exec, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    g.release();
    return fail(ctx, LISA_ERR_HIP, std::string("lisa_render_progressive: ") + hipGetErrorString(e));
  }
  HIP_TRY(ctx, hipEventElapsedTime(&ctx->render_ms, ctx->ev0, ctx->ev1));
  ctx->sum64_valid = f64;
  return LISA_OK;
}

lisa_status lisa_trace_subset(lisa_ctx* ctx, const lisa_render_params* p, const int32_t* pixels,
                              int32_t n_pixels, float* accum_rgba_out, float* sample_rgb,
                              int32_t* sample_bounces, int32_t* sample_occl,
                              uint32_t* sample_hash, int32_t* rad_hits, int32_t record_bounces) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  lisa_status rc = check_params(ctx, p);
  if (rc != LISA_OK) return rc;
  if (n_pixels <= 0 || !pixels || !accum_rgba_out)
    return fail(ctx, LISA_ERR_INVALID_ARG, "pixels / accum_rgba_out required, n_pixels > 0");
  if (record_bounces < 0) return fail(ctx, LISA_ERR_INVALID_ARG, "record_bounces < 0");
  const int64_t npx = (int64_t)p->width * p->height;
  for (int32_t i = 0; i < n_pixels; ++i)
    if (pixels[i] < 0 || pixels[i] >= npx)
      return fail(ctx, LISA_ERR_INVALID_ARG, "pixel id out of range");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t ns = (size_t)n_pixels * p->spp;
  Scratch S;
  int32_t* d_px = nullptr;
  float4* d_acc = nullptr;
  float* d_rgb = nullptr;
  int32_t *d_b = nullptr, *d_o = nullptr, *d_rh = nullptr;
  uint32_t* d_h = nullptr;
  HIP_TRY(ctx, S.alloc(&d_px, (size_t)n_pixels));
  HIP_TRY(ctx, S.alloc(&d_acc, (size_t)n_pixels));
  if (sample_rgb) HIP_TRY(ctx, S.alloc(&d_rgb, 3 * ns));
  if (sample_bounces) HIP_TRY(ctx, S.alloc(&d_b, ns));
  if (sample_occl) HIP_TRY(ctx, S.alloc(&d_o, ns));
  if (sample_hash) HIP_TRY(ctx, S.alloc(&d_h, ns));
  if (rad_hits && record_bounces > 0) HIP_TRY(ctx, S.alloc(&d_rh, ns * record_bounces));
  HIP_TRY(ctx, hipMemcpy(d_px, pixels, sizeof(int32_t) * n_pixels, hipMemcpyHostToDevice));
  PathArgs a = make_args(p);
  a.pixels = d_px;
  a.n_items = n_pixels;
  a.accum = d_acc;
  a.accum_prev = nullptr;
  a.s_rgb = d_rgb;
  a.s_b = d_b;
  a.s_o = d_o;
  a.s_h = d_h;
  a.rad_hits = d_rh;
  a.rb = d_rh ? record_bounces : 0;
  a.counters = ctx->d_counters;
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, hipMemsetAsync(ctx->d_counters, 0, LISA_NCOUNTERS * sizeof(unsigned long long), st));
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  if (a.fp64)
    k_path<true, true, true><<<grid_for(n_pixels, kPathBlock), kPathBlock, 0, st>>>(view(ctx), a);
  else
    k_path<true, true><<<grid_for(n_pixels, kPathBlock), kPathBlock, 0, st>>>(view(ctx), a);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  HIP_TRY(ctx, hipEventElapsedTime(&ctx->render_ms, ctx->ev0, ctx->ev1));
  HIP_TRY(ctx, hipMemcpy(accum_rgba_out, d_acc, sizeof(float4) * n_pixels, hipMemcpyDeviceToHost));
  if (d_rgb) HIP_TRY(ctx, hipMemcpy(sample_rgb, d_rgb, sizeof(float) * 3 * ns, hipMemcpyDeviceToHost));
  if (d_b) HIP_TRY(ctx, hipMemcpy(sample_bounces, d_b, sizeof(int32_t) * ns, hipMemcpyDeviceToHost));
  if (d_o) HIP_TRY(ctx, hipMemcpy(sample_occl, d_o, sizeof(int32_t) * ns, hipMemcpyDeviceToHost));
  if (d_h) HIP_TRY(ctx, hipMemcpy(sample_hash, d_h, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost));
  if (d_rh)
    HIP_TRY(ctx, hipMemcpy(rad_hits, d_rh, sizeof(int32_t) * ns * record_bounces, hipMemcpyDeviceToHost));
  unsigned long long h[LISA_NCOUNTERS];
  HIP_TRY(ctx, hipMemcpy(h, ctx->d_counters, sizeof(h), hipMemcpyDeviceToHost));
  for (int i = 0; i < LISA_NCOUNTERS; ++i) ctx->counters[i] = h[i];
  return LISA_OK;
}

lisa_status lisa_trace_rays(lisa_ctx* ctx, const float* O, const float* D, int32_t n, float tmin,
                            float tmax, int32_t* prim_out, float* t_out) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!ctx->has_scene) return fail(ctx, LISA_ERR_NO_SCENE, "lisa_set_scene has not been called");
  if (n < 0 || (n > 0 && (!O || !D || !prim_out || !t_out)))
    return fail(ctx, LISA_ERR_INVALID_ARG, "bad ray arrays");
  if (n == 0) return LISA_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  Scratch S;
  float *d_o = nullptr, *d_d = nullptr, *d_t = nullptr;
  int32_t* d_p = nullptr;
  HIP_TRY(ctx, S.alloc(&d_o, 3 * (size_t)n));
  HIP_TRY(ctx, S.alloc(&d_d, 3 * (size_t)n));
  HIP_TRY(ctx, S.alloc(&d_t, (size_t)n));
  HIP_TRY(ctx, S.alloc(&d_p, (size_t)n));
  HIP_TRY(ctx, hipMemcpy(d_o, O, 12 * (size_t)n, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(d_d, D, 12 * (size_t)n, hipMemcpyHostToDevice));
  hipStream_t st = ctx->stream;
  HIP_TRY(ctx, hipMemsetAsync(ctx->d_counters, 0, LISA_NCOUNTERS * sizeof(unsigned long long), st));
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  k_trace_rays<true><<<grid_for(n, kRayBlock), kRayBlock, 0, st>>>(view(ctx), d_o, d_d, n, tmin, tmax,
                                                                   d_p, d_t, ctx->d_counters);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  HIP_TRY(ctx, hipEventElapsedTime(&ctx->render_ms, ctx->ev0, ctx->ev1));
  HIP_TRY(ctx, hipMemcpy(prim_out, d_p, 4 * (size_t)n, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(t_out, d_t, 4 * (size_t)n, hipMemcpyDeviceToHost));
  unsigned long long h[LISA_NCOUNTERS];
  HIP_TRY(ctx, hipMemcpy(h, ctx->d_counters, sizeof(h), hipMemcpyDeviceToHost));
  for (int i = 0; i < LISA_NCOUNTERS; ++i) ctx->counters[i] = h[i];
  return LISA_OK;
}

lisa_status lisa_get_bvh(lisa_ctx* ctx, int32_t which, uint32_t* morton, uint32_t* prim,
                         uint32_t* child_refs, float* child_boxes) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!ctx->has_scene) return fail(ctx, LISA_ERR_NO_SCENE, "lisa_set_scene has not been called");
  if (which != 0 && which != 1) return fail(ctx, LISA_ERR_INVALID_ARG, "which must be 0 or 1");
  const DevBvh& b = which == 0 ? ctx->all : ctx->emit;
  if (b.T == 0) return LISA_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (morton) HIP_TRY(ctx, hipMemcpy(morton, b.morton, 4 * (size_t)b.T, hipMemcpyDeviceToHost));
  if (prim) HIP_TRY(ctx, hipMemcpy(prim, b.prim, 4 * (size_t)b.T, hipMemcpyDeviceToHost));
  if (child_refs || child_boxes) {
    std::vector<float> nodes(16 * (size_t)b.nI);
    HIP_TRY(ctx, hipMemcpy(nodes.data(), b.nodes, 64 * (size_t)b.nI, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < b.nI; ++i) {
      if (child_boxes) std::memcpy(child_boxes + 12 * (size_t)i, &nodes[16 * (size_t)i], 48);
      if (child_refs) std::memcpy(child_refs + 2 * (size_t)i, &nodes[16 * (size_t)i + 12], 8);
    }
  }
  return LISA_OK;
}

lisa_status lisa_copy_accum(lisa_ctx* ctx, float* rgba_out, int64_t n_pixels) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!rgba_out || n_pixels < 0 || (size_t)n_pixels != ctx->accum_px)
    return fail(ctx, LISA_ERR_INVALID_ARG, "rgba_out NULL or n_pixels != pixels of the last render");
  if (n_pixels == 0) return LISA_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemcpyAsync(rgba_out, ctx->d_accum, (size_t)n_pixels * sizeof(float4),
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LISA_OK;
}

lisa_status lisa_accum64_save(lisa_ctx* ctx, double* sums_out, int64_t n_pixels) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!sums_out || n_pixels <= 0 || (size_t)n_pixels != ctx->accum_px)
    return fail(ctx, LISA_ERR_INVALID_ARG, "sums_out NULL or n_pixels != pixels of the last render");
  if (!ctx->sum64_valid)
    return fail(ctx, LISA_ERR_INVALID_ARG, "no FP64 sums: the last render was not LISA_FLAG_FP64_ACCUM");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemcpyAsync(sums_out, ctx->c_sum64.p, (size_t)n_pixels * sizeof(double4), hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LISA_OK;
}

lisa_status lisa_accum64_load(lisa_ctx* ctx, const double* sums, int64_t n_pixels) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!sums || n_pixels <= 0 || n_pixels > (int64_t)1 << 31)
    return fail(ctx, LISA_ERR_INVALID_ARG, "sums NULL or n_pixels out of range");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t npx = (size_t)n_pixels;
  if (ctx->accum_px != npx) {
    lisa_status rc = ensure_accum(ctx, npx);
    if (rc != LISA_OK) return rc;
  }
  ctx->j_seq = false;
  HIP_TRY(ctx, ctx->c_sum64.ensure(npx));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->c_sum64.p, sums, npx * sizeof(double4), hipMemcpyHostToDevice, ctx->stream));
  k_mean64<<<grid_for((int64_t)npx, 256), 256, 0, ctx->stream>>>(ctx->c_sum64.p, (int64_t)npx, ctx->d_accum);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->sum64_valid = true;
  return LISA_OK;
}

lisa_status lisa_batch_create(lisa_batch** out, lisa_ctx* const* ctxs,
                              const lisa_render_params* params, int32_t n) {
  if (!out) return fail(nullptr, LISA_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (n <= 0 || !ctxs || !params) return fail(nullptr, LISA_ERR_INVALID_ARG, "empty batch");
  const int32_t dev = ctxs[0] ? ctxs[0]->device : -1;
  uint64_t total = 0;
  std::vector<BatchScene> hs(n);
  for (int32_t i = 0; i < n; ++i) {
    lisa_ctx* c = ctxs[i];
    if (!c) return fail(nullptr, LISA_ERR_INVALID_ARG, "NULL context in batch");
    if (c->device != dev) return fail(nullptr, LISA_ERR_INVALID_ARG, "batch contexts on different devices");
    lisa_status rc = check_params(c, &params[i]);
    if (rc != LISA_OK) {
      lisa::set_global_error("batch scene " + std::to_string(i) + ": " + c->err);
      return rc;
    }
    if (params[i].subframe > 0) return fail(nullptr, LISA_ERR_UNSUPPORTED, "batched renders use subframe 0");
    if (params[i].flags & LISA_FLAG_COUNT) return fail(nullptr, LISA_ERR_UNSUPPORTED, "no counters in batches");
    if (params[i].flags & LISA_FLAG_FP64_ACCUM)
      return fail(nullptr, LISA_ERR_UNSUPPORTED, "no FP64 accumulation in batches");
    for (int32_t k = 0; k < i; ++k)
      if (ctxs[k] == c) return fail(nullptr, LISA_ERR_INVALID_ARG, "a context appears twice in the batch");
    if (hipSetDevice(dev) != hipSuccess) return fail(nullptr, LISA_ERR_HIP, "hipSetDevice");
    const size_t npx = (size_t)params[i].width * params[i].height;
    if (c->accum_px != npx && ensure_accum(c, npx) != LISA_OK) {
      lisa::set_global_error("batch scene " + std::to_string(i) + ": " + c->err);
      return LISA_ERR_ALLOC;
    }
    PathArgs a = make_args(&params[i]);
    a.tiles_x = (params[i].width + 7) / 8;
    const uint32_t tiles_y = (params[i].height + 7) / 8;
    a.n_items = (int32_t)npx;
    a.accum = c->d_accum;
    a.accum_prev = c->d_accum;
    hs[i].sc = view(c);
    hs[i].p = a;
    hs[i].first = (uint32_t)total;
    total += (uint64_t)a.tiles_x * tiles_y * 64;
  }
  if (total >= (1ull << 32) - (1ull << 20)) return fail(nullptr, LISA_ERR_INVALID_ARG, "batch too large");
  lisa_batch* b = new lisa_batch;
  b->ctxs.assign(ctxs, ctxs + n);
  b->params.assign(params, params + n);
  b->accum_px.resize(n);
  b->gens.resize(n);
  for (int32_t i = 0; i < n; ++i) {
    b->accum_px[i] = ctxs[i]->accum_px;
    b->gens[i] = ctxs[i]->gen;
  }
  b->device = dev;
  b->total = (uint32_t)total;
  b->stream = ctxs[0]->stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) b->cus = prop.multiProcessorCount;
  b->cfg.grid_blocks = b->cus * 16;
  b->cfg.chunk = 1;
  hipError_t e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&b->d_scenes), sizeof(BatchScene) * n)) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&b->d_counter), 256)) != hipSuccess ||
      (e = hipMemcpy(b->d_scenes, hs.data(), sizeof(BatchScene) * n, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess) {
    lisa_batch_destroy(b);
    return fail(nullptr, LISA_ERR_HIP, std::string("lisa_batch_create: ") + hipGetErrorString(e));
  }
  b->n = n;
  *out = b;
  return LISA_OK;
}

void lisa_batch_destroy(lisa_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->exec2) (void)hipGraphExecDestroy(b->exec2);
  if (b->graph2) (void)hipGraphDestroy(b->graph2);
  b->wf.release();
  (void)hipFree(b->d_scenes);
  (void)hipFree(b->d_counter);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  delete b;
}

lisa_status lisa_batch_set_config(lisa_batch* b, const lisa_batch_config* cfg) {
  if (!b || !cfg) return fail(nullptr, LISA_ERR_INVALID_ARG, "NULL argument");
  if (cfg->grid_blocks <= 0 || cfg->chunk <= 0 || cfg->chunk > 4096 || cfg->form < 0 || cfg->form > 2 ||
      (cfg->graph != 0 && cfg->graph != 1))
    return fail(nullptr, LISA_ERR_INVALID_ARG,
                "grid_blocks and chunk must be positive (chunk <= 4096), form 0, 1 or 2, graph 0 or 1");
  if (cfg->grid_blocks != b->cfg.grid_blocks || cfg->chunk != b->cfg.chunk || cfg->form != b->cfg.form ||
      cfg->graph != b->cfg.graph) {
    b->cfg = *cfg;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->exec2) (void)hipGraphExecDestroy(b->exec2);
    if (b->graph2) (void)hipGraphDestroy(b->graph2);
    b->exec = nullptr;
    b->graph = nullptr;
    b->exec2 = nullptr;
    b->graph2 = nullptr;
  }
  return LISA_OK;
}

lisa_status lisa_batch_get_config(const lisa_batch* b, lisa_batch_config* cfg) {
  if (!b || !cfg) return fail(nullptr, LISA_ERR_INVALID_ARG, "NULL argument");
  *cfg = b->cfg;
  return LISA_OK;
}

// The batch's launch on its stream: the queue reset (a kernel, so a captured
// graph holds no memset node) and the persistent path kernel of its form.
static uint32_t wf_batch_slots(const lisa_batch* b) {
  return (uint32_t)std::min<uint64_t>((uint64_t)b->total, (uint64_t)b->cfg.grid_blocks * kWfBlock * b->cfg.chunk);
}
static hipError_t enqueue_batch_rounds(lisa_batch* b) {
  return wf_enqueue_rounds(b->stream, b->d_scenes, b->n, b->total, b->wf.ws, b->cfg.grid_blocks,
                           b->cus * kWfShadePerCu);
}
static hipError_t enqueue_batch(lisa_batch* b) {
  if (b->cfg.form == 2) {  // the start and the first block of rounds
    const hipError_t e = wf_enqueue_start(b->stream, b->d_scenes, b->n, b->total, b->wf.ws);
    return e != hipSuccess ? e : enqueue_batch_rounds(b);
  }
  k_queue_reset<<<1, 1, 0, b->stream>>>(b->d_counter);
  if (b->cfg.form == 1)
    k_path_tiles<<<b->cfg.grid_blocks, kPathBlock, 0, b->stream>>>(b->d_scenes, b->n, b->d_counter, b->total);
  else
    k_path_queue<<<b->cfg.grid_blocks, kPathBlock, 0, b->stream>>>(b->d_scenes, b->n, b->d_counter, b->total,
                                                                   (uint32_t)b->cfg.chunk);
  return hipGetLastError();
}

lisa_status lisa_batch_render(lisa_batch* b, float* elapsed_ms) {
  if (!b) return fail(nullptr, LISA_ERR_INVALID_ARG, "batch is NULL");
  for (int32_t i = 0; i < b->n; ++i)
    if (b->ctxs[i]->accum_px != b->accum_px[i] || !b->ctxs[i]->has_scene || b->ctxs[i]->gen != b->gens[i])
      return fail(nullptr, LISA_ERR_INVALID_ARG,
                  "a context of the batch changed (scene, vertices or accum buffer): re-create the batch");
  if (hipSetDevice(b->device) != hipSuccess) return fail(nullptr, LISA_ERR_HIP, "hipSetDevice");
  hipError_t e;
  if (!b->exec && b->cfg.graph) {  // capture reset + persistent launch once, replay afterwards
    if (b->cfg.form == 2 && (e = b->wf.ensure(wf_batch_slots(b))) != hipSuccess)
      return fail(nullptr, LISA_ERR_ALLOC, std::string("wavefront buffers: ") + hipGetErrorString(e));
    e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t ce = enqueue_batch(b);
      e = hipStreamEndCapture(b->stream, &b->graph);
      if (e == hipSuccess) e = ce;
      if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    }
    if (e == hipSuccess && b->cfg.form == 2) {
      e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        const hipError_t ce = enqueue_batch_rounds(b);
        e = hipStreamEndCapture(b->stream, &b->graph2);
        if (e == hipSuccess) e = ce;
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec2, b->graph2, nullptr, nullptr, 0);
      }
    }
    if (e != hipSuccess) {
      // capture refused (e.g. the caller is capturing this stream itself):
      // plain launches from now on
      (void)hipGetLastError();
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(b->stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
      }
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      if (b->exec2) (void)hipGraphExecDestroy(b->exec2);
      if (b->graph2) (void)hipGraphDestroy(b->graph2);
      b->exec = nullptr;
      b->graph = nullptr;
      b->exec2 = nullptr;
      b->graph2 = nullptr;
      b->cfg.graph = 0;
    }
  }
  if (b->cfg.form == 2 && (e = b->wf.ensure(wf_batch_slots(b))) != hipSuccess)
    return fail(nullptr, LISA_ERR_ALLOC, std::string("wavefront buffers: ") + hipGetErrorString(e));
  for (int32_t i = 0; i < b->n; ++i) {  // each accum buffer will hold a plain render
    b->ctxs[i]->j_seq = false;
    b->ctxs[i]->sum64_valid = false;
  }
  if ((e = hipEventRecord(b->ev0, b->stream)) != hipSuccess)
    return fail(nullptr, LISA_ERR_HIP, std::string("lisa_batch_render: ") + hipGetErrorString(e));
  if (b->exec)
    e = hipGraphLaunch(b->exec, b->stream);
  else
    e = enqueue_batch(b);
  if (e == hipSuccess && b->cfg.form == 2) {  // further blocks of rounds until every path has ended
    uint64_t max_spp_b = 1;
    for (const lisa_render_params& q : b->params)
      max_spp_b = std::max<uint64_t>(max_spp_b, (uint64_t)q.spp * ((uint64_t)q.bounces * (kNeeTries + 1) + 1));
    const uint32_t slots = b->wf.ws.slots;
    const uint64_t max_rounds = max_spp_b * (((uint64_t)b->total + slots - 1) / slots) + 2 * kWfRoundsPerCheck;
    for (uint64_t r = kWfRoundsPerCheck; e == hipSuccess; r += kWfRoundsPerCheck) {
      uint32_t pending = 0;
      if ((e = wf_pending(b->stream, b->wf.ws, &pending)) != hipSuccess || pending == 0) break;
      if (r > max_rounds) return fail(nullptr, LISA_ERR_HIP, "wavefront form: paths did not finish");
      e = b->exec2 ? hipGraphLaunch(b->exec2, b->stream) : enqueue_batch_rounds(b);
    }
  }
  if (e != hipSuccess || (e = hipEventRecord(b->ev1, b->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(b->stream)) != hipSuccess)
    return fail(nullptr, LISA_ERR_HIP, std::string("lisa_batch_render: ") + hipGetErrorString(e));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
  b->last_ms = ms;
  if (elapsed_ms) *elapsed_ms = ms;
  return LISA_OK;
}

lisa_status lisa_batch_autotune(lisa_batch* b, int32_t tune_spp, lisa_batch_config* best_out) {
  if (!b) return fail(nullptr, LISA_ERR_INVALID_ARG, "batch is NULL");
  if (tune_spp <= 0) tune_spp = 2;
  // A reduced-spp copy of the batch stands in for the real work.
  std::vector<lisa_render_params> tp(b->params);
  for (auto& q : tp) q.spp = std::min<uint32_t>(q.spp, (uint32_t)tune_spp);
  lisa_batch* t = nullptr;
  lisa_status rc = lisa_batch_create(&t, b->ctxs.data(), tp.data(), b->n);
  if (rc != LISA_OK) return rc;
  const int32_t per_cu[] = {4, 8, 12, 16, 20, 24, 32};
  const int32_t chunks[] = {1, 4, 16};
  float best = 1e30f;
  lisa_batch_config best_cfg = b->cfg;
  std::vector<lisa_batch_config> cands;
  for (int32_t pc : per_cu) {
    for (int32_t ch : chunks) cands.push_back(lisa_batch_config{b->cus * pc, ch, 0, b->cfg.graph});
    cands.push_back(lisa_batch_config{b->cus * pc, 1, 1, b->cfg.graph});
  }
  // form 2: the traversal kernel's persistent grid and the paths in flight
  // per traversal lane (queue depth)
  for (int32_t pc : {16, 32})
    for (int32_t ch : {1, 2, 4}) cands.push_back(lisa_batch_config{b->cus * pc, ch, 2, b->cfg.graph});
  for (const lisa_batch_config& c : cands) {
    float ms = 0.0f, ms2 = 0.0f;
    if ((rc = lisa_batch_set_config(t, &c)) != LISA_OK || (rc = lisa_batch_render(t, &ms)) != LISA_OK ||
        (rc = lisa_batch_render(t, &ms2)) != LISA_OK) {
      lisa_batch_destroy(t);
      return rc;
    }
    ms = std::min(ms, ms2);
    if (ms < best) {
      best = ms;
      best_cfg = c;
    }
  }
  lisa_batch_destroy(t);
  rc = lisa_batch_set_config(b, &best_cfg);
  if (rc != LISA_OK) return rc;
  if (best_out) *best_out = best_cfg;
  return LISA_OK;
}

// The prefix-path settings of a frame's cache parameters.
lisa_status prefix_args(lisa_ctx* ctx, const lisa_cache_frame_params* cp, PathArgs& a) {
  if (cp->prefix_bounces > 63) return fail(ctx, LISA_ERR_INVALID_ARG, "prefix_bounces must be in 0..63");
  if (cp->nee_loops > 16 || cp->nee_loops_deep > 16)
    return fail(ctx, LISA_ERR_INVALID_ARG, "nee_loops and nee_loops_deep must be in 0..16");
  const bool emit = (cp->flags & LISA_CACHE_EMISSION) != 0;
  if ((cp->nee_loops > 1 || cp->nee_loops_deep > 1) && !emit)
    return fail(ctx, LISA_ERR_INVALID_ARG, "nee_loops > 1 needs LISA_CACHE_EMISSION");
  a.emit_expect = emit ? 1 : 0;
  a.nee_loops = cp->nee_loops > 1 ? (int32_t)cp->nee_loops : 1;
  a.nee_loops_deep = cp->nee_loops_deep > 1 ? (int32_t)cp->nee_loops_deep : 1;
  return LISA_OK;
}

lisa_status lisa_render_cached(lisa_ctx* ctx, lisa_cache* cache, const lisa_render_params* p,
                               const lisa_cache_frame_params* cp, lisa_cache_frame_stats* stats) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!cache || !cp) return fail(ctx, LISA_ERR_INVALID_ARG, "cache / cache params NULL");
  lisa_status rc = check_params(ctx, p);
  if (rc != LISA_OK) return rc;
  if (p->flags & LISA_FLAG_COUNT) return fail(ctx, LISA_ERR_UNSUPPORTED, "no counters in cached renders");
  PathArgs pre{};  // only emit_expect, nee_loops and nee_loops_deep are read
  rc = prefix_args(ctx, cp, pre);
  if (rc != LISA_OK) return rc;
  if (cp->train_paths > (1u << 24)) return fail(ctx, LISA_ERR_INVALID_ARG, "train_paths must be at most 2^24");
  if (cp->train_paths > 0 && (cp->train_bounces == 0 || cp->train_bounces > 64))
    return fail(ctx, LISA_ERR_INVALID_ARG, "train_bounces must be in 1..64");
  if (cp->train_steps > 64) return fail(ctx, LISA_ERR_INVALID_ARG, "train_steps must be at most 64");
  const size_t npx = (size_t)p->width * p->height;
  if (npx * p->spp >= (1ull << 31)) return fail(ctx, LISA_ERR_INVALID_ARG, "width * height * spp >= 2^31");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const auto t_start = std::chrono::steady_clock::now();
  lisa_cache_frame_stats out_stats;
  std::memset(&out_stats, 0, sizeof(out_stats));
  if (cp->judge_pixels > npx || cp->judge_pixels > (1u << 20))
    return fail(ctx, LISA_ERR_INVALID_ARG, "judge_pixels must be at most the pixel count and 2^20");
  if (p->subframe > 0) {
    // frames after the first continue the sequence frame 0 started: same
    // image size and judge pixels (the judge's buffers are sized at frame 0)
    if (ctx->accum_px != npx) return fail(ctx, LISA_ERR_INVALID_ARG, "frame > 0 but the image size changed");
    if (!ctx->j_seq)
      return fail(ctx, LISA_ERR_INVALID_ARG,
                  "frame > 0 but no cached frame 0 started this sequence (the accum buffer was "
                  "written by another call since)");
    if (cp->judge_pixels != ctx->j_nj)
      return fail(ctx, LISA_ERR_INVALID_ARG, "judge_pixels changed within a cached sequence");
  }
  if (ctx->accum_px != npx) {
    rc = ensure_accum(ctx, npx);
    if (rc != LISA_OK) return rc;
  }
  ctx->sum64_valid = false;  // the accum buffer will hold cached frames
  if (p->subframe == 0) {  // a new sequence: the judge starts over
    ctx->j_seq = false;  // set again when this frame succeeds
    ctx->j_nj = cp->judge_pixels;
    ctx->j_from = -1;
    ctx->j_n = 0;
    ctx->j_f = 0;
    const size_t nj = cp->judge_pixels;
    ctx->j_pix.assign(nj, 0);
    ctx->j_sum.assign(3 * nj, 0.0);
    ctx->j_sq.assign(3 * nj, 0.0);
    ctx->j_csum.assign(3 * nj, 0.0);
    ctx->j_csq.assign(3 * nj, 0.0);
    uint64_t sm = cp->seed ^ 0x6A09E667F3BCC909ull;
    for (size_t i = 0; i < nj; ++i) {  // splitmix64
      sm += 0x9E3779B97F4A7C15ull;
      uint64_t z = sm;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      ctx->j_pix[i] = (int32_t)(z % npx);
    }
    if (nj > 0) {
      HIP_TRY(ctx, ctx->c_jpix.ensure(nj));
      HIP_TRY(ctx, ctx->c_jout.ensure(nj));
      HIP_TRY(ctx, hipMemcpy(ctx->c_jpix.p, ctx->j_pix.data(), 4 * nj, hipMemcpyHostToDevice));
    }
  }
  // plain frames from j_from on, once the judge found the cache inaccurate
  const bool taken_over = ctx->j_from >= 0 && (int64_t)p->subframe >= ctx->j_from;
  PathArgs base = cache_args(ctx, p);
  float loss = 0.0f;
  HIP_TRY(ctx, ctx->c_frame.ensure(npx));
  // The frame's prefix paths do not read the network, so without a fallback
  // decision (which waits for this frame's loss) they run on the context's
  // stream while the training steps run on the cache's.
  const bool overlap = !(cp->fallback_loss > 0.0f) && !(cp->flags & LISA_CACHE_SERIAL);
  PathArgs qa = base;
  qa.emit_expect = pre.emit_expect;
  qa.nee_loops = pre.nee_loops;
  qa.nee_loops_deep = pre.nee_loops_deep;
  const bool persistent = (cp->flags & LISA_CACHE_PERSISTENT) != 0;
  const bool full = (cp->flags & LISA_CACHE_FULL_PATHS) != 0;
  bool launched = false;
  // 1. online training on suffixes of full-length paths traced by the plain loop
  if (cp->train_paths > 0 && cp->train_steps > 0 && !taken_over && !full) {
    const auto t0 = std::chrono::steady_clock::now();
    // With the overlap the frame's prefix paths start first, on the
    // context's stream, and the training paths run beside them on a second
    // stream (the two write disjoint buffers; the scene is read-only).
    hipStream_t ts = st;
    if (overlap) {
      if (!ctx->stream2) HIP_TRY(ctx, hipStreamCreate(&ctx->stream2));
      ts = ctx->stream2;
      rc = trace_queries_launch(ctx, qa, nullptr, npx, cp->prefix_bounces, ctx->c_frame.p, persistent);
      if (rc != LISA_OK) return rc;
      launched = true;
    }
    const size_t nt = cp->train_paths;
    const size_t cap = cp->train_max > 0 ? cp->train_max : nt * cp->train_bounces;
    HIP_TRY(ctx, ctx->c_tpix.ensure(nt));
    k_train_pixels<<<grid_for((int64_t)nt, 256), 256, 0, ts>>>(
        ctx->c_tpix.p, (uint32_t)nt, cp->seed * 0x9E3779B97F4A7C15ull + p->subframe + 1, (uint32_t)npx);
    HIP_TRY(ctx, hipGetLastError());
    PathArgs a = base;
    a.spp = 1;
    a.bounces = cp->train_bounces;
    a.subframe = 0x40000000u ^ p->subframe;  // an RNG stream apart from the rendered samples
    uint32_t rows = 0;  // rows at the query depth: targets with the queries' remaining bounces
    rc = trace_train_rows(ctx, a, ctx->c_tpix.p, nt, cap, cp->prefix_bounces + 1, &rows, ts);
    if (rc != LISA_OK) return rc;
    rows = std::min<uint32_t>(rows, (uint32_t)cap);
    const auto t1 = std::chrono::steady_clock::now();
    if (rows > 0) {
      lisa_status r2 = lisa::cache_train_device(cache, ctx->c_tfeat.p, 12, ctx->c_ttgt.p, (int64_t)rows,
                                                &loss, (int)cp->train_steps);
      if (r2 != LISA_OK) return fail(ctx, r2, std::string("cache training: ") + lisa_cache_last_error(cache));
      out_stats.train_kernel_ms = lisa_cache_last_ms(cache);
    }
    const auto t2 = std::chrono::steady_clock::now();
    out_stats.trace_train_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    out_stats.train_ms = std::chrono::duration<float, std::milli>(t2 - t1).count();
    out_stats.train_samples = (int32_t)rows;
  }
  out_stats.loss = loss;
  // 2. this frame: cached paths, or the plain tracer as the fallback
  const bool fallback = taken_over || (!launched && cp->fallback_loss > 0.0f && loss > cp->fallback_loss);
  const auto t3 = std::chrono::steady_clock::now();
  if (fallback) {
    PathArgs a = base;
    a.tiles_x = (p->width + kTileW - 1) / kTileW;
    a.n_items = (int32_t)npx;
    a.accum = ctx->c_frame.p;
    a.accum_prev = nullptr;
    if (cp->flags & LISA_CACHE_PERSISTENT) {
      rc = launch_persistent<0>(ctx, a, 3);
      if (rc != LISA_OK) return rc;
    } else {
      const unsigned tiles = ((p->width + kTileW - 1) / kTileW) * ((p->height + kTileH - 1) / kTileH);
      if (ctx->all.wide)
        k_path<false, false, false, true><<<tiles, kPathBlock, 0, st>>>(view(ctx), a);
      else
        k_path<false, false><<<tiles, kPathBlock, 0, st>>>(view(ctx), a);
      HIP_TRY(ctx, hipGetLastError());
    }
    HIP_TRY(ctx, hipStreamSynchronize(st));
    out_stats.fell_back = 1;
  } else {
    uint32_t nq = 0;
    if (!launched) {
      rc = trace_queries_launch(ctx, qa, nullptr, npx, cp->prefix_bounces, ctx->c_frame.p, persistent, full);
      if (rc != LISA_OK) return rc;
    }
    rc = trace_queries_finish(ctx, qa, &nq);
    if (rc != LISA_OK) return rc;
    nq = std::min<uint32_t>(nq, (uint32_t)(npx * p->spp));
    const auto t4 = std::chrono::steady_clock::now();
    if (nq > 0) {
      lisa_status r2 = lisa::cache_infer_device(cache, ctx->c_qfeat.p, nq, ctx->c_qout.p);
      if (r2 != LISA_OK) return fail(ctx, r2, std::string("cache inference: ") + lisa_cache_last_error(cache));
      out_stats.infer_kernel_ms = lisa_cache_last_ms(cache);
      k_cache_add<<<(nq + 255) / 256, 256, 0, st>>>(ctx->c_qout.p, ctx->c_qatt.p, ctx->c_qitem.p, nq,
                                                    1.0f / (float)p->spp, ctx->c_frame.p);
      HIP_TRY(ctx, hipGetLastError());
    }
    HIP_TRY(ctx, hipStreamSynchronize(st));
    out_stats.infer_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t4).count();
    out_stats.queries = (int32_t)nq;
  }
  out_stats.render_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t3).count();
  // 3. running mean over frames
  // (after a takeover the mean restarts with the first plain frame)
  const uint32_t mean_from = taken_over ? (uint32_t)ctx->j_from : 0u;
  k_blend<<<grid_for((int64_t)npx, 256), 256, 0, st>>>(ctx->d_accum, ctx->c_frame.p, (int64_t)npx,
                                                      1.0f / (float)(p->subframe - mean_from + 1),
                                                      p->subframe == mean_from);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipStreamSynchronize(st));
  out_stats.judge_plain = taken_over ? 1 : 0;
  out_stats.judge_rmse_est = -1.0f;
  if (cp->judge_pixels > 0 && !taken_over) {
    // 4. the plain tracer judges: its samples on the judge pixels (an RNG
    // stream apart) give their mean m, the cached frames' values there in
    // the current block of kJudgeBlock frames their mean c (so the frames of
    // a network still training drop out); bias^2 = mean over pixels and
    // channels of (c - m)^2 - var(m) - var(c).  From the third block on,
    // with >= 4 frames in the block, past the bound: plain frames from the
    // next frame on.
    constexpr uint32_t kJudgeBlock = 8;
    if (p->subframe % kJudgeBlock == 0) {
      std::fill(ctx->j_csum.begin(), ctx->j_csum.end(), 0.0);
      std::fill(ctx->j_csq.begin(), ctx->j_csq.end(), 0.0);
      ctx->j_f = 0;
    }
    const int32_t nj = (int32_t)cp->judge_pixels;
    lisa_render_params jp = *p;
    jp.subframe = p->subframe ^ 0x20000000u;
    jp.flags = 0;
    const size_t ns = (size_t)nj * p->spp;
    HIP_TRY(ctx, ctx->c_jacc.ensure((size_t)nj));
    HIP_TRY(ctx, ctx->c_jrgb.ensure(3 * ns));
    PathArgs ja = make_args(&jp);  // the plain loop on the judge pixels, per-sample radiance kept
    ja.pixels = ctx->c_jpix.p;
    ja.n_items = nj;
    ja.accum = ctx->c_jacc.p;
    ja.accum_prev = nullptr;
    ja.s_rgb = ctx->c_jrgb.p;
    k_path<true, false><<<grid_for(nj, kPathBlock), kPathBlock, 0, st>>>(view(ctx), ja);
    HIP_TRY(ctx, hipGetLastError());
    std::vector<float> jrgb(3 * ns);
    HIP_TRY(ctx, hipMemcpyAsync(jrgb.data(), ctx->c_jrgb.p, 4 * 3 * ns, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    for (int32_t i = 0; i < nj; ++i)
      for (uint32_t s = 0; s < p->spp; ++s)
        for (int c = 0; c < 3; ++c) {
          const double x = jrgb[3 * ((size_t)i * p->spp + s) + c];
          ctx->j_sum[3 * (size_t)i + c] += x;
          ctx->j_sq[3 * (size_t)i + c] += x * x;
        }
    ctx->j_n += p->spp;
    k_pick<<<(nj + 255) / 256, 256, 0, st>>>(ctx->c_frame.p, ctx->c_jpix.p, nj, ctx->c_jout.p);
    HIP_TRY(ctx, hipGetLastError());
    std::vector<float> fv(4 * (size_t)nj);
    HIP_TRY(ctx, hipMemcpy(fv.data(), ctx->c_jout.p, 16 * (size_t)nj, hipMemcpyDeviceToHost));
    for (int32_t i = 0; i < nj; ++i)
      for (int c = 0; c < 3; ++c) {
        const double x = fv[4 * (size_t)i + c];
        ctx->j_csum[3 * (size_t)i + c] += x;
        ctx->j_csq[3 * (size_t)i + c] += x * x;
      }
    ctx->j_f += 1;
    const double n = (double)ctx->j_n, nf = (double)ctx->j_f;
    auto var_of_mean = [](double s, double q, double k) {
      if (k < 2.0) return 0.0;
      const double m = s / k;
      return (q / k - m * m) / (k - 1.0);  // sample variance / k
    };
    double e2 = 0.0;
    for (size_t k = 0; k < 3 * (size_t)nj; ++k) {
      const double d = ctx->j_csum[k] / nf - ctx->j_sum[k] / n;
      e2 += d * d - var_of_mean(ctx->j_sum[k], ctx->j_sq[k], n) -
            var_of_mean(ctx->j_csum[k], ctx->j_csq[k], nf);
    }
    e2 /= 3.0 * nj;
    out_stats.judge_rmse_est = (float)std::sqrt(std::max(e2, 0.0));
    if (p->subframe >= 2 * kJudgeBlock && ctx->j_f >= 4 && ctx->j_n >= 16 &&
        out_stats.judge_rmse_est > cp->judge_rmse)
      ctx->j_from = (int64_t)p->subframe + 1;
  }
  out_stats.total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (p->subframe == 0) ctx->j_seq = true;
  if (stats) *stats = out_stats;
  return LISA_OK;
}

namespace {
// Host copy of a pixel list, checked against the image size.
lisa_status upload_pixels(lisa_ctx* ctx, const lisa_render_params* p, const int32_t* pixels, int32_t n) {
  if (n <= 0 || !pixels) return fail(ctx, LISA_ERR_INVALID_ARG, "pixels required, n > 0");
  const int64_t npx = (int64_t)p->width * p->height;
  for (int32_t i = 0; i < n; ++i)
    if (pixels[i] < 0 || pixels[i] >= npx) return fail(ctx, LISA_ERR_INVALID_ARG, "pixel id out of range");
  HIP_TRY(ctx, ctx->c_tpix.ensure((size_t)n));
  HIP_TRY(ctx, hipMemcpy(ctx->c_tpix.p, pixels, 4 * (size_t)n, hipMemcpyHostToDevice));
  return LISA_OK;
}
}  // namespace

lisa_status lisa_cache_trace_queries(lisa_ctx* ctx, const lisa_render_params* p,
                                     const lisa_cache_frame_params* cp, const int32_t* pixels, int32_t n,
                                     float* accum_out, float* q_feat, float* q_att, int32_t* q_item,
                                     int32_t q_cap, int32_t* n_queries) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (!cp) return fail(ctx, LISA_ERR_INVALID_ARG, "cache params NULL");
  lisa_status rc = check_params(ctx, p);
  if (rc != LISA_OK) return rc;
  PathArgs pre{};  // only emit_expect, nee_loops and nee_loops_deep are read
  rc = prefix_args(ctx, cp, pre);
  if (rc != LISA_OK) return rc;
  if (!accum_out || !n_queries || q_cap < 0 || (q_cap > 0 && (!q_feat || !q_att || !q_item)))
    return fail(ctx, LISA_ERR_INVALID_ARG, "output arrays");
  if ((size_t)n * p->spp >= (1ull << 31)) return fail(ctx, LISA_ERR_INVALID_ARG, "n * spp >= 2^31");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  rc = upload_pixels(ctx, p, pixels, n);
  if (rc != LISA_OK) return rc;
  HIP_TRY(ctx, ctx->c_frame.ensure((size_t)n));
  PathArgs a = cache_args(ctx, p);
  a.emit_expect = pre.emit_expect;
  a.nee_loops = pre.nee_loops;
  a.nee_loops_deep = pre.nee_loops_deep;
  uint32_t nq = 0;
  rc = trace_queries(ctx, a, ctx->c_tpix.p, (size_t)n, cp->prefix_bounces, ctx->c_frame.p, &nq, false,
                     (cp->flags & LISA_CACHE_FULL_PATHS) != 0);
  if (rc != LISA_OK) return rc;
  HIP_TRY(ctx, hipMemcpy(accum_out, ctx->c_frame.p, sizeof(float4) * n, hipMemcpyDeviceToHost));
  const size_t k = std::min<size_t>(nq, (size_t)q_cap);
  if (k > 0) {
    HIP_TRY(ctx, hipMemcpy(q_feat, ctx->c_qfeat.p, 48 * k, hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(q_att, ctx->c_qatt.p, 12 * k, hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(q_item, ctx->c_qitem.p, 4 * k, hipMemcpyDeviceToHost));
  }
  *n_queries = (int32_t)nq;
  return LISA_OK;
}

lisa_status lisa_cache_trace_rows(lisa_ctx* ctx, const lisa_render_params* p, const int32_t* pixels,
                                  int32_t n, uint32_t depth, float* feat, float* tgt, int32_t cap,
                                  int32_t* n_rows) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  lisa_status rc = check_params(ctx, p);
  if (rc != LISA_OK) return rc;
  if (p->bounces == 0 || p->bounces > 64) return fail(ctx, LISA_ERR_INVALID_ARG, "bounces must be in 1..64");
  if (!n_rows || cap <= 0 || !feat || !tgt) return fail(ctx, LISA_ERR_INVALID_ARG, "output arrays");
  if ((size_t)n * p->spp * (depth > 0 ? 1 : p->bounces) >= (1ull << 31))
    return fail(ctx, LISA_ERR_INVALID_ARG, "n * spp * row slots per sample >= 2^31");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  rc = upload_pixels(ctx, p, pixels, n);
  if (rc != LISA_OK) return rc;
  PathArgs a = cache_args(ctx, p);
  uint32_t rows = 0;
  rc = trace_train_rows(ctx, a, ctx->c_tpix.p, (size_t)n, (size_t)cap, depth, &rows, ctx->stream);
  if (rc != LISA_OK) return rc;
  const size_t k = std::min<size_t>(rows, (size_t)cap);
  if (k > 0) {
    HIP_TRY(ctx, hipMemcpy(feat, ctx->c_tfeat.p, 48 * k, hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(tgt, ctx->c_ttgt.p, 12 * k, hipMemcpyDeviceToHost));
  }
  *n_rows = (int32_t)rows;
  return LISA_OK;
}

lisa_status lisa_synchronize(lisa_ctx* ctx) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LISA_OK;
}

lisa_status lisa_set_wide_threshold(lisa_ctx* ctx, int64_t min_triangles) {
  if (!ctx) return fail(nullptr, LISA_ERR_INVALID_ARG, "ctx is NULL");
  if (min_triangles < 0) return fail(ctx, LISA_ERR_INVALID_ARG, "min_triangles < 0");
  ctx->wide_min_tris = min_triangles;
  return LISA_OK;
}

lisa_status lisa_get_stats(const lisa_ctx* ctx, lisa_stats* out) {
  if (!ctx || !out) return fail(nullptr, LISA_ERR_INVALID_ARG, "NULL argument");
  out->render_ms = ctx->render_ms;
  out->bvh_build_ms = ctx->bvh_ms;
  out->num_tris = ctx->T;
  out->num_nodes = ctx->all.nI;
  out->num_emitters = ctx->emit.T;
  out->device_bytes = ctx->scene_bytes + ctx->accum_bytes;
  out->wide_tree = ctx->all.wide ? 1 : 0;
  out->reserved = 0;
  return LISA_OK;
}

lisa_status lisa_get_counters(const lisa_ctx* ctx, uint64_t* counters) {
  if (!ctx || !counters) return fail(nullptr, LISA_ERR_INVALID_ARG, "NULL argument");
  for (int i = 0; i < LISA_NCOUNTERS; ++i) counters[i] = ctx->counters[i];
  return LISA_OK;
}

void* lisa_get_stream(const lisa_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

}  // extern "C"
