// CPU oracle for the LiSA path-tracing hot path.  TEST INFRASTRUCTURE ONLY
// (see lisa_oracle.h): the product library never links or calls this file.
//
// Built with -O3 -ffp-contract=off and no fast-math, so every FP32 operation
// below is one IEEE-rounded operation in the order written.  The HIP kernels
// are written to the same operation order; parity tests compare the two.
#include "lisa_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// float3 algebra, op order as src/sutil/vec_math.h (dot 521, cross 527,
// normalize 539-543, reflect 552-555, faceforward 561-564, lerp 494-497,
// clamp 113, float3 / s 477-481).
// ---------------------------------------------------------------------------
struct V3 {
  float x, y, z;
};
inline V3 mk(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return mk(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return mk(a.x * s, a.y * s, a.z * s); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline V3 normalize(V3 v) {
  const float inv = 1.0f / std::sqrt(dot(v, v));
  return v * inv;
}
inline V3 reflect(V3 i, V3 n) { return i - (2.0f * n) * dot(n, i); }
inline V3 faceforward(V3 n, V3 i, V3 nref) { return n * std::copysign(1.0f, dot(i, nref)); }
inline V3 lerp(V3 a, V3 b, float t) { return a + t * (b - a); }
inline float clampf(float f, float a, float b) { return std::fmax(a, std::fmin(f, b)); }
inline float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
inline V3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
inline void st3(float* p, V3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

// ---------------------------------------------------------------------------
// RNG: src/cuda/random.h:31-46 (tea<16>: TEA, Wheeler & Needham 1994, as a
// seed hash per Zafar, Olano & Curtis, HPG 2010), :49-55 (lcg, Numerical
// Recipes constants), :64-67 (rnd);  src/LiSA/src/maths.cu:6-8 (rng).
// ---------------------------------------------------------------------------
inline uint32_t tea16(uint32_t v0, uint32_t v1) {
  uint32_t s0 = 0;
  for (int n = 0; n < 16; ++n) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
inline uint32_t lcg(uint32_t& s) {
  s = 1664525u * s + 1013904223u;
  return s & 0x00FFFFFFu;
}
inline float rnd(uint32_t& s) { return (float)lcg(s) / (float)0x01000000; }
inline float rng(uint32_t& s) { return rnd(s) * 2.0f - 1.0f; }

// maths.cu:10-15 -- x, then y, then z.
// in two steps: the three draws, then the normalised direction turned to
// the side of N (the NEE loop culls on the raw draws first).
inline V3 hemisphere_raw(uint32_t& s) {
  const float x = rng(s);
  const float y = rng(s);
  const float z = rng(s);
  return mk(x, y, z);
}
inline V3 hemisphere_dir(V3 raw, V3 N) {
  const V3 r = normalize(raw);
  return faceforward(r, N, r);
}
inline V3 hemisphere(V3 N, uint32_t& s) { return hemisphere_dir(hemisphere_raw(s), N); }
// maths.cu:17-20 (Schlick); pow(r,2) = r*r, pow(x,5) = ((x*x)*(x*x))*x.
inline float fresnel(float cosT, float eta) {
  const float r = (1.0f - eta) / (1.0f + eta);
  const float R0 = r * r;
  const float x = 1.0f - cosT;
  const float x2 = x * x;
  const float x5 = (x2 * x2) * x;
  return R0 + (1.0f - R0) * x5;
}
// maths.cu:22-30: zero vector on total internal reflection.
inline V3 refract(float cosI, V3 d, V3 N, float eta) {
  const float cost2 = 1.0f - eta * eta * (1.0f - cosI * cosI);
  const V3 t = eta * d + ((eta * cosI - std::sqrt(std::fabs(cost2))) * N);
  return t * (float)(cost2 > 0.0f);
}
// maths.cu:33-56.
inline V3 bary_normal(V3 P, V3 n1, V3 n2, V3 n3, V3 v1, V3 v2, V3 v3) {
  const V3 e1 = v2 - v1;
  const V3 e2 = v3 - v1;
  const V3 i = P - v1;
  const float d00 = dot(e1, e1);
  const float d01 = dot(e1, e2);
  const float d11 = dot(e2, e2);
  const float d20 = dot(i, e1);
  const float d21 = dot(i, e2);
  const float denom = d00 * d11 - d01 * d01;
  const float w = (d00 * d21 - d01 * d20) / denom;
  const float v = (d11 * d20 - d01 * d21) / denom;
  const float u = 1.0f - v - w;
  return normalize(u * n1 + v * n2 + w * n3);
}
// lambertian.cu:7-13 (sample), :15-22 (eval: (double)c / M_PI then float), :24-27.
inline V3 bsdf_bounce(V3 d, V3 N, uint32_t& s, float roughness) {
  const V3 r = reflect(d, N);
  const V3 h = hemisphere(N, s);
  return lerp(r, h, roughness);
}
// Density per steradian with which bsdf_bounce(d, N, ., roughness) returns
// the unit direction L (0 for roughness 0, a mirror).  The hemisphere draw h
// is a cube-uniform point normalised and folded to N's side, density
// q(h) = 1 / (12 max|h_i|^3); the bounce's direction is that of
// v = (1 - t) r + t h (t = roughness, r = reflect(d, N)), i.e. of a point on
// a unit sphere about c = (1 - t) r / t seen from the origin, so each
// preimage h of L adds q(h) (s / t)^2 / |h . L| with v = s L.  (This
// implementation's; lambertian.cu has no pdf.)
inline float bsdf_pdf(V3 d, V3 N, V3 L, float roughness) {
  if (!(roughness > 0.0f && roughness <= 1.0f)) return 0.0f;
  if (roughness == 1.0f) {  // h = L
    if (!(dot(N, L) > 0.0f)) return 0.0f;
    const float mw = std::fmax(std::fabs(L.x), std::fmax(std::fabs(L.y), std::fabs(L.z)));
    return 1.0f / (12.0f * mw * mw * mw);
  }
  const float t = roughness, u = 1.0f - roughness;
  const V3 r = reflect(d, N);
  const float a = u * dot(L, r);
  const float disc = a * a - u * u + t * t;
  if (disc < 0.0f) return 0.0f;
  const float sq = std::sqrt(disc);
  float p = 0.0f;
  for (int k = 0; k < 2; ++k) {
    const float s = k == 0 ? a + sq : a - sq;
    if (!(s > 0.0f)) continue;
    const V3 h = (s * L - u * r) * (1.0f / t);
    const float hl = std::fabs(dot(h, L));
    if (!(dot(h, N) > 0.0f) || !(hl > 0.0f)) continue;
    const float mh = std::fmax(std::fabs(h.x), std::fmax(std::fabs(h.y), std::fabs(h.z)));
    p += (1.0f / (12.0f * mh * mh * mh)) * ((s / t) * (s / t)) / hl;
    if (sq == 0.0f) break;
  }
  return p;
}
inline float bsdf_eval(V3 N, V3 L) {
  const float NdotL = clampf(dot(N, L), 0.0f, 1.0f);
  const float p = (float)((double)NdotL / 3.14159265358979323846);
  return NdotL * p;
}

// src/cuda/helpers.h:35-66 on the host.
inline float to_srgb(float c) {
  const float p = std::pow(c, 1.0f / 2.4f);
  return c < 0.0031308f ? 12.92f * c : 1.055f * p - 0.055f;
}
inline uint8_t quant8(float x) {
  x = clampf(x, 0.0f, 1.0f);
  const unsigned v = (unsigned)(x * 256.0f);
  return (uint8_t)std::min(v, 255u);
}

// ---------------------------------------------------------------------------
// LBVH (Karras, "Maximizing Parallelism in the Construction of BVHs,
// Octrees, and k-d Trees", HPG 2012) over the triangle soup.
// ---------------------------------------------------------------------------
constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
// 1 + 2*gamma(3) (Ize, "Robust BVH Ray Traversal", JCGT 2013): widens the
// far slab distance so rounding never culls a box the ray touches.
constexpr float kRobust = 0x1.000006p+0f;
// Boxes are culled against the best hit distance widened by 2^-16: a
// triangle whose hit lies on its box boundary (axis-aligned or coplanar
// geometry) can compute a t a few ulp below the box's slab entry, and must
// still be tested so that equal-t hits resolve to the lower triangle id.
constexpr float kWiden = 0x1.0001p+0f;

// Spread the low 10 bits of v to every third bit (bit-by-bit form).
inline uint32_t spread10(uint32_t v) {
  uint32_t r = 0;
  for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}
inline uint32_t quant10(float u) {
  float q = u * 1024.0f;
  q = std::fmin(std::fmax(q, 0.0f), 1023.0f);
  return (uint32_t)q;
}
inline uint32_t morton3(float x, float y, float z) {
  return (spread10(quant10(x)) << 2) | (spread10(quant10(y)) << 1) | spread10(quant10(z));
}

// One LBVH over a subset of the soup's triangles (all of them, or the
// emitters).  prim holds global triangle ids in Morton order.
struct Bvh {
  int32_t T = 0;
  std::vector<uint32_t> morton, prim;  // sorted
  std::vector<uint32_t> refs;          // 2 per internal node
  std::vector<float> boxes;            // 12 per internal node
  std::vector<float> tv;               // 9 per sorted triangle (traversal copy)
};

// Division-free cull of NEE's emitter query (this implementation's, not the
// reference's): a box around every emitter, centre c and half extents H,
// inflated by twice a margin m = S * 2^-12 per side, where S is the largest
// |coordinate| of the scene box.  A ray whose origin lies in the scene box
// and that misses the inflated box by a separating axis test evaluated in
// FP32 misses the box inflated by m in exact arithmetic; the slab tests of
// the emitter root's two child boxes (each inside the box) then miss in FP32
// too, since their rounding error is below 16 S u / |D_i| (u = 2^-24) per
// t value against a gap of at least m / |D_i| + m / |D_j| -- so the query
// answers "no emitter" exactly as the traversal would, without the ray's
// divisions.  Counted as no node visit.
struct EmitCull {
  float c[3] = {0.0f, 0.0f, 0.0f};
  float H[3] = {INFINITY, INFINITY, INFINITY};  // never culls
};

struct Scene {
  int32_t T = 0;
  std::vector<float> v, n;     // 9T each, original order
  std::vector<int32_t> mat;    // T
  std::vector<orc_material> mats;
  Bvh all;   // every triangle
  Bvh emit;  // the triangles whose material emits (light queries of NEE)
  EmitCull ecull;
  EmitCull ebox;  // the same box, never switched off (the light sample's bound)
  // area prefix sums of the emitting triangles in emitter-tree order, and the
  // total (the light sample of the cache's prefix-emission estimator)
  std::vector<float> ecdf;
  float earea = 0.0f;
};

// The emitter area prefix sums in tree order, in FP32, in the order the
// device sums them: chunks of kCdfChunk triangles are summed sequentially
// from 0 (p_k), the chunk totals sequentially into bases b_c, and
// cdf[k] = b_c + p_k (for one chunk: the plain sequential sums).
constexpr int32_t kCdfChunk = 64;
void emit_cdf(Scene& s) {
  const int32_t T = s.emit.T;
  s.ecdf.assign((size_t)T, 0.0f);
  float base = 0.0f;
  for (int32_t k0 = 0; k0 < T; k0 += kCdfChunk) {
    float acc = 0.0f;
    for (int32_t k = k0; k < std::min(k0 + kCdfChunk, T); ++k) {
      const float* t = &s.emit.tv[9 * (size_t)k];
      const V3 v0 = ld3(t), v1 = ld3(t + 3), v2 = ld3(t + 6);
      const V3 c = cross(v1 - v0, v2 - v0);
      acc = acc + 0.5f * std::sqrt(dot(c, c));
      s.ecdf[(size_t)k] = base + acc;
    }
    base = base + acc;
  }
  s.earea = T > 0 ? s.ecdf[(size_t)T - 1] : 0.0f;
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

// Leaf boxes are the triangle bounds widened outward by one ulp: a ray with
// a zero direction component whose origin lies exactly on a face plane of
// the tight box would evaluate 0 * inf = NaN in the slab test; with the
// widened box such an origin lies strictly inside, and an origin on the
// widened plane is outside all of the box's geometry.
void leaf_box(const Bvh& s, int64_t k, float* b) {
  const float* t = &s.tv[9 * k];
  for (int a = 0; a < 3; ++a) {
    b[a] = std::nextafter(std::fmin(std::fmin(t[a], t[3 + a]), t[6 + a]), -INFINITY);
    b[3 + a] = std::nextafter(std::fmax(std::fmax(t[a], t[3 + a]), t[6 + a]), INFINITY);
  }
}

// Refit: every child box from the current triangle copies (tv), post-order,
// child boxes stored in the parent.  Topology (refs) is unchanged.
void refit(Bvh& s) {
  if (s.T == 0) return;
  if (s.T == 1) {
    leaf_box(s, 0, &s.boxes[0]);
    return;
  }
  struct Frame {
    uint32_t node;
    int stage;
  };
  std::vector<Frame> st;
  st.push_back({0, 0});
  while (!st.empty()) {
    Frame& f = st.back();
    const uint32_t nd = f.node;
    if (f.stage < 2) {
      const uint32_t r = s.refs[2 * nd + f.stage];
      f.stage++;
      if (!(r & kLeaf)) st.push_back({r, 0});
      continue;
    }
    for (int ch = 0; ch < 2; ++ch) {
      const uint32_t r = s.refs[2 * nd + ch];
      float* b = &s.boxes[12 * nd + 6 * ch];
      if (r & kLeaf) {
        leaf_box(s, r & ~kLeaf, b);
      } else {
        const float* cb = &s.boxes[12 * r];
        for (int a = 0; a < 3; ++a) {
          b[a] = std::fmin(cb[a], cb[6 + a]);
          b[3 + a] = std::fmax(cb[3 + a], cb[9 + a]);
        }
      }
    }
    st.pop_back();
  }
}

// New vertex positions for the same triangles: refresh the triangle copies
// in Morton order and refit; Morton codes and topology stay those of the build.
void update_bvh(Bvh& s, const std::vector<float>& v) {
  for (int64_t k = 0; k < s.T; ++k)
    std::memcpy(&s.tv[9 * k], &v[9 * (size_t)s.prim[k]], 9 * sizeof(float));
  refit(s);
}

// Builds `s` over the triangles gid[0..n) of the soup `v` (9 floats each).
void build_bvh(Bvh& s, const std::vector<float>& v, const std::vector<uint32_t>& gid) {
  const int64_t n = (int64_t)gid.size();
  s.T = (int32_t)n;
  if (n == 0) return;
  // Centroids and their bounds.
  std::vector<float> c(3 * n);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t k = 0; k < n; ++k) {
    const float* t = &v[9 * (size_t)gid[k]];
    for (int a = 0; a < 3; ++a) {
      const float ca = (t[a] + t[3 + a] + t[6 + a]) / 3.0f;
      c[3 * k + a] = ca;
      lo[a] = std::fmin(lo[a], ca);
      hi[a] = std::fmax(hi[a], ca);
    }
  }
  std::vector<uint32_t> code(n);
  for (int64_t k = 0; k < n; ++k) {
    float u[3];
    for (int a = 0; a < 3; ++a) {
      const float ext = hi[a] - lo[a];
      u[a] = ext > 0.0f ? (c[3 * k + a] - lo[a]) / ext : 0.5f;
    }
    code[k] = morton3(u[0], u[1], u[2]);
  }
  std::vector<uint32_t> order(n);
  for (int64_t k = 0; k < n; ++k) order[k] = (uint32_t)k;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
  s.morton.resize(n);
  s.prim.resize(n);
  s.tv.resize(9 * n);
  for (int64_t k = 0; k < n; ++k) {
    s.morton[k] = code[order[k]];
    s.prim[k] = gid[order[k]];
    std::memcpy(&s.tv[9 * k], &v[9 * (size_t)s.prim[k]], 9 * sizeof(float));
  }

  const int64_t nI = std::max<int64_t>(n - 1, 1);
  s.refs.assign(2 * nI, kEmpty);
  s.boxes.assign(12 * nI, 0.0f);
  if (n == 1) {
    s.refs[0] = kLeaf | 0u;
    refit(s);
    return;
  }
  // delta(i, j): common prefix length of index-augmented keys, -1 outside.
  auto delta = [&](int64_t i, int64_t j) -> int {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = s.morton[i], b = s.morton[j];
    if (a != b) return clz32(a ^ b);
    return 32 + clz32((uint32_t)i ^ (uint32_t)j);
  };
  for (int64_t i = 0; i < n - 1; ++i) {
    const int d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(i, i - d);
    int64_t lmax = 2;
    while (delta(i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
      if (delta(i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(i, j);
    int64_t sp = 0, t = 0, div = 2;
    do {
      t = (l + div - 1) / div;
      if (delta(i, i + (sp + t) * d) > dnode) sp += t;
      div *= 2;
    } while (t > 1);
    const int64_t g = i + sp * d + std::min(d, 0);
    s.refs[2 * i + 0] = (std::min(i, j) == g) ? (kLeaf | (uint32_t)g) : (uint32_t)g;
    s.refs[2 * i + 1] = (std::max(i, j) == g + 1) ? (kLeaf | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
  }
  refit(s);
}

// ---------------------------------------------------------------------------
// Traversal + watertight ray/triangle test (Woop, Benthin & Wald, JCGT 2013).
// ---------------------------------------------------------------------------
struct Ray {
  V3 O, D, inv;
  int kx, ky, kz;
  float Sx, Sy, Sz;
  bool valid;
};

inline Ray make_ray(V3 O, V3 D) {
  Ray r;
  r.O = O;
  r.D = D;
  r.valid = std::isfinite(D.x) && std::isfinite(D.y) && std::isfinite(D.z) &&
            (D.x != 0.0f || D.y != 0.0f || D.z != 0.0f);
  r.inv = mk(1.0f / D.x, 1.0f / D.y, 1.0f / D.z);
  const float ax = std::fabs(D.x), ay = std::fabs(D.y), az = std::fabs(D.z);
  int kz = 0;
  float m = ax;
  if (ay > m) {
    kz = 1;
    m = ay;
  }
  if (az > m) kz = 2;
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  if (comp(D, kz) < 0.0f) std::swap(kx, ky);
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.Sx = comp(D, kx) / comp(D, kz);
  r.Sy = comp(D, ky) / comp(D, kz);
  r.Sz = 1.0f / comp(D, kz);
  return r;
}

inline bool watertight(const Ray& r, const float* tri, float& t) {
  const V3 A = ld3(tri) - r.O, B = ld3(tri + 3) - r.O, C = ld3(tri + 6) - r.O;
  const float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  const float Ax = comp(A, r.kx) - r.Sx * Akz;
  const float Ay = comp(A, r.ky) - r.Sy * Akz;
  const float Bx = comp(B, r.kx) - r.Sx * Bkz;
  const float By = comp(B, r.ky) - r.Sy * Bkz;
  const float Cx = comp(C, r.kx) - r.Sx * Ckz;
  const float Cy = comp(C, r.ky) - r.Sy * Ckz;
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

inline bool slab(const float* b, const Ray& r, float tmin, float tmax, float& tn_out) {
  const float tx0 = (b[0] - r.O.x) * r.inv.x, tx1 = (b[3] - r.O.x) * r.inv.x;
  const float ty0 = (b[1] - r.O.y) * r.inv.y, ty1 = (b[4] - r.O.y) * r.inv.y;
  const float tz0 = (b[2] - r.O.z) * r.inv.z, tz1 = (b[5] - r.O.z) * r.inv.z;
  const float tn = std::fmax(std::fmax(std::fmin(tx0, tx1), std::fmin(ty0, ty1)),
                             std::fmax(std::fmin(tz0, tz1), tmin));
  const float tfb = std::fmin(std::fmin(std::fmax(tx0, tx1), std::fmax(ty0, ty1)),
                              std::fmax(tz0, tz1)) * kRobust;
  const float tf = std::fmin(tfb, tmax);
  tn_out = tn;
  return tn <= tf;
}

struct Counters {
  uint64_t c[ORC_NCOUNTERS] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Closest hit (kAny = false) or any hit (kAny = true) over `b`: returns the
// global triangle id (-1 when none).  A triangle is accepted when
// t > tmin and (t < best_t or (t == best_t and id < best)), starting from
// best_t = tlim, best = lim_prim; the closest hit therefore resolves equal t
// to the lower triangle id, and the any-hit query answers "is any triangle
// strictly before (tlim, lim_prim) in that order".
template <bool kAny>
int32_t closest(const Bvh& b, const Ray& r, float tmin, float tlim, uint32_t lim_prim,
                float& t_out, Counters& cnt) {
  if (!r.valid || b.T == 0) return -1;
  float best_t = tlim;
  float cull_t = tlim * kWiden;
  uint32_t best = lim_prim;
  bool found = false;
  uint32_t stack[64];
  int sp = 0;
  uint32_t node = 0;
  auto tri_test = [&](uint32_t k) {
    cnt.c[3]++;
    float t;
    const uint32_t p = b.prim[k];
    if (watertight(r, &b.tv[9 * (size_t)k], t) && t > tmin &&
        (t < best_t || (t == best_t && p < best))) {
      best_t = t;
      cull_t = t * kWiden;
      best = p;
      found = true;
    }
  };
  for (;;) {
    cnt.c[2]++;
    const float* bx = &b.boxes[12 * (size_t)node];
    const uint32_t r0 = b.refs[2 * (size_t)node], r1 = b.refs[2 * (size_t)node + 1];
    float tn0 = 0.0f, tn1 = 0.0f;
    bool h0 = r0 != kEmpty && slab(bx, r, tmin, cull_t, tn0);
    bool h1 = r1 != kEmpty && slab(bx + 6, r, tmin, cull_t, tn1);
    if (h0 && (r0 & kLeaf)) {
      tri_test(r0 & ~kLeaf);
      h0 = false;
    }
    if (h1 && (r1 & kLeaf)) {
      tri_test(r1 & ~kLeaf);
      h1 = false;
    }
    if (kAny && found) break;
    if (h0 && h1) {
      uint32_t nr = r0, fr = r1;
      if (tn1 < tn0) {
        nr = r1;
        fr = r0;
      }
      stack[sp++] = fr;
      node = nr;
    } else if (h0) {
      node = r0;
    } else if (h1) {
      node = r1;
    } else {
      if (sp == 0) break;
      node = stack[--sp];
    }
  }
  if (!found) return -1;
  t_out = best_t;
  return (int32_t)best;
}

int32_t closest_brute(const Scene& s, const Ray& r, float tmin, float tmax, float& t_out) {
  if (!r.valid || s.T == 0) return -1;
  float best_t = tmax;
  uint32_t best = 0xFFFFFFFFu;
  for (int32_t p = 0; p < s.T; ++p) {
    float t;
    if (watertight(r, &s.v[9 * (size_t)p], t) && t > tmin &&
        (t < best_t || (t == best_t && (uint32_t)p < best))) {
      best_t = t;
      best = (uint32_t)p;
    }
  }
  if (best == 0xFFFFFFFFu) return -1;
  t_out = best_t;
  return (int32_t)best;
}

// ---------------------------------------------------------------------------
// Path loop: src/LiSA/src/shader.cu.
// ---------------------------------------------------------------------------
constexpr float kTmin = 1e-4f;  // shader.cu:114, :202
constexpr float kTmax = 1e16f;  // shader.cu:115, :202
constexpr int kNeeTries = 30;   // shader.cu:199
constexpr int kMaxRedraws = 64;  // conditioned bounces drawn again at most this often

struct RayState {  // shader.cu:30-42
  int32_t light_mat = -1;
  V3 normal{0, 0, 0};
  V3 origin{0, 0, 0};
  V3 att{1.0f, 1.0f, 1.0f};
  V3 color{0.0f, 0.0f, 0.0f};
  V3 direction{0, 0, 0};
  bool hit = false;
  bool done = false;
};

// shader.cu:196-209 with the occlusion programs 172-184.  `occlude(dir, hit)`
// returns the material id of the closest hit or -1 on a miss (given the
// current hit flag it may answer a non-emitting id for a miss when the two
// have the same effect).  The hit flag lives in RayState and is only
// cleared by a miss (Q1: after one success, a later NEE returns on its first
// non-missing ray with the stale light).  Returns true with the sampled
// direction in dir_out when a light is found.
template <class Sample, class Occlude, class Emits>
bool nee_loop(RayState& st, Sample sample_dir, Occlude occlude, Emits emits, V3& dir_out,
              int& tries) {
  for (int i = 0; i < kNeeTries; ++i) {
    V3 dir = sample_dir();  // occlude may finish it in place
    ++tries;
    const int32_t m = occlude(dir, st.hit);
    if (m < 0) {
      st.hit = false;
    } else if (emits(m)) {
      st.light_mat = m;
      st.hit = true;
    }
    if (st.hit) {
      dir_out = dir;
      return true;
    }
  }
  return false;
}

constexpr int32_t kNonEmitter = 0x7FFFFFFF;  // a material id that never emits

// Union of the root's child boxes.
void root_box(const Bvh& b, float lo[3], float hi[3]) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = INFINITY;
    hi[a] = -INFINITY;
  }
  for (int ch = 0; ch < 2; ++ch) {
    if (b.refs[ch] == kEmpty) continue;
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b.boxes[6 * ch + a]);
      hi[a] = std::max(hi[a], b.boxes[6 * ch + 3 + a]);
    }
  }
}

EmitCull emit_cull(const Scene& s) {
  EmitCull e;
  if (s.T == 0 || s.emit.T == 0) return e;
  float lo[3], hi[3], S = 0.0f;
  root_box(s.all, lo, hi);
  for (int a = 0; a < 3; ++a) S = std::max(S, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
  const float m2 = S * 0x1p-11f;
  root_box(s.emit, lo, hi);
  for (int a = 0; a < 3; ++a) {
    e.c[a] = 0.5f * (lo[a] + hi[a]);
    e.H[a] = 0.5f * (hi[a] - lo[a]) + m2;
  }
  return e;
}

// True when the line through O along d certainly misses every emitter: it
// passes outside the box on one of the axes d x e_x, d x e_y, d x e_z.  The
// test does not depend on the length or sign of d, so the NEE loop runs it
// on hemisphere()'s raw draws (parallel to the normalised direction up to a
// few ulps, an offset over the scene below 16 S u that the margin absorbs).
// A direction with a zero component is never culled.
bool line_culled(const EmitCull& e, V3 O, V3 d) {
  const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
  if (ax == 0.0f || ay == 0.0f || az == 0.0f) return false;
  const float vx = e.c[0] - O.x, vy = e.c[1] - O.y, vz = e.c[2] - O.z;
  if (std::fabs(vy * d.z - vz * d.y) > e.H[1] * az + e.H[2] * ay) return true;
  if (std::fabs(vz * d.x - vx * d.z) > e.H[2] * ax + e.H[0] * az) return true;
  if (std::fabs(vx * d.y - vy * d.x) > e.H[0] * ay + e.H[1] * ax) return true;
  return false;
}

// True when the ray (O, D) certainly misses every emitter: the box lies
// behind the origin along an axis the ray moves away on, or its line misses.
bool emit_culled(const EmitCull& e, V3 O, V3 D) {
  if (D.x == 0.0f || D.y == 0.0f || D.z == 0.0f) return false;
  const float vx = e.c[0] - O.x, vy = e.c[1] - O.y, vz = e.c[2] - O.z;
  if (D.x > 0.0f ? vx < -e.H[0] : vx > e.H[0]) return true;
  if (D.y > 0.0f ? vy < -e.H[1] : vy > e.H[1]) return true;
  if (D.z > 0.0f ? vz < -e.H[2] : vz > e.H[2]) return true;
  return line_culled(e, O, D);
}

// The prefix light sample's estimate p^ = q |cos_y| A / r^2 stays below 1/2
// at O: q <= 3 sqrt(3) / 12 < 0.4331, |cos_y| <= 1 and r^2 >= d^2 (d: O's
// distance to the emitters' box), so 0.9 A <= d^2 gives p^ < 0.49.  Depends
// on the vertex only, so choosing the estimator by it changes no expectation.
bool light_sample_bounded(const EmitCull& e, V3 O, float A) {
  const float dx = std::fmax(std::fabs(O.x - e.c[0]) - e.H[0], 0.0f);
  const float dy = std::fmax(std::fabs(O.y - e.c[1]) - e.H[1], 0.0f);
  const float dz = std::fmax(std::fabs(O.z - e.c[2]) - e.H[2], 0.0f);
  const float d2 = dx * dx + dy * dy + dz * dz;
  return 0.9f * A <= d2;
}

// The occlusion query of NEE.  Returns the material of the closest hit when
// it emits, -1 on a miss, kNonEmitter when the closest hit does not emit --
// exactly the outcomes shader.cu:172-184 distinguish, computed as:
//   1. closest emitter hit (t_E, E) from the emitter BVH;
//   2. if found: any triangle before (t_E, E) in (t, id) order? -> blocked;
//   3. if none and the hit flag is false, miss and non-emitter both leave
//      it false, so no further query is made;
//   4. otherwise any triangle at all? -> non-emitter, else miss.
// `code` is the per-ray record: the emitter's triangle id, or -5 when the
// flag was false and no light was found, -4 non-emitter, -1 miss.
int32_t occlusion_query(const Scene& s, V3 O, V3 D, bool hit_before, int32_t& code,
                        Counters& cnt) {
  const Ray r = make_ray(O, D);
  if (!r.valid) {
    cnt.c[6]++;
    code = hit_before ? -1 : -5;
    return -1;
  }
  float tE = 0.0f, tb = 0.0f;
  const int32_t e = emit_culled(s.ecull, O, D)
                        ? -1
                        : closest<false>(s.emit, r, kTmin, kTmax, 0xFFFFFFFFu, tE, cnt);
  if (e >= 0) {
    if (closest<true>(s.all, r, kTmin, tE, (uint32_t)e, tb, cnt) < 0) {
      code = e;
      return s.mat[e];
    }
    code = hit_before ? -4 : -5;
    return kNonEmitter;
  }
  if (!hit_before) {
    code = -5;
    return kNonEmitter;
  }
  const bool blocked = closest<true>(s.all, r, kTmin, kTmax, 0xFFFFFFFFu, tb, cnt) >= 0;
  code = blocked ? -4 : -1;
  return blocked ? kNonEmitter : -1;
}

// The same record code from one full closest-hit query (the reference
// semantics of an occlusion ray, for checking occlusion_query).
int32_t occlusion_reference(const Scene& s, V3 O, V3 D, bool hit_before, Counters& cnt) {
  const Ray r = make_ray(O, D);
  float t = 0.0f;
  const int32_t p = closest<false>(s.all, r, kTmin, kTmax, 0xFFFFFFFFu, t, cnt);
  if (p >= 0 && s.mats[s.mat[p]].emit) return p;
  if (!hit_before) return -5;
  return p < 0 ? -1 : -4;
}

struct SampleRec {
  int32_t bounces = 0;
  int32_t occl = 0;
  uint32_t hash = 2166136261u;
  int32_t* rad_hits = nullptr;
  int32_t rb = 0;
  void ray(int32_t prim, bool occlusion) {
    const uint32_t code = (uint32_t)prim ^ (occlusion ? 0x5bd1e995u : 0u);
    hash = (hash ^ code) * 16777619u;
  }
};

// Position normalisation of the radiance-cache features: the scene box is
// the union of the root's two child boxes (lisa_render_cached's set-up).
struct CacheBox {
  V3 bmin{0, 0, 0}, binv{1, 1, 1};
};

CacheBox cache_box(const Scene& s) {
  CacheBox c;
  if (s.T == 0) return c;
  const float* b = s.all.boxes.data();
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = s.T > 1 ? std::min(b[a], b[6 + a]) : b[a];
    hi[a] = s.T > 1 ? std::max(b[3 + a], b[9 + a]) : b[3 + a];
  }
  c.bmin = mk(lo[0], lo[1], lo[2]);
  c.binv = mk(1.0f / std::max(hi[0] - lo[0], 1e-20f), 1.0f / std::max(hi[1] - lo[1], 1e-20f),
              1.0f / std::max(hi[2] - lo[2], 1e-20f));
  return c;
}

inline float clamp01(float x) { return std::fmin(std::fmax(x, 0.0f), 1.0f); }

// The 12 cache inputs of a surface hit, each in [0, 1]: position in the
// scene box (3), incoming direction and shading normal as (theta/pi, phi/2pi)
// with theta measured from +y (2 + 2), albedo (3), roughness (0 for a
// dielectric), dielectric flag.
void cache_features(const CacheBox& c, V3 P, V3 D, V3 N, const orc_material& m, float* f) {
  f[0] = clamp01((P.x - c.bmin.x) * c.binv.x);
  f[1] = clamp01((P.y - c.bmin.y) * c.binv.y);
  f[2] = clamp01((P.z - c.bmin.z) * c.binv.z);
  const float dl = std::sqrt(dot(D, D));
  const float dy = dl > 0.0f ? D.y / dl : 0.0f;
  f[3] = std::acos(std::fmin(std::fmax(dy, -1.0f), 1.0f)) * 0.31830988618f;
  f[4] = (std::atan2(D.z, D.x) + 3.14159265359f) * 0.15915494309f;
  f[5] = std::acos(std::fmin(std::fmax(N.y, -1.0f), 1.0f)) * 0.31830988618f;
  f[6] = (std::atan2(N.z, N.x) + 3.14159265359f) * 0.15915494309f;
  const bool diel = m.alpha < 1.0f;
  f[7] = clamp01(m.diffuse[0]);
  f[8] = clamp01(m.diffuse[1]);
  f[9] = clamp01(m.diffuse[2]);
  f[10] = diel ? 0.0f : clamp01(m.roughness);
  f[11] = diel ? 1.0f : 0.0f;
}

// A surface hit of a training path: features, throughput on arrival and the
// radiance gathered before it.
struct CacheVertex {
  float f[12];
  V3 att, color;
};

struct Tracer {
  const Scene& s;
  Counters& cnt;
  uint32_t& seed;
  SampleRec& rec;
  // radiance-cache modes: cache_at > 0 ends the path at the surface hit of
  // radiance ray number cache_at with a query (features + throughput);
  // verts != nullptr records every surface hit (training paths).
  int32_t cache_at = 0;
  const CacheBox* box = nullptr;
  std::vector<CacheVertex>* verts = nullptr;
  int32_t vert_depth = 0;  // > 0: record only the hits of radiance ray number vert_depth
  bool queried = false;
  float qf[12] = {};
  V3 qatt{0, 0, 0};
  // prefix-emission estimator (queries only): at a roughness-1 Lambertian
  // vertex with a bounce left, one light sample takes the emission of the
  // bounce in expectation and the bounce is drawn on the condition that it
  // hits no emitter (cond; din and rough_v are the vertex's for redraws)
  bool emit_expect = false;
  int32_t nee_loops = 1;       // with emit_expect: NEE runs averaged at the camera ray's hit
  int32_t nee_loops_deep = 1;  // ... and at later vertices
  bool cond = false;
  int32_t rejects = 0, max_b = 0;
  V3 din{0, 0, 0};
  float rough_v = 0.0f;

  const orc_material& mat_of(int32_t prim) const { return s.mats[s.mat[prim]]; }

  int32_t trace_ray(V3 O, V3 D, float& t) {
    const Ray r = make_ray(O, D);
    if (!r.valid) {
      cnt.c[6]++;
      return -2;
    }
    return closest<false>(s.all, r, kTmin, kTmax, 0xFFFFFFFFu, t, cnt);
  }

  // shader.cu:196-209 (+ occlusion CH/miss 172-184).
  V3 shoot_ray_to_light(RayState& st) {
    V3 dir;
    int tries = 0;
    const bool found = nee_loop(
        st, [&]() { return hemisphere_raw(seed); },
        [&](V3& d, bool hit_before) {
          cnt.c[1]++;
          rec.occl++;
          int32_t code = -5;
          int32_t m = kNonEmitter;
          // with the flag false, a line that misses every emitter is resolved
          // on the raw draws (a non-emitter and a miss have the same effect)
          if (hit_before || !line_culled(s.ecull, st.origin, d)) {
            d = hemisphere_dir(d, st.normal);
            m = occlusion_query(s, st.origin, d, hit_before, code, cnt);
          }
          rec.ray(code, true);
          return m;
        },
        [&](int32_t m) { return m != kNonEmitter && s.mats[m].emit != 0; }, dir, tries);
    if (found) return ld3(s.mats[st.light_mat].emission) * bsdf_eval(st.normal, dir);
    return mk(0.0f, 0.0f, 0.0f);
  }

  // The light sample of the prefix-emission estimator at the current vertex.
  // The bounce (a hemisphere draw at roughness 1: density q = 1 / (12
  // max|w_i|^3) on N's side) would hit an emitter with probability p and add
  // Le * att, ending the path.  With y drawn by area over the emitters,
  // p^ = V q |cos_y| A / r^2 estimates p and Le p^ the emission without bias;
  // the path adds Le p^ att and goes on with att (1 - p^) behind a bounce
  // drawn on the condition that it hits no emitter, whose expectation is the
  // reference's.  V: y's triangle is the closest emitter along the line and
  // nothing lies before it.
  void emit_sample(RayState& st) {
    const float u0 = rnd(seed), u1 = rnd(seed), u2 = rnd(seed);
    // below the last prefix sum: the search ends on a triangle of positive area
    const float ta = std::fmin(u0 * s.earea, std::nextafter(s.earea, 0.0f));
    int32_t lo = 0, hi = s.emit.T - 1;  // first prefix sum above ta
    while (lo < hi) {
      const int32_t mid = (lo + hi) >> 1;
      if (ta < s.ecdf[(size_t)mid])
        hi = mid;
      else
        lo = mid + 1;
    }
    const float* tv = &s.emit.tv[9 * (size_t)lo];
    const V3 v0 = ld3(tv), v1 = ld3(tv + 3), v2 = ld3(tv + 6);
    const float su = std::sqrt(u1);
    const V3 y = v0 * (1.0f - su) + v1 * (su * (1.0f - u2)) + v2 * (su * u2);
    const V3 wv = y - st.origin;
    const float r2 = dot(wv, wv);
    const V3 w = wv * (1.0f / std::sqrt(r2));
    const Ray r = make_ray(st.origin, w);
    if (!(r2 > 0.0f && dot(st.normal, w) > 0.0f && r.valid)) return;
    const V3 nt = cross(v1 - v0, v2 - v0);
    const float cy = std::fabs(dot(nt, w)) / std::sqrt(dot(nt, nt));
    const float q = bsdf_pdf(st.direction, st.normal, w, 1.0f);  // the bounce at roughness 1
    float te = 0.0f, tb = 0.0f;
    const int32_t e = closest<false>(s.emit, r, kTmin, kTmax, 0xFFFFFFFFu, te, cnt);
    if (e != (int32_t)s.emit.prim[(size_t)lo]) return;
    const float lw = q * cy * s.earea / r2;
    const float ph = closest<true>(s.all, r, kTmin, te, (uint32_t)e, tb, cnt) >= 0 ? 0.0f : lw;
    st.color = st.color + (ld3(s.mats[s.mat[e]].emission) * ph) * st.att;
    st.att = st.att * (1.0f - ph);
  }

  // One radiance trace: shader.cu:77-98 + closest-hit 211-255 / miss 189-194.
  // Returns false for a conditioned bounce that hit an emitter and was drawn
  // again (not a bounce; after kMaxRedraws of them the path ends).
  bool radiance(RayState& st, V3 O, V3 D) {
    cnt.c[0]++;
    float t = 0.0f;
    const int32_t p = trace_ray(O, D, t);
    if (cond) {
      cond = false;
      if (p >= 0 && mat_of(p).emit) {
        if (++rejects == kMaxRedraws) {
          st.done = true;
        } else {
          st.direction = bsdf_bounce(din, st.normal, seed, rough_v);
          cond = true;
        }
        return false;
      }
    }
    rec.ray(p, false);
    if (rec.rad_hits && rec.bounces < rec.rb) rec.rad_hits[rec.bounces] = p;
    rec.bounces++;
    if (p < 0) {
      st.color = st.color + mk(0.0f, 0.0f, 0.0f);
      st.done = true;
      return true;
    }
    const orc_material& m = mat_of(p);
    if (m.emit) {
      st.color = st.color + ld3(m.emission) * st.att;
      st.done = true;
      return true;
    }
    cnt.c[4]++;
    st.origin = O + t * D;
    st.direction = D;
    {
      const float* v = &s.v[9 * (size_t)p];
      const float* n = &s.n[9 * (size_t)p];
      st.normal = bary_normal(st.origin, ld3(n), ld3(n + 3), ld3(n + 6), ld3(v), ld3(v + 3), ld3(v + 6));
    }
    if (cache_at > 0 && rec.bounces == cache_at) {
      cache_features(*box, st.origin, D, st.normal, m, qf);
      qatt = st.att;
      queried = true;
      st.done = true;
      return true;
    }
    if (verts && (vert_depth == 0 || rec.bounces == vert_depth)) {
      CacheVertex cv;
      cache_features(*box, st.origin, D, st.normal, m, cv.f);
      cv.att = st.att;
      cv.color = st.color;
      verts->push_back(cv);
    }
    if (m.alpha < 1.0f) {
      float cosI = dot(D, st.normal);
      float eta;
      V3 nn;
      if (cosI < 0.0f) {
        cosI = -cosI;
        eta = 1.0f / m.n;
        nn = st.normal;
      } else {
        st.att = st.att * ld3(m.diffuse);
        eta = m.n;
        nn = -st.normal;
      }
      if (eta == 1.0f) {
        st.direction = D;
      } else if (rnd(seed) <= fresnel(cosI, eta)) {
        st.direction = reflect(D, nn);
      } else {
        st.direction = refract(cosI, D, nn, eta);
      }
    } else {
      st.att = st.att * ld3(m.diffuse);
      const int32_t runs = rec.bounces > 1 ? nee_loops_deep : nee_loops;
      if (emit_expect && runs > 1) {
        // the mean of the vertex's NEE runs (nee_loops at the camera ray's
        // hit, nee_loops_deep later) from its entry flag and light, each
        // added as it ends; the last run's flag and light stay
        const bool e_hit = st.hit;
        const int32_t e_light = st.light_mat;
        const float inv = 1.0f / (float)runs;
        for (int32_t k = 0; k < runs; ++k) {
          st.hit = e_hit;
          st.light_mat = e_light;
          const V3 nee = shoot_ray_to_light(st);
          st.color = st.color + (nee * inv) * st.att;
        }
      } else {
        const V3 nee = shoot_ray_to_light(st);
        st.color = st.color + nee * st.att;
      }
      if (emit_expect && m.roughness == 1.0f && s.emit.T > 0 && s.earea > 0.0f && rec.bounces < max_b &&
          light_sample_bounded(s.ebox, st.origin, s.earea)) {
        emit_sample(st);
        cond = true;
        rejects = 0;
        din = st.direction;
        rough_v = m.roughness;
      }
      st.direction = bsdf_bounce(st.direction, st.normal, seed, m.roughness);
    }
    return true;
  }

  // shader.cu:102-124.
  V3 trace(V3 O, V3 D, int bounces) {
    RayState st;
    max_b = bounces;
    cond = false;
    for (int i = 0; i < bounces;) {
      if (radiance(st, O, D)) ++i;
      D = st.direction;
      O = st.origin;
      if (st.done) break;
    }
    return st.color;
  }
};

// shader.cu:126-167 for one pixel.
void render_pixel(const Scene& s, const orc_render_params& p, uint32_t pix, float* out4,
                  const float* prev4, float* srgb, int32_t* sb, int32_t* so, uint32_t* sh,
                  int32_t* rh, Counters& cnt) {
  const uint32_t x = pix % p.width, y = pix / p.width;
  uint32_t seed = tea16(y * p.width + x, p.subframe);
  const float sx = (float)p.width, sy = (float)p.height;
  const float ix = (float)x, iy = (float)y;
  const V3 eye = ld3(p.eye), U = ld3(p.U), V = ld3(p.V), W = ld3(p.W);
  V3 accum = mk(0.0f, 0.0f, 0.0f);
  for (uint32_t i = 0; i < p.spp; ++i) {
    const float jx = rng(seed);
    const float jy = rng(seed);
    const float dx = (2.0f * ix + jx) / sx - 1.0f;
    const float dy = (2.0f * iy + jy) / sy - 1.0f;
    const V3 D = normalize(dx * U + dy * V + W);
    SampleRec rec;
    rec.rb = p.record_bounces;
    if (rh) {
      rec.rad_hits = rh + (size_t)i * p.record_bounces;
      for (int b = 0; b < p.record_bounces; ++b) rec.rad_hits[b] = -3;
    }
    Tracer tr{s, cnt, seed, rec};
    const V3 c = tr.trace(eye, D, (int)p.bounces);
    cnt.c[5]++;
    if (srgb) st3(srgb + 3 * (size_t)i, c);
    if (sb) sb[i] = rec.bounces;
    if (so) so[i] = rec.occl;
    if (sh) sh[i] = rec.hash;
    accum = accum + c;
  }
  const float inv = 1.0f / (float)p.spp;
  accum = accum * inv;
  if (p.subframe > 0 && prev4) {  // shader.cu:160-164; without a previous buffer the subframe only seeds
    const float a = 1.0f / (float)(p.subframe + 1);
    accum = lerp(ld3(prev4), accum, a);
  }
  st3(out4, accum);
  out4[3] = 1.0f;
  cnt.c[7]++;
}

}  // namespace

struct orc_scene {
  Scene s;
};

extern "C" {

uint32_t orc_tea16(uint32_t v0, uint32_t v1) { return tea16(v0, v1); }
uint32_t orc_lcg(uint32_t* state) { return lcg(*state); }
float orc_rnd(uint32_t* state) { return rnd(*state); }
float orc_rng(uint32_t* state) { return rng(*state); }
void orc_hemisphere(const float* N, uint32_t* state, float* out3) { st3(out3, hemisphere(ld3(N), *state)); }
float orc_fresnel(float cosT, float eta) { return fresnel(cosT, eta); }
void orc_refract(float cosI, const float* d, const float* N, float eta, float* out3) {
  st3(out3, refract(cosI, ld3(d), ld3(N), eta));
}
void orc_reflect(const float* i, const float* n, float* out3) { st3(out3, reflect(ld3(i), ld3(n))); }
float orc_brdf(const float* N, const float* L) { return bsdf_eval(ld3(N), ld3(L)); }
void orc_bounce(const float* d, const float* N, uint32_t* state, float roughness, float* out3) {
  st3(out3, bsdf_bounce(ld3(d), ld3(N), *state, roughness));
}
void orc_bounce_many(const float* d, const float* N, uint32_t* state, float roughness, int32_t n,
                     float* out3) {
  for (int32_t i = 0; i < n; ++i) st3(out3 + 3 * (size_t)i, bsdf_bounce(ld3(d), ld3(N), *state, roughness));
}
void orc_bsdf_pdf(const float* d, const float* N, const float* L, int32_t n, float roughness, float* out) {
  for (int32_t i = 0; i < n; ++i) out[i] = bsdf_pdf(ld3(d), ld3(N), ld3(L + 3 * (size_t)i), roughness);
}
void orc_bary_normal(const float* P, const float* n1, const float* n2, const float* n3,
                     const float* v1, const float* v2, const float* v3, float* out3) {
  st3(out3, bary_normal(ld3(P), ld3(n1), ld3(n2), ld3(n3), ld3(v1), ld3(v2), ld3(v3)));
}
// src/sutil/Camera.cpp:34-45 with up = (0,1,0) (optix_wrapper.cc:434-442).
void orc_camera_uvw(const float* eye, const float* lookat, float fov, float aspect, float* U,
                    float* V, float* W) {
  const V3 e = ld3(eye), l = ld3(lookat), up = mk(0.0f, 1.0f, 0.0f);
  const V3 w = l - e;
  const float wlen = std::sqrt(dot(w, w));
  V3 u = normalize(cross(w, up));
  V3 v = normalize(cross(u, w));
  const float vlen = wlen * std::tan(0.5f * fov * 3.14159265358979323846f / 180.0f);
  v = v * vlen;
  const float ulen = vlen * aspect;
  u = u * ulen;
  st3(U, u);
  st3(V, v);
  st3(W, w);
}
void orc_srgb_quantize(const float* rgba, int n_pixels, uint8_t* rgba8) {
  for (int i = 0; i < n_pixels; ++i) {
    for (int c = 0; c < 3; ++c) rgba8[4 * i + c] = quant8(to_srgb(clampf(rgba[4 * i + c], 0.0f, 1.0f)));
    rgba8[4 * i + 3] = 255;
  }
}
// src/sutil/sutil.cpp:97-117 (P6 header) and :541-553 (row flip).
int orc_save_ppm(const char* path, const uint8_t* rgba8, int width, int height) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return -1;
  std::fprintf(f, "P6\n%d %d\n255\n", width, height);
  std::vector<uint8_t> row(3 * (size_t)width);
  for (int j = height - 1; j >= 0; --j) {
    for (int i = 0; i < width; ++i)
      for (int c = 0; c < 3; ++c) row[3 * i + c] = rgba8[4 * ((size_t)j * width + i) + c];
    std::fwrite(row.data(), 1, row.size(), f);
  }
  std::fclose(f);
  return 0;
}
uint32_t orc_morton3(float x, float y, float z) { return morton3(x, y, z); }

int32_t orc_nee_scripted(const int32_t* outcomes, int32_t n, int32_t hit_in, int32_t light_in,
                         int32_t* hit_out, int32_t* light_out) {
  RayState st;
  st.hit = hit_in != 0;
  st.light_mat = light_in;
  int i = 0, tries = 0;
  V3 dir;
  const bool found = nee_loop(
      st, [&]() { return mk((float)i, 0.0f, 0.0f); },
      [&](V3&, bool) {
        const int32_t m = i < n ? outcomes[i] : 0;
        ++i;
        return m;
      },
      [&](int32_t m) { return m > 0; }, dir, tries);
  *hit_out = st.hit ? 1 : 0;
  *light_out = st.light_mat;
  return found ? tries : -tries;
}

orc_scene* orc_scene_create(const float* verts, const float* normals, const int32_t* mat_idx,
                            int32_t T, const orc_material* mats, int32_t M) {
  orc_scene* o = new orc_scene;
  Scene& s = o->s;
  s.T = T;
  s.v.assign(verts, verts + 9 * (size_t)T);
  s.n.assign(normals, normals + 9 * (size_t)T);
  s.mat.assign(mat_idx, mat_idx + T);
  s.mats.assign(mats, mats + M);
  std::vector<uint32_t> gid_all(T), gid_emit;
  for (int32_t k = 0; k < T; ++k) {
    gid_all[k] = (uint32_t)k;
    if (s.mats[s.mat[k]].emit) gid_emit.push_back((uint32_t)k);
  }
  build_bvh(s.all, s.v, gid_all);
  build_bvh(s.emit, s.v, gid_emit);
  s.ecull = emit_cull(s);
  s.ebox = s.ecull;
  emit_cdf(s);
  return o;
}
void orc_scene_destroy(orc_scene* s) { delete s; }

void orc_scene_update_vertices(orc_scene* o, const float* verts, const float* normals) {
  Scene& s = o->s;
  s.v.assign(verts, verts + 9 * (size_t)s.T);
  if (normals) s.n.assign(normals, normals + 9 * (size_t)s.T);
  update_bvh(s.all, s.v);
  update_bvh(s.emit, s.v);
  s.ecull = emit_cull(s);
  s.ebox = s.ecull;
  emit_cdf(s);
}
int32_t orc_num_tris(const orc_scene* s) { return s->s.T; }

int32_t orc_num_emitters(const orc_scene* s) { return s->s.emit.T; }

void orc_raw_cull(const orc_scene* o, const float* O, const float* d, int32_t n, int32_t* culled) {
  for (int32_t i = 0; i < n; ++i)
    culled[i] = line_culled(o->s.ecull, ld3(O + 3 * (size_t)i), ld3(d + 3 * (size_t)i)) ? 1 : 0;
}

void orc_scene_set_emitter_cull(orc_scene* o, int32_t enabled) {
  o->s.ecull = enabled ? emit_cull(o->s) : EmitCull();
}

int orc_get_bvh(const orc_scene* o, int32_t which, uint32_t* morton_sorted, uint32_t* prim_sorted,
                uint32_t* child_refs, float* child_boxes) {
  const Bvh& s = which == 0 ? o->s.all : o->s.emit;
  if (s.T == 0) return 0;
  std::copy(s.morton.begin(), s.morton.end(), morton_sorted);
  std::copy(s.prim.begin(), s.prim.end(), prim_sorted);
  std::copy(s.refs.begin(), s.refs.end(), child_refs);
  std::copy(s.boxes.begin(), s.boxes.end(), child_boxes);
  return 0;
}

void orc_closest_hit(const orc_scene* o, const float* O, const float* D, int32_t n, float tmin,
                     float tmax, int32_t* prim_out, float* t_out, uint64_t* counters, int threads) {
  const Scene& s = o->s;
  Counters total;
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
  {
    Counters cnt;
#pragma omp for schedule(dynamic, 256)
    for (int32_t i = 0; i < n; ++i) {
      const Ray r = make_ray(ld3(O + 3 * (size_t)i), ld3(D + 3 * (size_t)i));
      float t = 0.0f;
      const int32_t p = closest<false>(s.all, r, tmin, tmax, 0xFFFFFFFFu, t, cnt);
      prim_out[i] = p;
      t_out[i] = p >= 0 ? t : 0.0f;
    }
#pragma omp critical
    for (int k = 0; k < ORC_NCOUNTERS; ++k) total.c[k] += cnt.c[k];
  }
  if (counters)
    for (int k = 0; k < ORC_NCOUNTERS; ++k) counters[k] = total.c[k];
}

void orc_occlusion(const orc_scene* o, const float* O, const float* D, int32_t n,
                   const int32_t* hit_before, int32_t* code_fast, int32_t* code_ref,
                   uint64_t* counters) {
  const Scene& s = o->s;
  Counters cf, cr;
  for (int32_t i = 0; i < n; ++i) {
    const V3 Oi = ld3(O + 3 * (size_t)i), Di = ld3(D + 3 * (size_t)i);
    int32_t code;
    occlusion_query(s, Oi, Di, hit_before[i] != 0, code, cf);
    code_fast[i] = code;
    code_ref[i] = occlusion_reference(s, Oi, Di, hit_before[i] != 0, cr);
  }
  if (counters)
    for (int k = 0; k < ORC_NCOUNTERS; ++k) counters[k] = cf.c[k];
}

void orc_closest_hit_brute(const orc_scene* o, const float* O, const float* D, int32_t n,
                           float tmin, float tmax, int32_t* prim_out, float* t_out) {
  const Scene& s = o->s;
  for (int32_t i = 0; i < n; ++i) {
    const Ray r = make_ray(ld3(O + 3 * (size_t)i), ld3(D + 3 * (size_t)i));
    float t = 0.0f;
    const int32_t p = closest_brute(s, r, tmin, tmax, t);
    prim_out[i] = p;
    t_out[i] = p >= 0 ? t : 0.0f;
  }
}

void orc_cache_bounds(const orc_scene* o, float* bmin, float* binv) {
  const CacheBox c = cache_box(o->s);
  st3(bmin, c.bmin);
  st3(binv, c.binv);
}

int orc_render_full_paths(const orc_scene* o, const orc_render_params* p, const int32_t* pixel_ids,
                          int32_t n, float* accum_out, int threads, int emit_expect, int nee_loops,
                          int nee_loops_deep) {
  const Scene& s = o->s;
  if (p->width == 0 || p->height == 0 || p->spp == 0 || n <= 0) return -1;
  if (nee_loops > 16 || nee_loops_deep > 16 || ((nee_loops > 1 || nee_loops_deep > 1) && !emit_expect))
    return -3;
  const int64_t npx = (int64_t)p->width * p->height;
  for (int32_t i = 0; i < n; ++i)
    if (pixel_ids[i] < 0 || pixel_ids[i] >= npx) return -2;
  const CacheBox box = cache_box(s);
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
  {
    Counters cnt;
#pragma omp for schedule(dynamic, 4)
    for (int32_t i = 0; i < n; ++i) {
      const uint32_t pix = (uint32_t)pixel_ids[i];
      const uint32_t x = pix % p->width, y = pix / p->width;
      uint32_t seed = tea16(y * p->width + x, p->subframe);
      const V3 eye = ld3(p->eye), U = ld3(p->U), V = ld3(p->V), W = ld3(p->W);
      V3 accum = mk(0.0f, 0.0f, 0.0f);
      for (uint32_t k = 0; k < p->spp; ++k) {
        const float jx = rng(seed);
        const float jy = rng(seed);
        const float dx = (2.0f * (float)x + jx) / (float)p->width - 1.0f;
        const float dy = (2.0f * (float)y + jy) / (float)p->height - 1.0f;
        const V3 D = normalize(dx * U + dy * V + W);
        SampleRec rec;
        Tracer tr{s, cnt, seed, rec};
        tr.cache_at = 0;  // no query: the whole path
        tr.box = &box;
        tr.emit_expect = emit_expect != 0;
        tr.nee_loops = nee_loops > 1 ? nee_loops : 1;
        tr.nee_loops_deep = nee_loops_deep > 1 ? nee_loops_deep : 1;
        accum = accum + tr.trace(eye, D, (int)p->bounces);
      }
      st3(accum_out + 4 * (size_t)i, accum * (1.0f / (float)p->spp));
      accum_out[4 * (size_t)i + 3] = 1.0f;
    }
  }
  return 0;
}

int orc_cache_queries(const orc_scene* o, const orc_render_params* p, uint32_t prefix_bounces,
                      const int32_t* pixel_ids, int32_t n, float* accum_out, float* q_feat,
                      float* q_att, int32_t* q_item, int32_t q_cap, int32_t* n_queries, int threads,
                      int emit_expect, int nee_loops, int nee_loops_deep) {
  const Scene& s = o->s;
  if (p->width == 0 || p->height == 0 || p->spp == 0 || n <= 0) return -1;
  if (nee_loops > 16 || nee_loops_deep > 16 || ((nee_loops > 1 || nee_loops_deep > 1) && !emit_expect))
    return -3;
  const int64_t npx = (int64_t)p->width * p->height;
  for (int32_t i = 0; i < n; ++i)
    if (pixel_ids[i] < 0 || pixel_ids[i] >= npx) return -2;
  const CacheBox box = cache_box(s);
  const size_t spp = p->spp;
  std::vector<uint8_t> has(spp * n, 0);
  std::vector<float> f(12 * spp * n), at(3 * spp * n);
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
  {
    Counters cnt;
#pragma omp for schedule(dynamic, 4)
    for (int32_t i = 0; i < n; ++i) {
      const uint32_t pix = (uint32_t)pixel_ids[i];
      const uint32_t x = pix % p->width, y = pix / p->width;
      uint32_t seed = tea16(y * p->width + x, p->subframe);
      const V3 eye = ld3(p->eye), U = ld3(p->U), V = ld3(p->V), W = ld3(p->W);
      V3 accum = mk(0.0f, 0.0f, 0.0f);
      for (size_t k = 0; k < spp; ++k) {
        const float jx = rng(seed);
        const float jy = rng(seed);
        const float dx = (2.0f * (float)x + jx) / (float)p->width - 1.0f;
        const float dy = (2.0f * (float)y + jy) / (float)p->height - 1.0f;
        const V3 D = normalize(dx * U + dy * V + W);
        SampleRec rec;
        Tracer tr{s, cnt, seed, rec};
        tr.cache_at = (int32_t)prefix_bounces + 1;
        tr.box = &box;
        tr.emit_expect = emit_expect != 0;
        tr.nee_loops = nee_loops > 1 ? nee_loops : 1;
        tr.nee_loops_deep = nee_loops_deep > 1 ? nee_loops_deep : 1;
        const V3 c = tr.trace(eye, D, (int)prefix_bounces + 1);
        accum = accum + c;
        if (tr.queried) {
          const size_t q = (size_t)i * spp + k;
          has[q] = 1;
          std::copy(tr.qf, tr.qf + 12, &f[12 * q]);
          st3(&at[3 * q], tr.qatt);
        }
      }
      st3(accum_out + 4 * (size_t)i, accum * (1.0f / (float)p->spp));
      accum_out[4 * (size_t)i + 3] = 1.0f;
    }
  }
  int32_t nq = 0;
  for (size_t q = 0; q < has.size(); ++q) {
    if (!has[q]) continue;
    if (nq < q_cap) {
      std::copy(&f[12 * q], &f[12 * q] + 12, q_feat + 12 * (size_t)nq);
      std::copy(&at[3 * q], &at[3 * q] + 3, q_att + 3 * (size_t)nq);
      q_item[nq] = (int32_t)(q / spp);
    }
    ++nq;
  }
  *n_queries = nq;
  return 0;
}

int orc_cache_train_rows(const orc_scene* o, const orc_render_params* p, const int32_t* pixel_ids,
                         int32_t n, uint32_t depth, float* feat, float* tgt, int32_t* row_path,
                         int32_t cap, int32_t* n_rows, int threads) {
  const Scene& s = o->s;
  if (p->width == 0 || p->height == 0 || p->spp == 0 || p->bounces == 0 || n <= 0) return -1;
  const int64_t npx = (int64_t)p->width * p->height;
  for (int32_t i = 0; i < n; ++i)
    if (pixel_ids[i] < 0 || pixel_ids[i] >= npx) return -2;
  const CacheBox box = cache_box(s);
  std::vector<std::vector<float>> rows(n);  // 15 floats per row: features, target
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
  {
    Counters cnt;
#pragma omp for schedule(dynamic, 4)
    for (int32_t i = 0; i < n; ++i) {
      const uint32_t pix = (uint32_t)pixel_ids[i];
      const uint32_t x = pix % p->width, y = pix / p->width;
      uint32_t seed = tea16(y * p->width + x, p->subframe);
      const V3 eye = ld3(p->eye), U = ld3(p->U), V = ld3(p->V), W = ld3(p->W);
      for (uint32_t k = 0; k < p->spp; ++k) {
        const float jx = rng(seed);
        const float jy = rng(seed);
        const float dx = (2.0f * (float)x + jx) / (float)p->width - 1.0f;
        const float dy = (2.0f * (float)y + jy) / (float)p->height - 1.0f;
        const V3 D = normalize(dx * U + dy * V + W);
        SampleRec rec;
        std::vector<CacheVertex> verts;
        Tracer tr{s, cnt, seed, rec};
        tr.box = &box;
        tr.verts = &verts;
        tr.vert_depth = (int32_t)depth;
        const V3 c = tr.trace(eye, D, (int)p->bounces);
        for (const CacheVertex& v : verts) {  // suffix radiance per unit throughput
          rows[i].insert(rows[i].end(), v.f, v.f + 12);
          rows[i].push_back((c.x - v.color.x) / v.att.x);
          rows[i].push_back((c.y - v.color.y) / v.att.y);
          rows[i].push_back((c.z - v.color.z) / v.att.z);
        }
      }
    }
  }
  int32_t nr = 0;
  for (int32_t i = 0; i < n; ++i)
    for (size_t r = 0; r < rows[i].size() / 15; ++r, ++nr) {
      if (nr >= cap) continue;
      std::copy(&rows[i][15 * r], &rows[i][15 * r] + 12, feat + 12 * (size_t)nr);
      std::copy(&rows[i][15 * r + 12], &rows[i][15 * r] + 15, tgt + 3 * (size_t)nr);
      if (row_path) row_path[nr] = i;
    }
  *n_rows = nr;
  return 0;
}

int orc_render(const orc_scene* o, const orc_render_params* p, const int32_t* pixel_ids,
               int32_t n_pixels, float* accum_out, const float* accum_prev, float* sample_rgb,
               int32_t* sample_bounces, int32_t* sample_occl, uint32_t* sample_hash,
               int32_t* rad_hits, uint64_t* counters, int threads) {
  const Scene& s = o->s;
  if (p->width == 0 || p->height == 0 || p->spp == 0) return -1;
  const int64_t npx = (int64_t)p->width * p->height;
  const int64_t n = pixel_ids ? n_pixels : npx;
  for (int64_t i = 0; i < n; ++i)
    if (pixel_ids && (pixel_ids[i] < 0 || pixel_ids[i] >= npx)) return -2;
  const size_t spp = p->spp, rb = (size_t)std::max(p->record_bounces, 0);
  Counters total;
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
  {
    Counters cnt;
#pragma omp for schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
      const uint32_t pix = pixel_ids ? (uint32_t)pixel_ids[i] : (uint32_t)i;
      render_pixel(s, *p, pix, accum_out + 4 * i, accum_prev ? accum_prev + 4 * i : nullptr,
                   sample_rgb ? sample_rgb + 3 * spp * i : nullptr,
                   sample_bounces ? sample_bounces + spp * i : nullptr,
                   sample_occl ? sample_occl + spp * i : nullptr,
                   sample_hash ? sample_hash + spp * i : nullptr,
                   (rad_hits && rb) ? rad_hits + spp * rb * i : nullptr, cnt);
    }
#pragma omp critical
    for (int k = 0; k < ORC_NCOUNTERS; ++k) total.c[k] += cnt.c[k];
  }
  if (counters)
    for (int k = 0; k < ORC_NCOUNTERS; ++k) counters[k] = total.c[k];
  return 0;
}

}  // extern "C"
