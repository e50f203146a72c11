// Device float3 algebra and the per-material BSDF of LiSA's hot path, written
// to the operation order of the reference so FP32 results are reproducible:
//   src/sutil/vec_math.h       dot 521, cross 527, normalize 539-543,
//                              reflect 552-555, faceforward 561-564, lerp 494-497,
//                              clamp 113, float3 / s 477-481
//   src/cuda/random.h          tea<16> 31-46, lcg 49-55, rnd 64-67
//   src/LiSA/src/maths.cu      rng 6-8, shoot_ray_hemisphere 10-15, fresnel 17-20,
//                              refract 22-30, barycentric_normal 33-56
//   src/LiSA/src/bsdfs/lambertian.cu  bounce 7-13, BRDF 15-22, BTDF 24-27
// Compiled with -ffp-contract=off and correctly rounded division / sqrt: one
// IEEE-rounded operation per operator, no FMA contraction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lisa {

struct F3 {
  float x, y, z;
};
__host__ __device__ __forceinline__ F3 mk(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 operator-(F3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 normalize(F3 v) {
  const float inv = 1.0f / sqrtf(dot(v, v));
  return v * inv;
}
__device__ __forceinline__ F3 reflect(F3 i, F3 n) { return i - (2.0f * n) * dot(n, i); }
__device__ __forceinline__ F3 faceforward(F3 n, F3 i, F3 nref) {
  return n * copysignf(1.0f, dot(i, nref));
}
__device__ __forceinline__ F3 lerp(F3 a, F3 b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float clampf(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }
__device__ __forceinline__ float sel(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ F3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

// ---- RNG ----
__device__ __forceinline__ uint32_t tea16(uint32_t v0, uint32_t v1) {
  uint32_t s0 = 0;
#pragma unroll
  for (int n = 0; n < 16; ++n) {
    s0 += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
__device__ __forceinline__ float rnd(uint32_t& s) {
  s = 1664525u * s + 1013904223u;
  return (float)(s & 0x00FFFFFFu) / (float)0x01000000;
}
// rnd(s) * 2 - 1 (maths.cu:6-8) as one fused multiply-add: k * 2^-24 * 2 - 1
// is exact for every 24-bit k (a multiple of 2^-23 in [-1, 1)), so the fused
// form gives the same bits with two instructions fewer.
__device__ __forceinline__ float rng(uint32_t& s) {
  s = 1664525u * s + 1013904223u;
  return __builtin_fmaf((float)(s & 0x00FFFFFFu), 0x1p-23f, -1.0f);
}

// ---- maths.cu ----
// shoot_ray_hemisphere in two steps: the three draws (x, y, z), then the
// normalised direction turned to the side of N.
__device__ __forceinline__ F3 hemisphere_raw(uint32_t& s) {
  const float x = rng(s);
  const float y = rng(s);
  const float z = rng(s);
  return mk(x, y, z);
}
__device__ __forceinline__ F3 hemisphere_dir(F3 raw, F3 N) {
  const F3 r = normalize(raw);
  return faceforward(r, N, r);
}
__device__ __forceinline__ F3 hemisphere(F3 N, uint32_t& s) { return hemisphere_dir(hemisphere_raw(s), N); }
__device__ __forceinline__ float fresnel(float cosT, float eta) {
  const float r = (1.0f - eta) / (1.0f + eta);
  const float R0 = r * r;
  const float x = 1.0f - cosT;
  const float x2 = x * x;
  const float x5 = (x2 * x2) * x;
  return R0 + (1.0f - R0) * x5;
}
__device__ __forceinline__ F3 refract(float cosI, F3 d, F3 N, float eta) {
  const float cost2 = 1.0f - eta * eta * (1.0f - cosI * cosI);
  const F3 t = eta * d + ((eta * cosI - sqrtf(fabsf(cost2))) * N);
  return t * (float)(cost2 > 0.0f);
}
__device__ __forceinline__ F3 bary_normal(F3 P, F3 n1, F3 n2, F3 n3, F3 v1, F3 v2, F3 v3) {
  const F3 e1 = v2 - v1;
  const F3 e2 = v3 - v1;
  const F3 i = P - v1;
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

// ---- the interchangeable BSDF (lambertian.cu) ----
// sample: lerp(reflect(d, N), hemisphere(N), roughness); not renormalised.
struct Lambertian {
  __device__ static __forceinline__ F3 sample(F3 d, F3 N, uint32_t& s, float roughness) {
    const F3 r = reflect(d, N);
    const F3 h = hemisphere(N, s);
    return lerp(r, h, roughness);
  }
  // Density per steradian with which sample(d, N, ., roughness) returns the
  // unit direction L (0 for roughness 0, a mirror).  The hemisphere draw h
  // is a cube-uniform point normalised and folded to N's side, density
  // q(h) = 1 / (12 max|h_i|^3); sample's direction is that of
  // v = (1 - t) r + t h (t = roughness, r = reflect(d, N)), i.e. of a point
  // on a unit sphere about c = (1 - t) r / t seen from the origin, so each
  // preimage h of L adds q(h) (s / t)^2 / |h . L| with v = s L.
  __device__ static __forceinline__ float pdf(F3 d, F3 N, F3 L, float roughness) {
    if (!(roughness > 0.0f && roughness <= 1.0f)) return 0.0f;
    if (roughness == 1.0f) {  // h = L
      if (!(dot(N, L) > 0.0f)) return 0.0f;
      const float mw = fmaxf(fabsf(L.x), fmaxf(fabsf(L.y), fabsf(L.z)));
      return 1.0f / (12.0f * mw * mw * mw);
    }
    const float t = roughness, u = 1.0f - roughness;
    const F3 r = reflect(d, N);
    const float a = u * dot(L, r);
    const float disc = a * a - u * u + t * t;
    if (disc < 0.0f) return 0.0f;
    const float sq = sqrtf(disc);
    float p = 0.0f;
    for (int k = 0; k < 2; ++k) {
      const float s = k == 0 ? a + sq : a - sq;
      if (!(s > 0.0f)) continue;
      const F3 h = (s * L - u * r) * (1.0f / t);
      const float hl = fabsf(dot(h, L));
      if (!(dot(h, N) > 0.0f) || !(hl > 0.0f)) continue;
      const float mh = fmaxf(fabsf(h.x), fmaxf(fabsf(h.y), fabsf(h.z)));
      p += (1.0f / (12.0f * mh * mh * mh)) * ((s / t) * (s / t)) / hl;
      if (sq == 0.0f) break;
    }
    return p;
  }
  // eval * pdf weight: NdotL * (float)((double)NdotL / M_PI).
  __device__ static __forceinline__ float eval(F3 N, F3 L) {
    const float NdotL = clampf(dot(N, L), 0.0f, 1.0f);
    const float p = (float)((double)NdotL / 3.14159265358979323846);
    return NdotL * p;
  }
};
// dielectric interface (BTDF = Schlick Fresnel) and the side logic of
// shader.cu:226-246.
struct Dielectric {
  __device__ static __forceinline__ float btdf(float cosI, float eta) { return fresnel(cosI, eta); }
  // Returns the continuation direction; multiplies att by the diffuse colour on exit.
  __device__ static __forceinline__ F3 scatter(F3 D, F3 N, float n_index, F3 diffuse, F3& att,
                                               uint32_t& s) {
    float cosI = dot(D, N);
    float eta;
    F3 nn;
    if (cosI < 0.0f) {
      cosI = -cosI;
      eta = 1.0f / n_index;
      nn = N;
    } else {
      att = att * diffuse;
      eta = n_index;
      nn = -N;
    }
    if (eta == 1.0f) return D;
    if (rnd(s) <= btdf(cosI, eta)) return reflect(D, nn);
    return refract(cosI, D, nn, eta);
  }
};

}  // namespace lisa
