// Learned radiance cache: the MLP of lisa_mlp.h on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) for gfx950.
//
//   k_mlp_infer      frequency encoding + 5-layer forward, 16 queries per
//                    wave-tile, weights and activations in registers
//   k_mlp_encode     the frequency encoding of a training call's rows, once
//                    for all of its steps (sample- and feature-major copies)
//   k_mlp_train      the forward pass with the relative-L2 loss gradient and
//                    the backward data pass (dZ_l = (dZ_{l+1} W_{l+1}^T) *
//                    [A_{l+1} > 0]); writes the layer inputs and dZ
//                    feature-major for the weight GEMMs
//   k_mlp_wgrad      dW_l^T = dZ_l * in_l^T over the batch (split-K partials)
//   k_mlp_adam       the partials' sum, Adam on the FP32 master weights, BF16
//                    copies for the forward (W^T) and backward (W) layouts
//
// MFMA 16x16x32 lane layout (64 lanes, q = lane / 16, r = lane % 16):
//   A (16 x 32): lane holds A[r][8q .. 8q+7]
//   B (32 x 16): lane holds B[8q .. 8q+7][r]
//   D (16 x 16): lane holds D[4q .. 4q+3][r]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lisa.h"
#include "lisa_internal.h"
#include "lisa_mlp.h"

using namespace lisa::mlp;

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

// Round-to-nearest-even FP32 -> BF16 (NaN kept quiet).
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline bf16x8 ld8(const uint16_t* p) {  // 16-byte aligned
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS row strides (elements): padded by 8 so the 16 rows a wave reads
// together fall on different banks; 16-byte alignment is kept.
constexpr int kHStride = kHidden + 8;  // 72
constexpr int kGStride = kHidden + 8;  // dZ tiles
__host__ __device__ constexpr int wt_stride(int l) { return kInDim[l] + 8; }
__host__ __device__ constexpr int wt_lds_offset(int l) {
  int o = 0;
  for (int k = 0; k < l; ++k) o += kOutDim[k] * wt_stride(k);
  return o;
}
constexpr int kWtLds = wt_lds_offset(kLayers);  // 25728 elements
// Backward layout W_l[in][out_pad2] for layers 1..4 (out padded to a multiple of 32).
__host__ __device__ constexpr int wb_out(int l) { return l == kLayers - 1 ? 32 : kHidden; }
__host__ __device__ constexpr int wb_global_offset(int l) {
  int o = 0;
  for (int k = 1; k < l; ++k) o += kHidden * wb_out(k);
  return o;
}
constexpr int kWbGlobal = wb_global_offset(kLayers);  // 14336
__host__ __device__ constexpr int wb_stride(int l) { return wb_out(l) + 8; }
__host__ __device__ constexpr int wb_lds_offset(int l) {
  int o = 0;
  for (int k = 1; k < l; ++k) o += kHidden * wb_stride(k);
  return o;
}
constexpr int kWbLds = wb_lds_offset(kLayers);  // 16384 elements

// Copies the forward weights W^T (global, row stride in_pad) and the
// backward copies W (layers 1..4) into LDS (row strides + 8).  The 16-byte
// chunks are numbered over both arrays; a thread of the kThreads-thread block
// loads all of its chunks before it stores any (one memory round trip).
template <int kThreads>
__device__ void load_weights(const uint16_t* __restrict__ wt, const uint16_t* __restrict__ wb,
                             uint16_t* s_wt, uint16_t* s_wb) {
  constexpr int kWtChunks = kParams / 8, kChunks = (kParams + kWbGlobal) / 8;
  constexpr int kPer = (kChunks + kThreads - 1) / kThreads;
  u32x4 v[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = (int)threadIdx.x + k * kThreads;
    if (e < kWtChunks)
      v[k] = reinterpret_cast<const u32x4*>(wt)[e];
    else if (e < kChunks)
      v[k] = reinterpret_cast<const u32x4*>(wb)[e - kWtChunks];
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int e = (int)threadIdx.x + k * kThreads;
    if (e < kWtChunks) {
      const int g = 8 * e;
      int l = 0;
      while (g >= kOffset[l + 1]) ++l;
      const int r = (g - kOffset[l]) / kInDim[l], c = (g - kOffset[l]) % kInDim[l];
      *reinterpret_cast<u32x4*>(s_wt + wt_lds_offset(l) + r * wt_stride(l) + c) = v[k];
    } else if (e < kChunks) {
      const int g = 8 * (e - kWtChunks);
      int l = 1;
      while (l < kLayers - 1 && g >= wb_global_offset(l + 1)) ++l;
      const int r = (g - wb_global_offset(l)) / wb_out(l), c = (g - wb_global_offset(l)) % wb_out(l);
      *reinterpret_cast<u32x4*>(s_wb + wb_lds_offset(l) + r * wb_stride(l) + c) = v[k];
    }
  }
}

// ---- inference with the weights and activations in registers ----
// Each layer computes Z^T = W^T X^T: the weights are the A operand (the
// hidden layers' held in VGPRs for the whole launch, layer 0's in LDS), the
// activations the B operand.  Lane (q, r)
// then holds D[4q .. 4q+3][r] = Z[row r][16 nt + 4q .. 4q+3], which after
// ReLU and BF16 rounding is its own B operand of the next layer if the K
// order of that layer's weights is permuted to match: K-step ks, lane q,
// slot j <-> hidden feature
//   h(ks, q, j) = 32 ks + 4 q + j           (j < 4)
//               = 32 ks + 16 + 4 q + j - 4  (j >= 4).
// Layer 0 uses e(ks, q, j) = 36 q + 8 ks + j (slots with 8 ks + j >= 36 are
// zero): lane q encodes inputs 3q .. 3q+2 of row r, the (sin, cos) pairs of
// 6 octaves, which the encoded-feature layout 12 i + 2 b + {0, 1} places at
// 36 q .. 36 q + 35.  Activations never leave the lane.
constexpr int kInferWaves = 4;

// FP32 -> BF16 bits, round to nearest even (the hardware conversion; the
// same bits as f2bf for every non-NaN value).
__device__ inline uint16_t bf16_rne(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }

__device__ inline uint32_t pack_bf16(float a, float b) {
  bf16x2 v;
  v[0] = (__bf16)a;  // round to nearest even
  v[1] = (__bf16)b;
  return __builtin_bit_cast(uint32_t, v);
}

// A fragment of hidden layer l (1..4): row 16 nt + r of W_l^T, K-step ks.
__device__ inline u32x4 wfrag_hidden(const uint16_t* __restrict__ wt, int l, int nt, int ks, int q,
                                     int r) {
  const uint16_t* row = wt + kOffset[l] + (nt * 16 + r) * kInDim[l] + 32 * ks + 4 * q;
  const uint2 a = *reinterpret_cast<const uint2*>(row);
  const uint2 b = *reinterpret_cast<const uint2*>(row + 16);
  return u32x4{a.x, a.y, b.x, b.y};
}

// A fragment of layer 0: row 16 nt + r of W_0^T at encoded features
// 36 q + 8 ks .. + 7 (only 4 of them for ks = 4).
__device__ inline u32x4 wfrag_in(const uint16_t* __restrict__ wt, int nt, int ks, int q, int r) {
  const uint32_t* row =
      reinterpret_cast<const uint32_t*>(wt + (nt * 16 + r) * kEncPad + 36 * q + 8 * ks);
  if (ks < 4) return u32x4{row[0], row[1], row[2], row[3]};
  return u32x4{row[0], row[1], 0u, 0u};
}

// sin(pi x) and cos(pi x) for the encoder's inputs (x in [0, 1]): x = n + r
// with n = rint(x), r in [-1/2, 1/2]; the Taylor series of sin(pi r) to r^13
// and of cos(pi r) to r^12 in Horner form (|error| < 3e-7 in FP32, against a
// BF16 rounding of 2^-9 for the features); (-1)^n for the integer part.
__device__ inline void sincospi_unit(float x, float& s, float& c) {
  const float n = rintf(x);
  const float r = x - n;
  const float z = r * r;
  float ps = 4.6630279e-4f;
  ps = fmaf(ps, z, -7.3704310e-3f);
  ps = fmaf(ps, z, 8.2145885e-2f);
  ps = fmaf(ps, z, -5.9926450e-1f);
  ps = fmaf(ps, z, 2.5501640e+0f);
  ps = fmaf(ps, z, -5.1677127e+0f);
  ps = fmaf(ps, z, 3.1415927e+0f);
  float pc = 1.9295743e-3f;
  pc = fmaf(pc, z, -2.5806891e-2f);
  pc = fmaf(pc, z, 2.3533063e-1f);
  pc = fmaf(pc, z, -1.3352628e+0f);
  pc = fmaf(pc, z, 4.0587121e+0f);
  pc = fmaf(pc, z, -4.9348021e+0f);
  pc = fmaf(pc, z, 1.0f);
  const float sign = ((int)n & 1) ? -1.0f : 1.0f;
  s = sign * (ps * r);
  c = sign * pc;
}

// Frequency encoding of one input p: e[2 b], e[2 b + 1] = (sin, cos)(pi 2^b p)
// for b = 0..5 (input i's features are 12 i + 2 b + {0, 1}): one sincospi,
// then the double-angle recurrence sin 2x = 2 sin x cos x,
// cos 2x = cos^2 x - sin^2 x.  Training and inference both encode with it.
__device__ inline void encode_input(float p, float* e) {
  float sn, cs;
  sincospi_unit(p, sn, cs);
#pragma unroll
  for (int b = 0; b < kBands; ++b) {
    e[2 * b] = sn;
    e[2 * b + 1] = cs;
    const float s2 = 2.0f * sn * cs;
    const float c2 = cs * cs - sn * sn;
    sn = s2;
    cs = c2;
  }
}

// ReLU as a signed integer max on the float's bits (v_max_i32 with 0): a
// float with its sign bit set becomes +0, others are kept; one instruction,
// where fmaxf under IEEE mode needs a quieting step as well.
__device__ inline float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

// ReLU + BF16 of two D tiles (N tiles 2 ks, 2 ks + 1) into the B operand of K-step ks.
__device__ inline u32x4 relu_pack(f32x4 lo, f32x4 hi) {
  return u32x4{pack_bf16(relu(lo[0]), relu(lo[1])), pack_bf16(relu(lo[2]), relu(lo[3])),
               pack_bf16(relu(hi[0]), relu(hi[1])), pack_bf16(relu(hi[2]), relu(hi[3]))};
}

__device__ inline bf16x8 bf(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(const float* __restrict__ feat,
                                                               int64_t n,
                                                               const uint16_t* __restrict__ wt,
                                                               float* __restrict__ out) {
  // layer-0 A fragments (80 VGPRs' worth) in LDS, [nt][ks][lane]: one
  // conflict-free 16-byte read per lane per MFMA; the hidden layers' stay in
  // registers
  __shared__ u32x4 s_w0[4 * 5 * 64];
  for (int e = threadIdx.x; e < 4 * 5 * 64; e += blockDim.x) {
    const int l = e & 63, f = e >> 6;
    s_w0[e] = wfrag_in(wt, f / 5, f % 5, l >> 4, l & 15);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int64_t tiles = (n + 15) / 16;
  const int64_t first = (int64_t)blockIdx.x * kInferWaves + (threadIdx.x >> 6);
  const int64_t stride = (int64_t)gridDim.x * kInferWaves;
  if (first >= tiles) return;
  u32x4 w1[4][2], w2[4][2], w3[4][2], w4[2];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      w1[nt][ks] = wfrag_hidden(wt, 1, nt, ks, q, r);
      w2[nt][ks] = wfrag_hidden(wt, 2, nt, ks, q, r);
      w3[nt][ks] = wfrag_hidden(wt, 3, nt, ks, q, r);
    }
  }
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) w4[ks] = wfrag_hidden(wt, 4, 0, ks, q, r);

  // inputs 3q .. 3q+2 of row r of a tile
  auto load_inputs = [&](int64_t t, float* p) {
    const int64_t row = t * 16 + r;
#pragma unroll
    for (int k = 0; k < 3; ++k) p[k] = row < n ? feat[row * kInputs + 3 * q + k] : 0.0f;
  };
  float p[3];
  load_inputs(first, p);
  for (int64_t t = first; t < tiles; t += stride) {
    float pn[3] = {0.0f, 0.0f, 0.0f};
    if (t + stride < tiles) load_inputs(t + stride, pn);
    // frequency encoding of the lane's three inputs
    float e[40];
#pragma unroll
    for (int k = 0; k < 3; ++k) encode_input(p[k], e + 12 * k);
    e[36] = e[37] = e[38] = e[39] = 0.0f;
    u32x4 x[5];
#pragma unroll
    for (int ks = 0; ks < 5; ++ks)
      x[ks] = u32x4{pack_bf16(e[8 * ks], e[8 * ks + 1]), pack_bf16(e[8 * ks + 2], e[8 * ks + 3]),
                    pack_bf16(e[8 * ks + 4], e[8 * ks + 5]), pack_bf16(e[8 * ks + 6], e[8 * ks + 7])};
    // the lane's LDS offset, opaque to the compiler on every tile: the layer-0
    // fragments are read here each time instead of being hoisted into
    // registers (which would push the accumulators into AGPRs)
    int w0_lane = lane;
    asm volatile("" : "+v"(w0_lane));
    const u32x4* w0 = s_w0 + w0_lane;
    f32x4 z[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      z[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 5; ++ks) z[nt] = mfma(bf(w0[(nt * 5 + ks) * 64]), bf(x[ks]), z[nt]);
    }
    u32x4 h[2];
    auto hidden = [&](u32x4 (&w)[4][2]) {
      h[0] = relu_pack(z[0], z[1]);
      h[1] = relu_pack(z[2], z[3]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        z[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) z[nt] = mfma(bf(w[nt][ks]), bf(h[ks]), z[nt]);
      }
    };
    hidden(w1);
    hidden(w2);
    hidden(w3);
    h[0] = relu_pack(z[0], z[1]);
    h[1] = relu_pack(z[2], z[3]);
    f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
    y = mfma(bf(w4[0]), bf(h[0]), y);
    y = mfma(bf(w4[1]), bf(h[1]), y);
    const int64_t row = t * 16 + r;
    if (q == 0 && row < n) {
      out[row * kOut] = y[0];
      out[row * kOut + 1] = y[1];
      out[row * kOut + 2] = y[2];
    }
    p[0] = pn[0];
    p[1] = pn[1];
    p[2] = pn[2];
  }
}

// Blocks of the inference launch: enough waves for every tile, at most three
// 4-wave blocks per CU (the kernel's occupancy; each wave loads its 27 KB of
// hidden-layer fragments and each block 20 KB of layer-0 fragments once).
int infer_blocks(int cus, int64_t n) {
  const int64_t tiles = (n + 15) / 16;
  return (int)std::max<int64_t>(1, std::min<int64_t>((tiles + kInferWaves - 1) / kInferWaves,
                                                     3 * (int64_t)cus));
}

// Frequency encoding of a training call's rows, once for all of its steps:
// thread (row m, input i) writes the input's 12 features into the
// sample-major copy xs (row stride kEncPad; the layer-0 A fragments of
// k_mlp_train) and the feature-major copy xf (row stride npad; layer 0's
// input for the weight gradients).  Rows n <= m < npad encode p = 0 (their
// loss gradient is zero).  Threads of input 0 also write features 144..159.
constexpr int kEncRows = 64;  // rows per block; blockDim = kEncRows * kInputs
__global__ __launch_bounds__(kEncRows * kInputs) void k_mlp_encode(
    const float* __restrict__ feat, int feat_stride, int64_t n, int64_t npad,
    uint16_t* __restrict__ xs, uint16_t* __restrict__ xf) {
  const int i = threadIdx.x / kEncRows;
  const int64_t m = (int64_t)blockIdx.x * kEncRows + threadIdx.x % kEncRows;
  if (m >= npad) return;
  const float p = m < n ? feat[m * feat_stride + i] : 0.0f;
  float e[2 * kBands];
  encode_input(p, e);
  for (int j = 0; j < 2 * kBands; ++j) {
    const uint16_t h = bf16_rne(e[j]);
    xs[m * kEncPad + 2 * kBands * i + j] = h;
    xf[(int64_t)(2 * kBands * i + j) * npad + m] = h;
  }
  if (i == 0) {
    for (int f = kEnc; f < kEncPad; ++f) {
      xs[m * kEncPad + f] = 0;
      xf[(int64_t)f * npad + m] = 0;
    }
  }
}

// Waves per training workgroup (two per SIMD): they share one LDS copy of
// the weights in both layouts (84 KB) and hold 6.9 KB each (two activation
// tiles used in turn through the forward pass and one dZ tile), 140 KB in
// all, so one workgroup per CU; the grid is persistent (at most one
// workgroup per CU), each wave looping over 16-row tiles.
constexpr int kTrainWaves = 8;
constexpr int kLossSlots = 64;  // one loss per step of a training call
constexpr int kMaxSplit = 64;   // K splits of the weight-gradient GEMMs (partials, no atomics)

// One feature of a 16-row tile (src[i * stride], i < 16) to 16 consecutive
// BF16 values at dst (32-byte aligned), as two 16-byte stores.
__device__ inline void store16(const uint16_t* src, int stride, uint16_t* dst) {
  u32x4 v[2];
  for (int h = 0; h < 2; ++h)
    for (int j = 0; j < 4; ++j)
      v[h][j] = (uint32_t)src[(8 * h + 2 * j) * stride] | ((uint32_t)src[(8 * h + 2 * j + 1) * stride] << 16);
  reinterpret_cast<u32x4*>(dst)[0] = v[0];
  reinterpret_cast<u32x4*>(dst)[1] = v[1];
}
// Four consecutive BF16 values to dst (8-byte aligned), one store.
__device__ inline void store4(const uint16_t* v, uint16_t* dst) {
  *reinterpret_cast<uint2*>(dst) = make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16),
                                             (uint32_t)v[2] | ((uint32_t)v[3] << 16));
}

// Forward + relative-L2 loss + backward data pass for 16 samples per wave
// tile.  Layer 0's input comes from xs (k_mlp_encode); writes feature-major
// (row = feature, column = sample, row stride npad):
//   act_l (l = 1..4): the input of layer l (A1..A4), BF16
//   dz_l  (l = 0..4): dLoss/dZ_l, BF16 (dz_4 has 16 rows, 3 used)
// and the wave's share of the loss into loss_part[wave of the grid].  The
// ReLU masks of the backward pass stay in registers: lane (q, r) holds
// the D elements [4q + i][16 nt + r] of every layer, in the forward and the
// backward pass alike.
__global__ __launch_bounds__(64 * kTrainWaves) void k_mlp_train(
    const uint16_t* __restrict__ xs, const float* __restrict__ target, int64_t n, int64_t npad,
    float inv_count, const uint16_t* __restrict__ wt, const uint16_t* __restrict__ wb,
    uint16_t* __restrict__ act, uint16_t* __restrict__ dz, float* __restrict__ loss_part) {
  __shared__ __attribute__((aligned(16))) uint16_t s_wt[kWtLds];
  __shared__ __attribute__((aligned(16))) uint16_t s_wb[kWbLds];
  __shared__ __attribute__((aligned(16))) uint16_t s_a[kTrainWaves][2][16 * kHStride];
  __shared__ __attribute__((aligned(16))) uint16_t s_g[kTrainWaves][16 * kGStride];
  load_weights<64 * kTrainWaves>(wt, wb, s_wt, s_wb);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = lane >> 4, r = lane & 15;
  uint16_t* G = s_g[wave];
  // feature-major global buffers
  uint16_t* act_l[kLayers];
  uint16_t* dz_l[kLayers];
  {
    int64_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
      act_l[l] = act + off;
      off += (int64_t)kInDim[l] * npad;
    }
    off = 0;
    for (int l = 0; l < kLayers; ++l) {
      dz_l[l] = dz + off;
      off += (int64_t)kOutDim[l] * npad;
    }
  }
  float loss_acc = 0.0f;
  const int64_t tiles = npad / 16;
  const int64_t t_step = (int64_t)gridDim.x * kTrainWaves;
  // layer 0's A fragments straight from the sample-major encoding, one tile ahead
  bf16x8 x[kEncPad / 32], xn[kEncPad / 32];
  auto load_x = [&](int64_t t, bf16x8 (&dst)[kEncPad / 32]) {
#pragma unroll
    for (int ks = 0; ks < kEncPad / 32; ++ks) dst[ks] = ld8(xs + (t * 16 + r) * kEncPad + ks * 32 + 8 * q);
  };
  const int64_t t0 = (int64_t)blockIdx.x * kTrainWaves + wave;
  if (t0 < tiles) load_x(t0, xn);
  for (int64_t t = t0; t < tiles; t += t_step) {
    const int64_t base = t * 16;
#pragma unroll
    for (int ks = 0; ks < kEncPad / 32; ++ks) x[ks] = xn[ks];
    if (t + t_step < tiles) load_x(t + t_step, xn);
    // mask bit 16 l + 4 nt + i: [A_{l+1}[4q + i][16 nt + r] > 0], l = 0..3
    uint64_t mask = 0;
    f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int l = 0; l < kLayers; ++l) {
      const uint16_t* w = s_wt + wt_lds_offset(l);
      const uint16_t* in = s_a[wave][(l + 1) & 1];  // output of layer l - 1
      uint16_t* out = s_a[wave][l & 1];
      const int ntiles = kOutDim[l] / 16;
      for (int nt = 0; nt < ntiles; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        if (l == 0) {
#pragma unroll
          for (int ks = 0; ks < kEncPad / 32; ++ks)
            acc = mfma(x[ks], ld8(w + (nt * 16 + r) * wt_stride(0) + ks * 32 + 8 * q), acc);
        } else {
#pragma unroll
          for (int ks = 0; ks < kHidden / 32; ++ks)
            acc = mfma(ld8(in + r * kHStride + ks * 32 + 8 * q),
                       ld8(w + (nt * 16 + r) * wt_stride(l) + ks * 32 + 8 * q), acc);
        }
        if (l < kLayers - 1) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const uint16_t h = bf16_rne(relu(acc[i]));
            out[(4 * q + i) * kHStride + nt * 16 + r] = h;
            if (h != 0) mask |= 1ull << (16 * l + 4 * nt + i);
          }
        } else {
          y = acc;
        }
      }
      __builtin_amdgcn_wave_barrier();
      // the layer's output is the input of layer l + 1: feature-major to HBM
      if (l < kLayers - 1) store16(out + lane, kHStride, act_l[l + 1] + (int64_t)lane * npad + base);
      __builtin_amdgcn_wave_barrier();
    }
    // dLoss/dY for L = mean over samples and channels of (y - t)^2 / (y^2 + 0.01),
    // the denominator held constant (relative L2)
    uint16_t gq[4];
    for (int i = 0; i < 4; ++i) {
      const int64_t m = base + 4 * q + i;
      float g = 0.0f;
      if (r < kOut && m < n) {
        const float tv = target[m * kOut + r];
        if (isfinite(tv)) {  // non-finite targets mark unused samples
          const float d = y[i] - tv;
          const float den = y[i] * y[i] + 0.01f;
          g = 2.0f * d / den * inv_count;
          loss_acc += d * d / den * inv_count;
        }
      }
      gq[i] = bf16_rne(g);
      G[(4 * q + i) * kGStride + r] = gq[i];
    }
    store4(gq, dz_l[kLayers - 1] + (int64_t)r * npad + base + 4 * q);
    for (int e = lane; e < 16 * 16; e += 64) G[(e / 16) * kGStride + 16 + e % 16] = 0;  // K pad 16..31
    __builtin_amdgcn_wave_barrier();
    // backward data pass: dZ_{l-1} = (dZ_l W_l^T) * [A_l > 0], l = 4 .. 1
    for (int l = kLayers - 1; l >= 1; --l) {
      const int K = wb_out(l);
      const uint16_t* wl = s_wb + wb_lds_offset(l);
      f32x4 acc[4];
      for (int nt = 0; nt < 4; ++nt) {
        acc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < K / 32; ++ks) {
          const bf16x8 a = ld8(G + r * kGStride + ks * 32 + 8 * q);
          const bf16x8 b = ld8(wl + (nt * 16 + r) * wb_stride(l) + ks * 32 + 8 * q);
          acc[nt] = mfma(a, b, acc[nt]);
        }
      }
      __builtin_amdgcn_wave_barrier();
      for (int nt = 0; nt < 4; ++nt) {
        uint16_t gq[4];
        const int col = nt * 16 + r;
        for (int i = 0; i < 4; ++i) {
          const float g = (mask >> (16 * (l - 1) + 4 * nt + i)) & 1u ? acc[nt][i] : 0.0f;
          gq[i] = bf16_rne(g);
          G[(4 * q + i) * kGStride + col] = gq[i];
        }
        store4(gq, dz_l[l - 1] + (int64_t)col * npad + base + 4 * q);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  // the wave's loss: one plain store per wave (summed by k_mlp_adam / k_mlp_reduce)
  for (int o = 32; o > 0; o >>= 1) loss_acc += __shfl_xor(loss_acc, o, 64);
  if (lane == 0) loss_part[blockIdx.x * kTrainWaves + wave] = loss_acc;
}

// dW_l^T (out_pad x in_pad) = dZ_l (out_pad x npad) * in_l^T (npad x in_pad)
// for every layer in one launch.  One wave per (layer, column block, K
// split): a column block is all of the layer's output rows (4 tiles of 16, 1
// for the last layer) times kWgIT input tiles, so each dZ fragment serves
// kWgIT MFMAs and each input fragment up to 4; waves [first[l], first[l + 1])
// belong to layer l.  Split s writes its partial sums to part[s * kParams + p]
// (plain stores, every p of the block).
constexpr int kWgIT = 2;
static_assert(kEncPad % (16 * kWgIT) == 0 && kHidden % (16 * kWgIT) == 0, "whole column blocks");
struct WgradPlan {
  int64_t off_dz[kLayers], off_in[kLayers];
  int first[kLayers + 1];
  int ksplit;
};
__global__ __launch_bounds__(256) void k_mlp_wgrad(const uint16_t* __restrict__ dz_all,
                                                   const uint16_t* __restrict__ in_all, int64_t npad,
                                                   WgradPlan plan, float* __restrict__ part) {
  int wave = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (wave >= plan.first[kLayers]) return;
  int l = 0;
  while (wave >= plan.first[l + 1]) ++l;
  wave -= plan.first[l];
  const uint16_t* __restrict__ dz = dz_all + plan.off_dz[l];
  const uint16_t* __restrict__ in = in_all + plan.off_in[l];
  const int in_pad = kInDim[l], ksplit = plan.ksplit;
  const int n_ot = kOutDim[l] / 16;  // wave-uniform: 4, or 1 for the last layer
  const int q = lane >> 4, r = lane & 15;
  const int blk = wave / ksplit, split = wave % ksplit;
  float* __restrict__ dwt = part + (int64_t)split * kParams + kOffset[l];
  const int64_t steps = npad / 32;
  const int64_t s0 = steps * split / ksplit, s1 = steps * (split + 1) / ksplit;
  f32x4 acc[4][kWgIT];
#pragma unroll
  for (int o = 0; o < 4; ++o)
#pragma unroll
    for (int j = 0; j < kWgIT; ++j) acc[o][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const uint16_t* arow = dz + (int64_t)r * npad + 8 * q;
  const uint16_t* brow = in + (int64_t)(blk * kWgIT * 16 + r) * npad + 8 * q;
  auto step = [&](int64_t s, int nu) {
    // nu K-steps from s: every load first, then the MFMAs
    bf16x8 a[4][4], b[4][kWgIT];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (u < nu) {
#pragma unroll
        for (int o = 0; o < 4; ++o)
          if (o < n_ot) a[u][o] = ld8(arow + (int64_t)o * 16 * npad + 32 * (s + u));
#pragma unroll
        for (int j = 0; j < kWgIT; ++j) b[u][j] = ld8(brow + (int64_t)j * 16 * npad + 32 * (s + u));
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (u < nu) {
#pragma unroll
        for (int o = 0; o < 4; ++o)
          if (o < n_ot) {
#pragma unroll
            for (int j = 0; j < kWgIT; ++j) acc[o][j] = mfma(a[u][o], b[u][j], acc[o][j]);
          }
      }
    }
  };
  int64_t s = s0;
  for (; s + 4 <= s1; s += 4) step(s, 4);
  if (s < s1) step(s, (int)(s1 - s));
#pragma unroll
  for (int o = 0; o < 4; ++o)
    if (o < n_ot) {
#pragma unroll
      for (int j = 0; j < kWgIT; ++j)
        for (int i = 0; i < 4; ++i)
          dwt[(int64_t)(o * 16 + 4 * q + i) * in_pad + (blk * kWgIT + j) * 16 + r] = acc[o][j][i];
    }
}

// Sum of n floats at p (the block's 256 threads; result valid in thread 0).
__device__ float block_sum(const float* __restrict__ p, int n, float* s_red) {
  float a = 0.0f;
  for (int k = threadIdx.x; k < n; k += 256) a += p[k];
  for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = a;
  __syncthreads();
  return s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// The weight gradient of p from the K-split partials: each of four
// neighbouring lanes (k = lane % 4) sums the splits k, k + 4, ... in order,
// then the four sums are added in a fixed tree; every lane of the group
// returns the total.  kParamsPerBlock weights per 256-thread block.
constexpr int kParamsPerBlock = 64;
static_assert(kParams % kParamsPerBlock == 0, "whole blocks of weights");
__device__ inline float grad_of(const float* __restrict__ part, int ksplit, int p) {
  const int k = threadIdx.x & 3;
  float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int s = k;
  for (; s + 12 < ksplit; s += 16) {
#pragma unroll
    for (int u = 0; u < 4; ++u) g[u] += part[(int64_t)(s + 4 * u) * kParams + p];
  }
  for (; s < ksplit; s += 4) g[0] += part[(int64_t)s * kParams + p];
  float t = (g[0] + g[1]) + (g[2] + g[3]);
  t += __shfl_xor(t, 1, 64);
  t += __shfl_xor(t, 2, 64);
  return t;
}

// Gradients and loss of one k_mlp_train + k_mlp_wgrad pass into g[0..kParams)
// and *loss (lisa_cache_gradients).
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int ksplit,
                                                    const float* __restrict__ loss_part,
                                                    int n_loss, float* __restrict__ g,
                                                    float* __restrict__ loss) {
  __shared__ float s_red[4];
  if (blockIdx.x == 0) {
    const float l = block_sum(loss_part, n_loss, s_red);
    if (threadIdx.x == 0) *loss = l;
  }
  const int p = blockIdx.x * kParamsPerBlock + (threadIdx.x >> 2);
  const float gr = grad_of(part, ksplit, p);  // kParams is a multiple of kParamsPerBlock
  if ((threadIdx.x & 3) == 0) g[p] = gr;
}

// Adam (Kingma & Ba) on the FP32 master weights (packed W^T layout) with the
// gradients summed from the K-split partials; padded entries stay zero.
// Writes the BF16 forward (W^T) and backward (W) copies; block 0 also sums
// the step's loss into *loss.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ w, float* __restrict__ m,
                                                  float* __restrict__ v,
                                                  const float* __restrict__ part, int ksplit,
                                                  const float* __restrict__ loss_part, int n_loss,
                                                  float* __restrict__ loss, float lr, float b1,
                                                  float b2, float eps, float c1, float c2,
                                                  uint16_t* __restrict__ wt,
                                                  uint16_t* __restrict__ wb) {
  __shared__ float s_red[4];
  if (blockIdx.x == 0) {
    const float l = block_sum(loss_part, n_loss, s_red);
    if (threadIdx.x == 0) *loss = l;
  }
  const int p = blockIdx.x * kParamsPerBlock + (threadIdx.x >> 2);
  const float gr = grad_of(part, ksplit, p);  // kParams is a multiple of kParamsPerBlock
  if ((threadIdx.x & 3) != 0) return;
  int l = 0;
  while (p >= kOffset[l + 1]) ++l;
  const int o = (p - kOffset[l]) / kInDim[l], i = (p - kOffset[l]) % kInDim[l];
  if (o >= real_out(l) || i >= real_in(l)) return;
  const float mm = b1 * m[p] + (1.0f - b1) * gr;
  const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
  m[p] = mm;
  v[p] = vv;
  const float ww = w[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  w[p] = ww;
  const uint16_t h = f2bf(ww);
  wt[p] = h;
  if (l >= 1) wb[wb_global_offset(l) + i * wb_out(l) + o] = h;
}

}  // namespace

// ===========================================================================
// C ABI: lisa_cache_*
// ===========================================================================
struct lisa_cache {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *loss = nullptr;
  float* part = nullptr;       // kMaxSplit x kParams weight-gradient partials
  float* loss_part = nullptr;  // kLossSlots x (cus * kTrainWaves) per-wave losses
  uint16_t *wt = nullptr, *wb = nullptr;
  float lr = 1e-3f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int64_t step = 0;
  // training scratch (grown on demand)
  int64_t cap = 0;
  uint16_t *act = nullptr, *dz = nullptr, *xs = nullptr;
  float *feat = nullptr, *tgt = nullptr, *out = nullptr;
  int64_t io_cap = 0;
  float last_ms = 0.0f;
  int cus = 256;  // compute units of the device (inference grid)
};

namespace {

lisa_status cfail(lisa_cache* c, lisa_status code, const std::string& msg) {
  if (c)
    c->err = msg;
  else
    lisa::set_global_error(msg);
  return code;
}

#define CHIP(c, expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return cfail((c), LISA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// He-uniform initial weights (bound sqrt(6 / fan_in)) from a splitmix64 stream.
void init_weights(std::vector<float>& w, uint64_t seed) {
  w.assign(kParams, 0.0f);
  uint64_t s = seed;
  auto next = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / (double)(1ull << 53);
  };
  for (int l = 0; l < kLayers; ++l) {
    const float bound = std::sqrt(6.0f / (float)real_in(l));
    for (int o = 0; o < real_out(l); ++o)
      for (int i = 0; i < real_in(l); ++i)
        w[kOffset[l] + o * kInDim[l] + i] = (float)((2.0 * next() - 1.0) * bound);
  }
}

lisa_status upload_weights(lisa_cache* c, const std::vector<float>& w) {
  std::vector<uint16_t> wt(kParams), wb(kWbGlobal, 0);
  for (int l = 0; l < kLayers; ++l)
    for (int o = 0; o < kOutDim[l]; ++o)
      for (int i = 0; i < kInDim[l]; ++i) {
        const int p = kOffset[l] + o * kInDim[l] + i;
        const bool real = o < real_out(l) && i < real_in(l);
        const float x = real ? w[p] : 0.0f;
        wt[p] = f2bf(x);
        if (l >= 1) wb[wb_global_offset(l) + i * wb_out(l) + o] = f2bf(x);
      }
  std::vector<float> wm(w);
  for (int l = 0; l < kLayers; ++l)
    for (int o = 0; o < kOutDim[l]; ++o)
      for (int i = 0; i < kInDim[l]; ++i)
        if (!(o < real_out(l) && i < real_in(l))) wm[kOffset[l] + o * kInDim[l] + i] = 0.0f;
  CHIP(c, hipMemcpy(c->w, wm.data(), 4 * kParams, hipMemcpyHostToDevice));
  CHIP(c, hipMemcpy(c->wt, wt.data(), 2 * kParams, hipMemcpyHostToDevice));
  CHIP(c, hipMemcpy(c->wb, wb.data(), 2 * kWbGlobal, hipMemcpyHostToDevice));
  return LISA_OK;
}

lisa_status ensure_io(lisa_cache* c, int64_t n) {
  if (n <= c->io_cap) return LISA_OK;
  (void)hipFree(c->feat);
  (void)hipFree(c->tgt);
  (void)hipFree(c->out);
  c->feat = c->tgt = c->out = nullptr;
  c->io_cap = 0;
  const int64_t cap = std::max<int64_t>(n, 1 << 16);
  CHIP(c, hipMalloc(reinterpret_cast<void**>(&c->feat), 4 * kInputs * (size_t)cap));
  CHIP(c, hipMalloc(reinterpret_cast<void**>(&c->tgt), 4 * kOut * (size_t)cap));
  CHIP(c, hipMalloc(reinterpret_cast<void**>(&c->out), 4 * kOut * (size_t)cap));
  c->io_cap = cap;
  return LISA_OK;
}

lisa_status ensure_train(lisa_cache* c, int64_t npad) {
  if (npad <= c->cap) return LISA_OK;
  (void)hipFree(c->act);
  (void)hipFree(c->dz);
  (void)hipFree(c->xs);
  c->act = c->dz = c->xs = nullptr;
  c->cap = 0;
  int64_t rows_act = 0, rows_dz = 0;
  for (int l = 0; l < kLayers; ++l) {
    rows_act += kInDim[l];
    rows_dz += kOutDim[l];
  }
  CHIP(c, hipMalloc(reinterpret_cast<void**>(&c->act), 2 * (size_t)rows_act * npad));
  CHIP(c, hipMalloc(reinterpret_cast<void**>(&c->dz), 2 * (size_t)rows_dz * npad));
  CHIP(c, hipMalloc(reinterpret_cast<void**>(&c->xs), 2 * (size_t)kEncPad * npad));
  c->cap = npad;
  return LISA_OK;
}

// Encodes n device-resident samples (features with row stride feat_stride)
// for the steps that follow (k_mlp_encode into xs and layer 0's feature-major
// input).
lisa_status encode_device(lisa_cache* c, const float* d_feat, int feat_stride, int64_t n) {
  const int64_t npad = (n + 31) / 32 * 32;
  lisa_status rc = ensure_train(c, npad);
  if (rc != LISA_OK) return rc;
  k_mlp_encode<<<(int)((npad + kEncRows - 1) / kEncRows), kEncRows * kInputs, 0, c->stream>>>(
      d_feat, feat_stride, n, npad, c->xs, c->act);
  CHIP(c, hipGetLastError());
  return LISA_OK;
}

// One gradient pass over the n encoded samples (targets n x 3): the weight
// gradients as ksplit partials in c->part, the per-wave losses in
// loss_part[0 .. *n_loss).  Two launches: k_mlp_train, k_mlp_wgrad.
lisa_status grads_step(lisa_cache* c, const float* d_tgt, int64_t n, float* loss_part,
                       int* ksplit, int* n_loss) {
  const int64_t npad = (n + 31) / 32 * 32;
  hipStream_t st = c->stream;
  const int64_t tiles = npad / 16;
  const int blocks = (int)std::min<int64_t>((tiles + kTrainWaves - 1) / kTrainWaves, c->cus);
  k_mlp_train<<<blocks, 64 * kTrainWaves, 0, st>>>(c->xs, d_tgt, n, npad, 1.0f / (3.0f * (float)n),
                                                    c->wt, c->wb, c->act, c->dz, loss_part);
  WgradPlan plan;
  plan.ksplit = (int)std::max<int64_t>(1, std::min<int64_t>(npad / 32 / 16, kMaxSplit));
  int64_t off_a = 0, off_d = 0;
  plan.first[0] = 0;
  for (int l = 0; l < kLayers; ++l) {
    plan.off_dz[l] = off_d;
    plan.off_in[l] = off_a;
    plan.first[l + 1] = plan.first[l] + kInDim[l] / 16 / kWgIT * plan.ksplit;
    off_a += (int64_t)kInDim[l] * npad;
    off_d += (int64_t)kOutDim[l] * npad;
  }
  k_mlp_wgrad<<<(plan.first[kLayers] + 3) / 4, 256, 0, st>>>(c->dz, c->act, npad, plan, c->part);
  CHIP(c, hipGetLastError());
  *ksplit = plan.ksplit;
  *n_loss = blocks * kTrainWaves;
  return LISA_OK;
}

}  // namespace

extern "C" {

lisa_status lisa_cache_create(lisa_cache** out, int32_t device, uint64_t seed, float lr) {
  if (!out) return cfail(nullptr, LISA_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (!(lr > 0.0f)) return cfail(nullptr, LISA_ERR_INVALID_ARG, "lr must be > 0");
  lisa_cache* c = new lisa_cache;
  c->device = device;
  c->lr = lr;
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreate(&c->stream)) != hipSuccess ||
      (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->w), 4 * kParams)) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->m), 4 * kParams)) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->v), 4 * kParams)) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->g), 4 * (kParams + kLossSlots))) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->part), 4 * (size_t)kMaxSplit * kParams)) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->wt), 2 * kParams)) != hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->wb), 2 * kWbGlobal)) != hipSuccess ||
      (e = hipMemset(c->m, 0, 4 * kParams)) != hipSuccess ||
      (e = hipMemset(c->v, 0, 4 * kParams)) != hipSuccess ||
      (e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device)) !=
          hipSuccess ||
      (e = hipMalloc(reinterpret_cast<void**>(&c->loss_part),
                     4 * (size_t)kLossSlots * c->cus * kTrainWaves)) != hipSuccess) {
    const std::string msg = std::string("lisa_cache_create: ") + hipGetErrorString(e);
    lisa_cache_destroy(c);
    return cfail(nullptr, LISA_ERR_HIP, msg);
  }
  c->loss = c->g + kParams;  // the per-step losses follow the gradients
  std::vector<float> w;
  init_weights(w, seed);
  lisa_status rc = upload_weights(c, w);
  if (rc != LISA_OK) {
    const std::string msg = c->err;
    lisa_cache_destroy(c);
    return cfail(nullptr, rc, msg);
  }
  *out = c;
  return LISA_OK;
}

void lisa_cache_destroy(lisa_cache* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : {(void*)c->w, (void*)c->m, (void*)c->v, (void*)c->g, (void*)c->part,
                  (void*)c->loss_part, (void*)c->wt, (void*)c->wb, (void*)c->act, (void*)c->dz,
                  (void*)c->xs, (void*)c->feat, (void*)c->tgt, (void*)c->out})
    (void)hipFree(p);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* lisa_cache_last_error(const lisa_cache* c) { return c ? c->err.c_str() : lisa::global_error(); }

lisa_status lisa_cache_set_weights(lisa_cache* c, const float* w) {
  if (!c || !w) return cfail(c, LISA_ERR_INVALID_ARG, "NULL argument");
  CHIP(c, hipSetDevice(c->device));
  std::vector<float> v(w, w + kParams);
  lisa_status rc = upload_weights(c, v);
  if (rc != LISA_OK) return rc;
  CHIP(c, hipMemset(c->m, 0, 4 * kParams));
  CHIP(c, hipMemset(c->v, 0, 4 * kParams));
  c->step = 0;
  return LISA_OK;
}

lisa_status lisa_cache_get_weights(lisa_cache* c, float* w) {
  if (!c || !w) return cfail(c, LISA_ERR_INVALID_ARG, "NULL argument");
  CHIP(c, hipSetDevice(c->device));
  CHIP(c, hipStreamSynchronize(c->stream));
  CHIP(c, hipMemcpy(w, c->w, 4 * kParams, hipMemcpyDeviceToHost));
  return LISA_OK;
}

lisa_status lisa_cache_infer(lisa_cache* c, const float* feat, int64_t n, float* out) {
  if (!c) return cfail(nullptr, LISA_ERR_INVALID_ARG, "cache is NULL");
  if (n < 0 || (n > 0 && (!feat || !out))) return cfail(c, LISA_ERR_INVALID_ARG, "bad arrays");
  if (n == 0) return LISA_OK;
  CHIP(c, hipSetDevice(c->device));
  lisa_status rc = ensure_io(c, n);
  if (rc != LISA_OK) return rc;
  CHIP(c, hipMemcpy(c->feat, feat, 4 * kInputs * (size_t)n, hipMemcpyHostToDevice));
  CHIP(c, hipEventRecord(c->ev0, c->stream));
  k_mlp_infer<<<infer_blocks(c->cus, n), 64 * kInferWaves, 0, c->stream>>>(c->feat, n, c->wt, c->out);
  CHIP(c, hipGetLastError());
  CHIP(c, hipEventRecord(c->ev1, c->stream));
  CHIP(c, hipStreamSynchronize(c->stream));
  CHIP(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  CHIP(c, hipMemcpy(out, c->out, 4 * kOut * (size_t)n, hipMemcpyDeviceToHost));
  return LISA_OK;
}

lisa_status lisa_cache_gradients(lisa_cache* c, const float* feat, const float* target, int64_t n,
                                 float* grad, float* loss) {
  if (!c) return cfail(nullptr, LISA_ERR_INVALID_ARG, "cache is NULL");
  if (n <= 0 || !feat || !target) return cfail(c, LISA_ERR_INVALID_ARG, "bad arrays");
  CHIP(c, hipSetDevice(c->device));
  lisa_status rc = ensure_io(c, n);
  if (rc != LISA_OK) return rc;
  CHIP(c, hipMemcpy(c->feat, feat, 4 * kInputs * (size_t)n, hipMemcpyHostToDevice));
  CHIP(c, hipMemcpy(c->tgt, target, 4 * kOut * (size_t)n, hipMemcpyHostToDevice));
  rc = encode_device(c, c->feat, kInputs, n);
  if (rc != LISA_OK) return rc;
  int ksplit = 0, n_loss = 0;
  rc = grads_step(c, c->tgt, n, c->loss_part, &ksplit, &n_loss);
  if (rc != LISA_OK) return rc;
  k_mlp_reduce<<<kParams / kParamsPerBlock, 256, 0, c->stream>>>(c->part, ksplit, c->loss_part, n_loss,
                                                             c->g, c->loss);
  CHIP(c, hipGetLastError());
  CHIP(c, hipStreamSynchronize(c->stream));
  if (grad) CHIP(c, hipMemcpy(grad, c->g, 4 * kParams, hipMemcpyDeviceToHost));
  if (loss) CHIP(c, hipMemcpy(loss, c->loss, 4, hipMemcpyDeviceToHost));
  return LISA_OK;
}

}  // extern "C"

// `steps` Adam steps on the same n device-resident samples (features with
// row stride feat_stride, targets n x 3), enqueued back to back; one wait at
// the end; *loss = the last step's loss, last_ms = all steps (HIP events).
lisa_status lisa::cache_train_device(lisa_cache* c, const float* d_feat, int feat_stride,
                                     const float* d_tgt, int64_t n, float* loss, int steps) {
  if (!c) return cfail(nullptr, LISA_ERR_INVALID_ARG, "cache is NULL");
  if (n <= 0 || steps <= 0) return cfail(c, LISA_ERR_INVALID_ARG, "n and steps must be > 0");
  if (steps > kLossSlots) return cfail(c, LISA_ERR_INVALID_ARG, "at most 64 training steps per call");
  CHIP(c, hipSetDevice(c->device));
  CHIP(c, hipEventRecord(c->ev0, c->stream));
  // the encoding once, then three launches per step
  lisa_status rc = encode_device(c, d_feat, feat_stride, n);
  if (rc != LISA_OK) return rc;
  for (int k = 0; k < steps; ++k) {
    float* lp = c->loss_part + (size_t)k * c->cus * kTrainWaves;
    int ksplit = 0, n_loss = 0;
    rc = grads_step(c, d_tgt, n, lp, &ksplit, &n_loss);
    if (rc != LISA_OK) return rc;
    ++c->step;
    const float c1 = 1.0f - std::pow(c->b1, (float)c->step),
                c2 = 1.0f - std::pow(c->b2, (float)c->step);
    k_mlp_adam<<<kParams / kParamsPerBlock, 256, 0, c->stream>>>(
        c->w, c->m, c->v, c->part, ksplit, lp, n_loss, c->loss + k, c->lr, c->b1, c->b2, c->eps, c1,
        c2, c->wt, c->wb);
    CHIP(c, hipGetLastError());
  }
  CHIP(c, hipEventRecord(c->ev1, c->stream));
  if (loss) {
    CHIP(c, hipMemcpyAsync(loss, c->loss + (steps - 1), 4, hipMemcpyDeviceToHost, c->stream));
  }
  CHIP(c, hipStreamSynchronize(c->stream));
  CHIP(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  return LISA_OK;
}

extern "C" {

lisa_status lisa_cache_train(lisa_cache* c, const float* feat, const float* target, int64_t n,
                             float* loss) {
  if (!c) return cfail(nullptr, LISA_ERR_INVALID_ARG, "cache is NULL");
  if (n <= 0 || !feat || !target) return cfail(c, LISA_ERR_INVALID_ARG, "bad arrays");
  CHIP(c, hipSetDevice(c->device));
  lisa_status rc = ensure_io(c, n);
  if (rc != LISA_OK) return rc;
  CHIP(c, hipMemcpy(c->feat, feat, 4 * kInputs * (size_t)n, hipMemcpyHostToDevice));
  CHIP(c, hipMemcpy(c->tgt, target, 4 * kOut * (size_t)n, hipMemcpyHostToDevice));
  return lisa::cache_train_device(c, c->feat, kInputs, c->tgt, n, loss);
}

float lisa_cache_last_ms(const lisa_cache* c) { return c ? c->last_ms : 0.0f; }

}  // extern "C"

// Forward pass on n device-resident feature rows (12 floats) into d_out (3 per row).
lisa_status lisa::cache_infer_device(lisa_cache* c, const float* d_feat, int64_t n, float* d_out) {
  if (!c) return cfail(nullptr, LISA_ERR_INVALID_ARG, "cache is NULL");
  if (n <= 0) return LISA_OK;
  CHIP(c, hipSetDevice(c->device));
  CHIP(c, hipEventRecord(c->ev0, c->stream));
  k_mlp_infer<<<infer_blocks(c->cus, n), 64 * kInferWaves, 0, c->stream>>>(d_feat, n, c->wt, d_out);
  CHIP(c, hipGetLastError());
  CHIP(c, hipEventRecord(c->ev1, c->stream));
  CHIP(c, hipStreamSynchronize(c->stream));
  CHIP(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  return LISA_OK;
}
