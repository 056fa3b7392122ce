// Learned warm start of the FIM: a 22-64-64-4 ReLU MLP per element copy,
// run and trained with BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32
// accumulate). No reference counterpart.
//
// Input of one tet copy (22 values, zero-padded to 32): the four corner phi
// values and the six edge vectors AB, AC, AD, BC, BD, CD in the copy's local
// frame, all divided by the copy's scale s = |AB|. Output (4 values, padded
// to 16): y_i = (D_i - |phi_i|) / s, D_i the corner's distance after
// redistancing. The predicted distance is max(0, |phi_i| + s * y_i); a
// vertex's warm value is the mean over its element copies. A triangle copy
// uses the same layout with its three corners and edges AB, AC, BC, the
// fourth corner's phi and edges zero and its output unused (a context holds
// one kind of mesh, so its network sees one kind of input).
//
// MFMA operand layout (64 lanes, g = lane / 16, r = lane % 16):
//   A (16x16, rows m, cols k):   lane holds A[r][4g + j], j = 0..3
//   B (16x16, rows k, cols n):   lane holds B[4g + j][r]
//   D (16x16):                   lane holds D[4g + j][r]
// A wave processes 16 samples at a time; activations pass between layers
// through the wave's own LDS tile.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsamd {
namespace mlp {

constexpr int kIn = 22, kInPad = 32, kHid = 64, kOut = 4, kOutPad = 16;
// Flat FP32 parameter vector: W1 (kIn x kHid, row-major by input), b1,
// W2 (kHid x kHid), b2, W3 (kHid x kOut), b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kOut;
constexpr int kNumParams = kOffB3 + kOut;  // 5892

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 -> BF16 (finite inputs).
__host__ __device__ inline uint16_t to_bf16(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

// Weights as MFMA B operands, BF16, padded: W1 [kInPad][kHid],
// W2 [kHid][kHid], W3 [kHid][kOutPad]; biases FP32 (b3 padded to kOutPad).
struct Weights {
  const uint16_t* w1;
  const uint16_t* w2;
  const uint16_t* w3;
  const float* b1;
  const float* b2;
  const float* b3;
};

// Features of one tet copy from its corner phi values and local coordinates
// A=(0,0,0), B=(L0,0,0), C=(L1,L2,0), D=(L3,L4,L5). Returns the scale s.
__host__ __device__ inline float features(const float* ph, const float* L, float* x) {
  const float s = L[0];
  const float inv = 1.0f / s;
  for (int i = 0; i < 4; ++i) x[i] = ph[i] * inv;
  const float AB[3] = {L[0], 0.f, 0.f}, AC[3] = {L[1], L[2], 0.f}, AD[3] = {L[3], L[4], L[5]};
  for (int k = 0; k < 3; ++k) {
    x[4 + k] = AB[k] * inv;
    x[7 + k] = AC[k] * inv;
    x[10 + k] = AD[k] * inv;
    x[13 + k] = (AC[k] - AB[k]) * inv;
    x[16 + k] = (AD[k] - AB[k]) * inv;
    x[19 + k] = (AD[k] - AC[k]) * inv;
  }
  return s;
}

// Triangle copy, local frame A=(0,0), B=(c0,0), C=(c1,c2).
__host__ __device__ inline float features_tri(const float* ph, const float* c, float* x) {
  const float s = c[0];
  const float inv = 1.0f / s;
  for (int i = 0; i < 3; ++i) x[i] = ph[i] * inv;
  x[3] = 0.f;
  const float AB[3] = {c[0], 0.f, 0.f}, AC[3] = {c[1], c[2], 0.f};
  for (int k = 0; k < 3; ++k) {
    x[4 + k] = AB[k] * inv;
    x[7 + k] = AC[k] * inv;
    x[10 + k] = 0.f;
    x[13 + k] = (AC[k] - AB[k]) * inv;
    x[16 + k] = 0.f;
    x[19 + k] = 0.f;
  }
  return s;
}

// Features of an element copy with NC corners (ph: NC values, L: its local
// coordinates).
template <int NC>
__host__ __device__ inline float copy_features(const float* ph, const float* L, float* x) {
  if constexpr (NC == 4) return features(ph, L, x);
  else return features_tri(ph, L, x);
}

__device__ inline bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 v;
  v[0] = (short)to_bf16(a);
  v[1] = (short)to_bf16(b);
  v[2] = (short)to_bf16(c);
  v[3] = (short)to_bf16(d);
  return v;
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// B operand of a row-major BF16 matrix W[k][n] (ld columns), tile (ks, nt).
__device__ inline bf16x4 load_b(const uint16_t* W, int ld, int ks, int nt, int lane) {
  const int g = lane >> 4, r = lane & 15;
  const uint16_t* p = W + (ks * 16 + 4 * g) * ld + nt * 16 + r;
  bf16x4 v;
  v[0] = (short)p[0];
  v[1] = (short)p[ld];
  v[2] = (short)p[2 * ld];
  v[3] = (short)p[3 * ld];
  return v;
}

// A operand from a wave's FP32 LDS tile H[16][ld] (rows = samples), k-step ks.
__device__ inline bf16x4 load_a(const float* H, int ld, int ks, int lane) {
  const int g = lane >> 4, r = lane & 15;
  const float* p = H + r * ld + ks * 16 + 4 * g;
  return pack4(p[0], p[1], p[2], p[3]);
}

// Writes an MFMA result tile (nt) into H[16][ld] with bias and optional ReLU.
__device__ inline void store_d(float* H, int ld, int nt, f32x4 d, const float* bias, bool relu, int lane) {
  const int g = lane >> 4, r = lane & 15;
  const float bb = bias ? bias[nt * 16 + r] : 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float v = d[j] + bb;
    if (relu) v = v > 0.f ? v : 0.f;
    H[(4 * g + j) * ld + nt * 16 + r] = v;
  }
}

// Orders a wave's LDS writes before its later LDS reads of other lanes' data.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward pass of 16 samples whose padded inputs sit in X[16][kInPad].
// Leaves H1 and H2 (post-ReLU) and the output O[16][kOutPad] (with b3) in
// the wave's LDS tiles.
__device__ inline void forward16(const Weights& w, const float* X, float* H1, float* H2, float* O, int lane) {
#pragma unroll
  for (int nt = 0; nt < kHid / 16; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kInPad / 16; ++ks) acc = mfma(load_a(X, kInPad, ks, lane), load_b(w.w1, kHid, ks, nt, lane), acc);
    store_d(H1, kHid, nt, acc, w.b1, true, lane);
  }
  wave_sync();
#pragma unroll
  for (int nt = 0; nt < kHid / 16; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kHid / 16; ++ks) acc = mfma(load_a(H1, kHid, ks, lane), load_b(w.w2, kHid, ks, nt, lane), acc);
    store_d(H2, kHid, nt, acc, w.b2, true, lane);
  }
  wave_sync();
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < kHid / 16; ++ks) acc = mfma(load_a(H2, kHid, ks, lane), load_b(w.w3, kOutPad, ks, 0, lane), acc);
  store_d(O, kOutPad, 0, acc, w.b3, false, lane);
  wave_sync();
}

// LDS floats one wave needs: X, H1, H2, O tiles of 16 rows.
constexpr int kWaveTile = 16 * (kInPad + kHid + kHid + kOutPad);

}  // namespace mlp
}  // namespace lsamd
