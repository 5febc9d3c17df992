// Surrogate MLP building blocks on CDNA4 matrix cores: BF16 operands, FP32
// accumulation with v_mfma_f32_16x16x16_bf16 (one wave computes a 16x16 tile).
//
// Lane layout of the 16x16x16 BF16 MFMA (wave64): lane l holds
//   A: row l%16, k = 4*(l/16) .. 4*(l/16)+3
//   B: column l%16, k = 4*(l/16) .. 4*(l/16)+3
//   C: rows 4*(l/16) .. 4*(l/16)+3, column l%16
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mind {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (NaN stays a quiet NaN).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Exact GELU, x * Phi(x), and its derivative.
__host__ __device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__host__ __device__ inline float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// Layer geometry of the surrogate: n_in -> hidden (x layers) -> n_out, padded to 16.
struct MlpShape {
  int n_in, hidden, layers, n_out;  // layers = number of hidden layers
  __host__ __device__ int nlin() const { return layers + 1; }
  __host__ __device__ int k(int l) const { return l == 0 ? n_in : hidden; }
  __host__ __device__ int n(int l) const { return l == layers ? n_out : hidden; }
  __host__ __device__ static int pad(int x) { return (x + 15) / 16 * 16; }
};

}  // namespace mind
