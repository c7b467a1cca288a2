// The candidate ranker's inference on one tile of 64 blocks, for the learned
// search's fused rank-and-select kernel (bc7_encode.hip); the helpers for
// bf16 conversion are shared with the training kernels (mlp.hip).  A
// workgroup of four waves computes each layer as 32x32x16 bf16 MFMA tiles
// whose accumulation runs over K in the same order as gemm_nt_kernel, so
// every logit is bit-identical to the training forward pass; the bf16 hidden
// activations stay in LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace fh {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// float -> bfloat16 bits, round to nearest even (v_cvt_pk_bf16_f32): for
// finite values and infinities the same bits as adding 0x7FFF plus the
// kept part's lowest bit and truncating.
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// b / 255.0f, correctly rounded, for a byte b: b / 255 is 0, 1 or not a
// dyadic rational, so it lies further (relative 2^-41) from every float
// rounding boundary than the double product (within 2^-52 of it), which
// therefore rounds to the same float.
__device__ __forceinline__ float byte_unit(uint32_t b) {
  return (float)((double)b * (1.0 / 255.0));
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int kRankOffB1 = 64 * 256, kRankOffB2 = kRankOffB1 + 256 + 256 * 256,
              kRankOffB3 = kRankOffB2 + 256 + 256 * 512;  // biases in the parameter vector
constexpr int kRankRows = 64;
constexpr int kXs = 72;    // LDS row strides in bf16 elements: 16-byte multiples whose
constexpr int kHs = 264;   // rows fall on different banks for the operand loads
constexpr int kLdsX = 0, kLdsH1 = kRankRows * kXs * 2, kLdsH2 = kLdsH1 + kRankRows * kHs * 2;
constexpr int kLdsBytes = kLdsH2 + kRankRows * kHs * 2;

// Features of blocks tile0 .. tile0 + 63 of r (n real): one texel row of one
// block per thread (256 threads), RGBA / 255 in bf16, as features_kernel
// computes them; rows past n are zero.
__device__ __forceinline__ void rank_features(const BlockRange &r, uint32_t n, uint32_t tile0,
                                              int tid, uint16_t *xs) {
  const int row = tid >> 2, y = tid & 3;
  const bool real = tile0 + row < n;
  uint4 px = make_uint4(0, 0, 0, 0);
  if (real) {
    const uint32_t b = r.first_block + tile0 + row;
    const uint32_t bx = b % r.blocks_wide, by = b / r.blocks_wide;
    px = *reinterpret_cast<const uint4 *>(r.img + (size_t)(by * 4 + y) * r.width + bx * 4);
  }
  const uint32_t p4[4] = {px.x, px.y, px.z, px.w};
  uint32_t packed[8];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int c = 0; c < 4; c += 2) {
      const uint16_t f0 = real ? f2bf(byte_unit((p4[i] >> (8 * c)) & 0xFFu)) : 0;
      const uint16_t f1 = real ? f2bf(byte_unit((p4[i] >> (8 * c + 8)) & 0xFFu)) : 0;
      packed[2 * i + c / 2] = f0 | ((uint32_t)f1 << 16);
    }
  *reinterpret_cast<uint4 *>(xs + row * kXs + 16 * y) =
      make_uint4(packed[0], packed[1], packed[2], packed[3]);
  *reinterpret_cast<uint4 *>(xs + row * kXs + 16 * y + 8) =
      make_uint4(packed[4], packed[5], packed[6], packed[7]);
}

// acc[rb][t] += IN[64 x K] Wt^T for wave column tiles c0 + 32 t (rows 32 rb ..).
template <int K, int SIN, int NT>
__device__ __forceinline__ void rank_mma(const uint16_t *in, const uint16_t *Wt, int c0,
                                         int lane, floatx16 (&acc)[2][NT]) {
  const int r = lane & 31, kh = (lane >> 5) * 8;
#pragma unroll 4
  for (int k = 0; k < K; k += 16) {
    const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(in + r * SIN + k + kh);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(in + (32 + r) * SIN + k + kh);
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const bf16x8 b =
          *reinterpret_cast<const bf16x8 *>(Wt + (size_t)(c0 + t * 32 + r) * K + k + kh);
      acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc[1][t], 0, 0, 0);
    }
  }
}

// Row of accumulator element v (of 16) in row block rb for this lane.
__device__ __forceinline__ int acc_row(int rb, int v, int lane) {
  return rb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
}

// H[64 x NT*32 per wave] = relu(IN[64 x K] W^T + b) for wave w's column
// tiles c0 + 32 t, rounded to bf16 into LDS.
template <int K, int SIN, int NT>
__device__ __forceinline__ void rank_hidden(const uint16_t *in, const uint16_t *Wt,
                                            const float *bias, int c0, int lane, uint16_t *hout) {
  const int r = lane & 31;
  floatx16 acc[2][NT] = {};
  rank_mma<K, SIN, NT>(in, Wt, c0, lane, acc);
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int col = c0 + t * 32 + r;
    const float bv = bias[col];
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int row = acc_row(rb, v, lane);
        float x = acc[rb][t][v] + bv;
        if (x < 0.0f) x = 0.0f;
        hout[row * kHs + col] = f2bf(x);
      }
  }
}

}  // namespace mlp
}  // namespace fh
