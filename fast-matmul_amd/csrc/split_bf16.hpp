// split_bf16.hpp -- the three-slice bf16 split of an FP32 value, shared by
// the slice-plane preparation (csrc/gemm_split.hip) and the two-level S/T
// stages that write slice planes directly (csrc/stage_kernels.hpp), so that
// both give the same bits.
//
// x = x0 + x1 + x2 with x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)
// (round to nearest even; every difference is exact in FP32).  Finite
// values whose rounding to bf16 would overflow (|x| at or above the largest
// bf16 plus half an ulp) take the truncated value as the leading slice
// instead, so their residuals stay finite (x - x0 is then exact: both lie in
// the top binade).  Inf and NaN go through the leading slice with zero
// residuals (a NaN keeps a set mantissa bit).
//
// Slice-plane layout of an operand of `rows` x K (A: row r = A[r, :];
// B: row r = B[:, r]): element (r, k), slice p at byte
//   (k / 32 * rows_pad + r) * 192 + p * 64 + (k % 32) * 2
// of the operand's planes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// round-to-nearest-even float -> bf16 (finite inputs; fm_split1 handles the rest)
__device__ __forceinline__ uint16_t fm_bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float fm_bf16_to_f32(uint16_t b) {
  return __uint_as_float((uint32_t)b << 16);
}

// the three slices of one value
__device__ __forceinline__ void fm_split1(float x, uint32_t& b0, uint32_t& b1, uint32_t& b2) {
  const uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) {
    b0 = (u >> 16) | ((u & 0x007FFFFFu) ? 0x0040u : 0u);
    b1 = b2 = 0;
  } else {
    uint16_t h = fm_bf16_bits(x);
    if ((h & 0x7F80u) == 0x7F80u) h = (uint16_t)(u >> 16);  // rounded to Inf: truncate
    const float r1 = __fsub_rn(x, fm_bf16_to_f32(h));
    const uint16_t m = fm_bf16_bits(r1);
    const float r2 = __fsub_rn(r1, fm_bf16_to_f32(m));
    b0 = h;
    b1 = m;
    b2 = fm_bf16_bits(r2);
  }
}
