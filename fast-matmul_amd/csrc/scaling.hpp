// scaling.hpp -- diagonal scaling kernels (csrc/scaling.hip) used by
// fm_fastmatmul_scaled.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// A (m x k) and B (k x n) are scaled in place (they are the caller's
// operands copied into the workspace); r (m) and s (n) receive the row and
// column factors of the product; f, f2 hold per-pass factors and u1, u2 the
// per-pass maxima (max(m, k, n) entries each).
struct FmScaleArgs {
  void* A;
  void* B;
  int64_t lda, ldb;
  int m, k, n;
  int type;   // 0 OUTSIDE_INSIDE, 1 INSIDE_OUTSIDE, 2 INSIDE, 3 OUTSIDE
  int steps;
  void* r;
  void* s;
  void* f;
  void* f2;
  void* u1;
  void* u2;
};

hipError_t fm_scaling_run(int dtype_f64, const FmScaleArgs& a, hipStream_t s);
hipError_t fm_scaling_copy(int dtype_f64, const void* X, int64_t ldx, void* Y, int64_t ldy,
                           int rows, int cols, hipStream_t s);
hipError_t fm_scaling_post(int dtype_f64, const void* P, int64_t ldp, void* C, int64_t ldc,
                           int rows, int cols, const void* r, const void* sv, double alpha,
                           double beta, hipStream_t s);
