// gemm_ozaki.hpp -- descriptors of the int8-digit (Ozaki) FP64 leaf products.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "alg_desc.hpp"

// One product C = alpha * A B + beta * C (FP64, column-major) computed from
// int8 digit planes that the preparation kernels write into the workspace.
struct FmOzakiDesc {
  FmOp a, b, c;                 // FP64 operands
  int32_t M, N, K, Kp;          // Kp: K rounded up to 64 (digit row length)
  int32_t rows_pad, cols_pad;   // M rounded up to 64, N to the tile width (64 or 128)
  int32_t tiles_m, tiles_n, tile_begin, pad;
  int64_t a_slc, b_slc;         // byte offsets of the digit planes in the workspace
  int64_t a_exp, b_exp;         // byte offsets of the int32 row / column exponents
  double alpha, beta;
};

struct FmOzakiArgs {
  FmBases bases;
  const FmOzakiDesc* descs;
  int32_t ndesc;
  int32_t total_tiles;
  int32_t variant;  // 0: 16 x 16 x 64 MFMA kernel, 64-deep slices; 1: 32 x 32 x 32, two 32-deep
                    // slices per stage, 64 x 64 tiles; 2: as 1 on 64 x 128 tiles
  int32_t pad;
};

// Largest K (rounded to 64) for which every per-scale int32 sum is exact.
inline int64_t fm_ozaki_max_kp(int slices) { return (int64_t)2147483647 / (slices * 127 * 127); }

hipError_t fm_ozaki_launch(int slices, const FmOzakiArgs& a, int max_rows, int max_cols,
                           int phases, hipStream_t s);
