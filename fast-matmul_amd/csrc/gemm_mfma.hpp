// gemm_mfma.hpp -- grouped GEMM descriptors shared by the host plan builder
// and the MFMA leaf kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "alg_desc.hpp"

// One product C = alpha * A * B + beta * C, column-major, 'N','N'
// (the leaf call of linalg/linalg.hpp:261-269 / linalg/blas_wrap.hpp:109-121).
struct FmGemmDesc {
  FmOp a, b, c;
  int32_t M, N, K;
  int32_t tiles_m, tiles_n;
  int32_t tile_begin;  // first tile of this product in the launch
  double alpha, beta;
};

struct FmGemmArgs {
  FmBases bases;
  const FmGemmDesc* descs;
  int32_t ndesc;
  int32_t total_tiles;
  int32_t uniform;  // all products have one shape: tile -> product by division
  int32_t pad;
};

enum FmTileCfg : int { FM_TILE_128x128 = 0, FM_TILE_64x64 = 1, FM_TILE_128x64 = 2, FM_TILE_64x128 = 3 };

// tile sizes of each configuration
inline void fm_tile_dims(int cfg, int* bm, int* bn) {
  switch (cfg) {
    case FM_TILE_128x128: *bm = 128; *bn = 128; break;
    case FM_TILE_64x64: *bm = 64; *bn = 64; break;
    case FM_TILE_128x64: *bm = 128; *bn = 64; break;
    default: *bm = 64; *bn = 128; break;
  }
}

hipError_t fm_gemm_launch_f64(int cfg, const FmGemmArgs& a, hipStream_t s);
hipError_t fm_gemm_launch_f32(int cfg, const FmGemmArgs& a, hipStream_t s);
