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
  int32_t vec;      // every A/B column start is 16-byte aligned: vector loads allowed
  int32_t grid;     // > 0: persistent launch with this many work groups
  int32_t full;     // every product is a whole number of tiles and K slices
};

// Tile configurations of the leaf kernel: BM x BN output tile, BK-deep K
// slices, WM x WN waves (64 * WM * WN threads).
enum FmTileCfg : int {
  FM_TILE_128x128 = 0,         // BK 16, 2x2 waves
  FM_TILE_64x64 = 1,           // BK 16, 2x2 waves
  FM_TILE_128x64 = 2,          // BK 16, 2x2 waves
  FM_TILE_64x128 = 3,          // BK 16, 2x2 waves
  FM_TILE_128x128_K32 = 4,     // BK 32, 2x2 waves
  FM_TILE_128x128_W8 = 5,      // BK 16, 2x4 waves
  FM_TILE_128x128_K32_W8 = 6,  // BK 32, 2x4 waves
  FM_TILE_256x128_W8 = 7,      // BK 16, 4x2 waves
  FM_TILE_128x256_W8 = 8,      // BK 16, 2x4 waves
  FM_TILE_256x256_W16 = 9,     // BK 16, 4x4 waves
  FM_TILE_64x64_K32 = 10,      // BK 32, 2x2 waves
  FM_TILE_128x128_W16 = 11,    // BK 16, 4x4 waves
  FM_TILE_128x64_W8 = 12,      // BK 16, 2x4 waves
  FM_TILE_64x128_W8 = 13,      // BK 16, 4x2 waves
  FM_TILE_128x128_O2 = 14,     // BK 16, 2x2 waves, registers for 2 waves per SIMD
  FM_TILE_256x128_W8_U1 = 15,  // BK 16, 4x2 waves, K steps not unrolled
  FM_TILE_128x128_O2_U1 = 16,  // as 14, K steps not unrolled
  FM_TILE_128x256_W8_U1 = 17,  // BK 16, 2x4 waves, K steps not unrolled
  // software-pipelined main loop: LDS fragments double-buffered, the last
  // K step of a slice issued after the barrier
  FM_TILE_PP_256x128_W8 = 18,  // BK 16, 4x2 waves
  FM_TILE_PP_128x256_W8 = 19,  // BK 16, 2x4 waves
  FM_TILE_PP_128x128_W8 = 20,  // BK 16, 2x4 waves
  FM_TILE_PP_128x128_O2 = 21,  // BK 16, 2x2 waves, registers for 2 waves per SIMD
  // as PP, with LDS reads interleaved with MFMAs by scheduling-group barriers
  FM_TILE_PI_128x128_O2 = 22,  // BK 16, 2x2 waves, registers for 2 waves per SIMD
  FM_TILE_PI_128x256_W8 = 23,  // BK 16, 2x4 waves
  FM_TILE_NUM = 24
};

// K slice depth of each configuration
inline int fm_tile_bk(int cfg) {
  return (cfg == FM_TILE_128x128_K32 || cfg == FM_TILE_128x128_K32_W8 || cfg == FM_TILE_64x64_K32)
             ? 32
             : 16;
}

// output tile of each configuration
inline void fm_tile_dims(int cfg, int* bm, int* bn) {
  static const int d[FM_TILE_NUM][2] = {{128, 128}, {64, 64}, {128, 64}, {64, 128},
                                        {128, 128}, {128, 128}, {128, 128}, {256, 128},
                                        {128, 256}, {256, 256}, {64, 64},  {128, 128},
                                        {128, 64},  {64, 128},  {128, 128}, {256, 128},
                                        {128, 128}, {128, 256}, {256, 128}, {128, 256},
                                        {128, 128}, {128, 128}, {128, 128}, {128, 256}};
  *bm = d[cfg][0];
  *bn = d[cfg][1];
}

hipError_t fm_gemm_launch_f64(int cfg, const FmGemmArgs& a, hipStream_t s);
hipError_t fm_gemm_launch_f32(int cfg, const FmGemmArgs& a, hipStream_t s);
// FP32 data, split-bf16 slice products (csrc/gemm_split.hip)
hipError_t fm_gemm_launch_bf16x3(int cfg, const FmGemmArgs& a, hipStream_t s);
