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
  int32_t vec;      // the base pointers are 16-byte aligned: vector loads allowed for the
                    // products whose own offsets and leading dimensions keep columns aligned
  int32_t grid;     // > 0: persistent launch with this many work groups
  int32_t full;     // every product is a whole number of tiles and K slices
  int32_t even;     // every product has M and K multiples of the vector width and
                    // 16-byte aligned columns (edge tiles and a partial last K slice)
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
  // split-bf16 leaves on bf16 slice planes prepared once per operand
  // (FP32 data only): BK 32
  FM_TILE_SP_128x128 = 24,     // 2x2 waves, one LDS buffer (two work groups per CU)
  FM_TILE_SP_256x128 = 25,     // 4x2 waves, LDS double-buffered
  FM_TILE_SP_128x256 = 26,     // 2x4 waves, LDS double-buffered
  FM_TILE_SP_128x128_DB = 27,  // 2x2 waves, LDS double-buffered
  FM_TILE_SP_256x128_IL = 28,  // as 25, fragment reads interleaved with the MFMAs
  FM_TILE_SP_128x256_IL = 29,  // as 26, fragment reads interleaved with the MFMAs
  // plain FP64 / FP32 again: 96-column strips (e.g. 2400 = 18 x 128 + 96)
  FM_TILE_PI_128x96_O2 = 30,   // BK 16, 2x2 waves of 64 x 48, registers for 2 waves per SIMD
  // split-bf16 on slice planes, B fragments loaded straight into registers
  FM_TILE_SP_128x256_BD = 31,  // 2x4 waves, A through double-buffered LDS
  // as 22 and 23, with the next slice's LDS stores and global loads spread
  // over the K steps (edge launches too, on products whose M and K are
  // multiples of the vector width and whose columns are 16-byte aligned)
  FM_TILE_PS_128x128_O2 = 32,
  FM_TILE_PS_128x256_W8 = 33,
  FM_TILE_PS_128x96_O2 = 34,   // as 30, spread
  FM_TILE_NUM = 35
};

// configurations that run on prepared slice planes
inline bool fm_tile_planes(int cfg) {
  return (cfg >= FM_TILE_SP_128x128 && cfg <= FM_TILE_SP_128x256_IL) || cfg == FM_TILE_SP_128x256_BD;
}

// K slice depth of each configuration
inline int fm_tile_bk(int cfg) {
  return (cfg == FM_TILE_128x128_K32 || cfg == FM_TILE_128x128_K32_W8 || cfg == FM_TILE_64x64_K32 ||
          fm_tile_planes(cfg))
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
                                        {128, 128}, {128, 128}, {128, 128}, {128, 256},
                                        {128, 128}, {256, 128}, {128, 256}, {128, 128},
                                        {256, 128}, {128, 256}, {128, 96},  {128, 256},
                                        {128, 128}, {128, 256}, {128, 96}};
  *bm = d[cfg][0];
  *bn = d[cfg][1];
}

hipError_t fm_gemm_launch_f64(int cfg, const FmGemmArgs& a, hipStream_t s);
hipError_t fm_gemm_launch_f32(int cfg, const FmGemmArgs& a, hipStream_t s);
// FP32 data, split-bf16 slice products (csrc/gemm_split.hip)
hipError_t fm_gemm_launch_bf16x3(int cfg, const FmGemmArgs& a, hipStream_t s);

// ---- split-bf16 products on prepared slice planes ----------------------------
// An FP32 operand of `rows` x K (A: rows = M, row r = A[r, :]; B: rows = N,
// row r = B[:, r]) is cut once into three bf16 slice planes in the
// workspace.  Element (r, k), slice p (0 = leading) is the bf16 at byte
//   dst + (k / 32 * rows_pad + r) * 192 + p * 64 + (k % 32) * 2,
// so each 32-deep K slice of a row is one 192-byte chunk holding the three
// planes, and the chunks of all rows of one K slice are consecutive; rows
// [rows, rows_pad) and k in [K, Kp) hold zeros.
struct FmSplitPrep {
  FmOp src;             // FP32 source, column-major
  int32_t rows, K;
  int32_t rows_pad, Kp;  // rows rounded up to the tile, K rounded up to 32
  int32_t kcontig;       // 1: K contiguous in the source (B), 0: rows contiguous (A)
  int32_t wg_begin;      // first work group of this operand in the preparation launch
  int64_t dst;           // byte offset in the workspace
};

// K slices of 32 per preparation work group (64 rows each)
constexpr int FM_PREP_KCH = 4;
inline int fm_prep_wgs(const FmSplitPrep& p) {
  return (p.rows_pad / 64) * ((p.Kp / 32 + FM_PREP_KCH - 1) / FM_PREP_KCH);
}

// One product C = alpha * A B + beta * C from the planes of its operands.
struct FmSplitDesc {
  FmOp c;
  int32_t M, N, K, Kp;
  int32_t tiles_m, tiles_n, tile_begin, pad;
  int64_t a_pl, b_pl;   // workspace byte offsets of the A and B planes
  double alpha, beta;
};

struct FmSplitArgs {
  FmBases bases;
  const FmSplitDesc* descs;
  const FmSplitPrep* preps;
  int32_t ndesc, total_tiles;
  int32_t uniform, grid;        // as in FmGemmArgs
  int32_t nprep;
  int32_t prep_wgs;             // work groups of the preparation launch (all operands)
  int32_t even_k;               // 1: every product has an even number of K slices
};

// the preparation launch, then the grouped product launch (phases: FmLeafPhase)
hipError_t fm_split_launch(int cfg, const FmSplitArgs& a, int phases, hipStream_t s);
