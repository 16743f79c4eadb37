// costmodel.hpp -- the autotuner's cost model of this engine on MI355X:
// estimated device time of one multiply for a given algorithm and depth,
// from rates measured on the GPU and kept in a small text table
// (fastmm/tuning/mi355x.txt, written by python -m fastmm.calibrate).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "plan.hpp"

namespace fmplan {

struct CostTable {
  // grouped leaf GEMM rate (TFLOP/s of 2mnk) against the leaf's size
  // (cube root of m n k), per dtype (0 FP64, 1 FP32) and leaf mode
  // (plain, split_bf16, ozaki_int8); sorted by size, interpolated in log size
  std::vector<std::pair<double, double>> leaf[2][3];
  double stage_gbps[2] = {5000.0, 5000.0};  // S/T/W stage bandwidth per dtype
  double launch_us = 4.0;                   // one kernel launch
  double level_us[2] = {20.0, 20.0};        // one recursion level of a batched problem,
                                            // beyond its bytes and launches, per dtype
  double tile_k_us[2] = {0.1, 0.1};         // one peeling tile's time per unit of K: a
                                            // launch lasts at least its longest tile
  std::string source = "built-in";
};

// The table in use: loaded at first use from the file named by FM_TUNING_FILE
// or tuning/mi355x.txt beside the library, else built-in values.
const CostTable& cost_table();
// Replace the table by the file at `path` (the same format).
fm_status cost_table_load(const char* path);

// Estimated device time (us) of one multiply with `alg` at `steps` levels
// (alg == nullptr or steps == 0: the classical product).
double estimate_us(fm_dtype dt, fm_mode mode, const FmAlgDesc* alg, int64_t m, int64_t k, int64_t n,
                   int steps);

}  // namespace fmplan
