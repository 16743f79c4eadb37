// plan.hpp -- execution plans: the reference's recursion unrolled into
// batched launches over all nodes of each level.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fastmm.h"
#include "alg_desc.hpp"
#include "gemm_mfma.hpp"
#include "gemm_ozaki.hpp"

namespace fmplan {

fm_status set_err(fm_status s, const std::string& msg);
fm_status hip_err(hipError_t e, const char* what);
const std::string& last_error();
void clear_error();

enum LaunchKind { LK_S, LK_T, LK_W, LK_GEMM, LK_OZAKI, LK_S2, LK_T2, LK_W2, LK_SPLIT, LK_NONE };

// One kernel launch.  Descriptors stay on the host until Plan::upload packs
// them into the plan's device buffer.
struct Launch {
  LaunchKind kind = LK_GEMM;
  int stage = 0;                   // fm_stage_kind, for profiling
  const FmAlgDesc* alg = nullptr;  // stage launches: whose combine functions
  FmStageArgs st{};
  FmGemmArgs gm{};
  FmOzakiArgs oz{};
  FmSplitArgs sp{};
  int cfg = 0;
  int vec_ok = 1;
  int planes = 0;  // LK_S2 / LK_T2 writing split-bf16 slice planes (byte offsets in st)
  int oz_slices = 0, oz_rows = 0, oz_cols = 0;
  std::vector<FmOp> ops;
  std::vector<FmStageNode> snodes;  // grouped stage launch: per-node descriptors
  std::vector<FmGemmDesc> gdescs;
  std::vector<FmOzakiDesc> odescs;
  std::vector<FmSplitDesc> sdescs;
  std::vector<FmSplitPrep> spreps;
  size_t dev_off = 0, dev_off2 = 0;
};

struct Plan {
  fm_dtype dtype = FM_F64;
  fm_mode mode = FM_MODE_PLAIN;
  int slices = 0;
  int tile = -1;       // forced leaf tile configuration (-1: cost model)
  int split_rem = 1;   // whole-tile carve-out and remainder launch
  int dfs_levels = 0;  // depth-first top levels the plan runs
  std::vector<Launch> launches;
  std::vector<void*> ext;          // role table (roles >= 4) of batched plans
  void* dbuf = nullptr;            // descriptors + role table on the device (owned)
  void* desc_base = nullptr;       // where the launches read them (dbuf, or a caller's copy)
  size_t ext_off = 0;
  size_t ws_bytes = 0;
  ~Plan();
  void pack(std::vector<char>& buf);
  void bind(void* dev);
  fm_status upload();  // pack, copy into a new dbuf, bind
  // Launch every kernel on `s`; A/B/C/ws fill roles 0-3.
  fm_status run(const void* A, const void* B, void* C, void* ws, hipStream_t s) const;
};

// One multiply C = alpha op(A) op(B) + beta C, as the reference's recursion.
struct Problem {
  const FmAlgDesc* alg = nullptr;  // nullptr: classical product (numsteps ignored)
  int steps = 0;
  int64_t m = 0, k = 0, n = 0;
  int64_t lda = 0, ldb = 0, ldc = 0;
  int role_a = FM_ROLE_A, role_b = FM_ROLE_B, role_c = FM_ROLE_C;
  double x = 1e-8, alpha = 1.0, beta = 0.0;
  int64_t off_a = 0, off_b = 0, off_c = 0;  // element offsets within the roles
  bool fuse = true;  // two-level S/T stages where the shape allows them
  // split_bf16 leaves on slice planes: the last two-level S/T stage writes
  // the leaf operands' planes itself (FP32 data; single problems only)
  bool fuse_planes = false;
  int dfs = 0;       // top levels run depth-first (one child sub-plan at a time)
  fm_mode mode = FM_MODE_PLAIN;  // leaf mode and digit count, for depth-first children
  int slices = 0;
  int tile = -1, split_rem = 1;  // leaf launch options, for depth-first children
};

struct Dims {
  int64_t m, k, n;
};

// The pieces of one problem's plan, before launches of several problems
// are merged.
struct SubPlan {
  std::vector<Launch> pre;                          // S/T stages, level order
  std::vector<FmGemmDesc> leaves;                   // base-case products
  std::vector<Launch> w;                            // W stage of level l at w[l]
  std::vector<std::vector<FmGemmDesc>> peels;       // peeling products of level l
  int depth = 0;
  int64_t ws_elems = 0;
  Dims leaf{0, 0, 0};
};

// Build the sub-plan of `p`, placing its workspace at element offset ws_base.
fm_status build_subplan(fm_dtype dt, const Problem& p, int64_t ws_base, SubPlan& out);
fm_status build_dfs(fm_dtype dt, const Problem& p, const std::vector<Dims>& dims, int64_t ws_base,
                    SubPlan& out);

// Merge sub-plans into one launch sequence:
//   all S/T stages; one grouped leaf launch; per level (deepest first) the W
//   stages, then one grouped launch of that level's peeling products.
// persistent_grid > 0 runs the grouped leaf launch with that many work
// groups, each looping over tiles (no carving or remainder split then);
// 0 launches one work group per tile.
fm_status assemble(Plan& P, std::vector<SubPlan>& subs, int64_t ws_elems, int persistent_grid = 0);

int pick_tile(fm_dtype dt, fm_mode mode, int tile, const std::vector<FmGemmDesc>& prods);
// move the products of a part-empty last round of tiles to `rest`
void split_remainder(const Plan& P, int cfg, std::vector<FmGemmDesc>& prods,
                     std::vector<FmGemmDesc>& rest, int* cfg_rest, int* cfg_head);
// process defaults of the per-call options (fm_set_* change them)
extern fm_options g_defaults;
fm_status check_options(const fm_options& o);
fm_status check_args(fm_dtype dt, int m, int k, int n, const void* A, int64_t lda, const void* B,
                     int64_t ldb, const void* C, int64_t ldc);
fm_status check_mode(fm_dtype dt, fm_mode mode);
// One multiply's plan (alg == nullptr: the classical product) under options
// `o`, not uploaded.
fm_status build_single(Plan& P, const FmAlgDesc* alg, fm_dtype dt, fm_mode mode, int64_t m,
                       int64_t k, int64_t n, int64_t lda, int64_t ldb, int64_t ldc, int steps,
                       double x, double alpha, double beta, const fm_options& o);

// profiling (fm_profile_enable / fm_profile_read)
extern bool g_prof_on;
fm_status profile_read(fm_profile* out);

int elem_size(fm_dtype dt);
double round_scalar(double v, fm_dtype dt);
int64_t align_up(int64_t v, int64_t a);

}  // namespace fmplan
