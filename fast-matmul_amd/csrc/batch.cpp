// batch.cpp -- many independent multiplies in one launch sequence, with a
// per-problem choice of algorithm and depth and optional hipGraph capture.
//
// Each problem is planned like a single call (build_subplan); assemble()
// then merges them: every problem's S/T stages, ONE grouped launch holding
// the base-case products of all problems (found per tile by binary search;
// edge strips and remainder tiles in one or two more), and, deepest level first, the W stages followed by one grouped launch of
// all peeling products of that level.  Operands of problem p are roles
// 4 + 3p (A), 5 + 3p (B), 6 + 3p (C) resolved through the plan's device
// table, so one launch can address every problem's matrices.
//
// The autotuner prices every candidate (algorithm, depth) with the cost
// model of costmodel.hpp -- leaf flops (tile-padded) over the leaf rate
// measured at that leaf size, stage bytes over the measured stage
// bandwidth, a measured per-level and per-launch cost -- and keeps the
// cheapest.  The rates come from a table that python -m fastmm.calibrate
// writes on the GPU.  The reference's counterpart is its HYBRID policy
// (util/par_util.hpp:17-95), which fixes the schedule of one algorithm;
// here the algorithm and depth are chosen per shape as well.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "costmodel.hpp"
#include "plan.hpp"

using namespace fmplan;

extern std::mutex& fm_api_mutex();
extern const FmAlgDesc* fm_find_alg(const char* name);

namespace {

// Strassen, <4,2,3;20> in all six orientations (the reference ships each
// permutation as its own table; for an uneven problem the one whose block
// split matches the shape's aspect needs the fewest flops and no peeling)
// and <3,3,3;23>
const char* const kDefaultCandidates[] = {"strassen",       "grey423_20_144", "grey432_20_144",
                                          "grey243_20_144", "grey234_20_144", "grey324_20_144",
                                          "grey342_20_144", "smirnov333_23_139"};
constexpr int kNumDefaultCandidates = (int)(sizeof(kDefaultCandidates) / sizeof(kDefaultCandidates[0]));

// Estimated time of a recursive candidate is multiplied by this before the
// comparison.  Measured on 256 uneven FP32 problems (512-4096): with no
// margin the calibrated model gave 19 problems to <4,2,3;20> orientations
// (each predicted at most 6% faster) and the batch fell from 49 to 57 ms;
// with 1.02, 5 problems, 119 -> 109-113 TFLOP/s.  Their odd block shapes add
// edge tiles and an extra leaf launch, and their stage launches (one per
// orientation) run nearly empty, none of which a per-problem model sees.
constexpr double kRecursionMargin = 1.06;

}  // namespace

struct fm_batch_s {
  Plan plan;
  std::vector<std::string> alg_names;
  std::vector<int> steps;
  void* ws = nullptr;
  bool own_ws = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~fm_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (own_ws && ws) (void)hipFree(ws);
  }
};

extern "C" {

fm_status fm_autotune_estimate(fm_dtype dtype, fm_mode mode, const char* alg, int m, int k, int n,
                               int num_steps, double* us) {
  if (!us || m <= 0 || k <= 0 || n <= 0 || num_steps < 0)
    return set_err(FM_ERR_INVALID_ARG, "bad arguments");
  const FmAlgDesc* a = nullptr;
  if (alg) {
    a = fm_find_alg(alg);
    if (!a) return set_err(FM_ERR_UNKNOWN_ALG, std::string("unknown algorithm: ") + alg);
  }
  *us = estimate_us(dtype, mode, a, m, k, n, num_steps);
  return FM_OK;
}

fm_status fm_autotune_load(const char* path) {
  clear_error();
  return cost_table_load(path);
}

const char* fm_autotune_table(void) { return cost_table().source.c_str(); }

fm_status fm_batch_create(fm_dtype dtype, fm_mode mode, const fm_problem* problems, int count,
                          const char* const* candidates, int num_candidates, int max_steps,
                          int use_graph, void* workspace, size_t workspace_bytes,
                          fm_batch* out) {
  clear_error();
  if (!out || !problems || count <= 0) return set_err(FM_ERR_INVALID_ARG, "bad batch arguments");
  *out = nullptr;
  if (dtype != FM_F64 && dtype != FM_F32) return set_err(FM_ERR_INVALID_ARG, "unknown dtype");
  if (!(mode == FM_MODE_PLAIN || (mode == FM_MODE_SPLIT_BF16 && dtype == FM_F32) ||
        (mode == FM_MODE_OZAKI_INT8 && dtype == FM_F64)))
    return set_err(FM_ERR_UNSUPPORTED, "leaf mode not available for this dtype");
  if (!candidates || num_candidates <= 0) {
    candidates = kDefaultCandidates;
    num_candidates = kNumDefaultCandidates;
  }
  if (max_steps < 0) max_steps = 2;
  std::vector<const FmAlgDesc*> cands;
  for (int i = 0; i < num_candidates; ++i) {
    const FmAlgDesc* a = fm_find_alg(candidates[i]);
    if (!a) return set_err(FM_ERR_UNKNOWN_ALG, std::string("unknown algorithm: ") + candidates[i]);
    cands.push_back(a);
  }
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  std::unique_ptr<fm_batch_s> b(new fm_batch_s());
  Plan& P = b->plan;
  P.dtype = dtype;
  P.mode = mode;
  const fm_options opt = g_defaults;
  P.slices = mode == FM_MODE_OZAKI_INT8 ? opt.ozaki_slices : 0;
  P.tile = opt.gemm_tile;
  P.split_rem = opt.split_remainder != 0;
  std::vector<SubPlan> subs(count);
  int64_t ws = 0;
  for (int i = 0; i < count; ++i) {
    const fm_problem& q = problems[i];
    if (q.m < 0 || q.k < 0 || q.n < 0 || q.lda < std::max(1, q.m) || q.ldb < std::max(1, q.k) ||
        q.ldc < std::max(1, q.m) || !q.A || !q.B || !q.C)
      return set_err(FM_ERR_INVALID_ARG, "problem " + std::to_string(i) + ": bad dimensions or pointers");
    if (q.m == 0 || q.n == 0 || q.k == 0)
      return set_err(FM_ERR_SHAPE, "problem " + std::to_string(i) + ": empty dimension");
    const FmAlgDesc* alg = nullptr;
    int steps = q.num_steps;
    if (q.alg) {
      alg = fm_find_alg(q.alg);
      if (!alg) return set_err(FM_ERR_UNKNOWN_ALG, std::string("unknown algorithm: ") + q.alg);
    }
    if (!alg || steps < 0) {  // autotune over the candidates (or depths of the given algorithm)
      double best = INFINITY;
      const FmAlgDesc* best_alg = nullptr;
      int best_steps = 0;
      std::vector<const FmAlgDesc*> pool = alg ? std::vector<const FmAlgDesc*>{alg} : cands;
      const int smin = steps >= 0 ? steps : 0, smax = steps >= 0 ? steps : max_steps;
      for (const FmAlgDesc* a : pool)
        for (int s = smin; s <= smax; ++s) {
          double t = estimate_us(dtype, mode, s == 0 ? nullptr : a, q.m, q.k, q.n, s);
          // a recursion must beat the classical product by a clear margin:
          // in a batch its odd-shaped leaves share the grouped launch with
          // every other problem's, which the per-problem model does not see
          if (s > 0) t *= kRecursionMargin;
          if (t < best) {
            best = t;
            best_alg = s == 0 ? nullptr : a;
            best_steps = s;
          }
        }
      alg = best_alg;
      steps = best_steps;
    }
    Problem p;
    p.alg = steps == 0 ? nullptr : alg;  // depth 0: classical product (beta honoured)
    p.steps = steps;
    p.m = q.m;
    p.k = q.k;
    p.n = q.n;
    p.lda = q.lda;
    p.ldb = q.ldb;
    p.ldc = q.ldc;
    p.role_a = 4 + 3 * i;
    p.role_b = 5 + 3 * i;
    p.role_c = 6 + 3 * i;
    p.alpha = q.alpha;
    p.beta = q.beta;
    p.fuse = opt.stage_fusion != 0;
    p.mode = mode;
    p.slices = P.slices;
    p.tile = P.tile;
    p.split_rem = P.split_rem;
    P.ext.push_back(const_cast<void*>(q.A));
    P.ext.push_back(const_cast<void*>(q.B));
    P.ext.push_back(q.C);
    fm_status st = build_subplan(dtype, p, ws, subs[i]);
    if (st != FM_OK) return set_err(st, "problem " + std::to_string(i) + ": " + last_error());
    ws += subs[i].ws_elems;
    b->alg_names.push_back(p.alg ? p.alg->name : "classical");
    b->steps.push_back(p.alg ? steps : 0);
  }
  // one grouped launch for every problem's base-case products, one work
  // group per tile: the tiles' K runs 512-4096 here, so the hardware
  // dispatcher, which starts a tile wherever a slot frees, balances the CUs
  // better than a persistent grid taking every grid-th tile (measured on 256
  // problems: 107.6 against 103.4 TFLOP/s with 2048 persistent work groups).
  // Whole-tile carving and the remainder split apply as for one multiply.
  fm_status st = assemble(P, subs, ws, 0);
  if (st != FM_OK) return st;
  st = P.upload();
  if (st != FM_OK) return st;
  if (P.ws_bytes > 0) {
    if (workspace) {
      if (workspace_bytes < P.ws_bytes)
        return set_err(FM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(P.ws_bytes) + " bytes");
      b->ws = workspace;
    } else {
      hipError_t e = hipMalloc(&b->ws, P.ws_bytes);
      if (e != hipSuccess) {
        b->ws = nullptr;
        return hip_err(e, "hipMalloc(batch workspace)");
      }
      b->own_ws = true;
    }
  }
  if (use_graph) {
    hipStream_t cs;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_err(e, "hipStreamCreate");
    const bool prof = g_prof_on;
    g_prof_on = false;  // no event records inside the graph
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      st = P.run(nullptr, nullptr, nullptr, b->ws, cs);
      hipError_t e2 = hipStreamEndCapture(cs, &b->graph);
      if (st == FM_OK && e2 != hipSuccess) st = hip_err(e2, "hipStreamEndCapture");
      if (st == FM_OK) {
        e2 = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e2 != hipSuccess) st = hip_err(e2, "hipGraphInstantiate");
      }
    } else {
      st = hip_err(e, "hipStreamBeginCapture");
    }
    g_prof_on = prof;
    (void)hipStreamDestroy(cs);
    if (st != FM_OK) return st;
  }
  *out = b.release();
  return FM_OK;
}

fm_status fm_batch_run(fm_batch b, void* stream, float* ms_out) {
  clear_error();
  if (!b) return set_err(FM_ERR_INVALID_ARG, "null batch");
  if (ms_out) *ms_out = 0.f;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms_out) {
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
      return set_err(FM_ERR_HIP, "hipEventCreate");
    (void)hipEventRecord(e0, s);
  }
  fm_status st = FM_OK;
  if (b->exec) {
    hipError_t e = hipGraphLaunch(b->exec, s);
    if (e != hipSuccess) st = hip_err(e, "hipGraphLaunch");
  } else {
    std::lock_guard<std::mutex> lock(fm_api_mutex());
    st = b->plan.run(nullptr, nullptr, nullptr, b->ws, s);
  }
  if (ms_out) {
    (void)hipEventRecord(e1, s);
    if (st == FM_OK) {
      hipError_t e = hipEventSynchronize(e1);
      if (e == hipSuccess) (void)hipEventElapsedTime(ms_out, e0, e1);
      else st = hip_err(e, "hipEventSynchronize");
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return st;
}

fm_status fm_batch_choice(fm_batch b, int index, const char** alg, int* num_steps) {
  if (!b || index < 0 || index >= (int)b->steps.size())
    return set_err(FM_ERR_INVALID_ARG, "bad batch index");
  if (alg) *alg = b->alg_names[index].c_str();
  if (num_steps) *num_steps = b->steps[index];
  return FM_OK;
}

size_t fm_batch_workspace_bytes(fm_batch b) { return b ? b->plan.ws_bytes : 0; }

void fm_batch_destroy(fm_batch b) { delete b; }

}  // extern "C"
