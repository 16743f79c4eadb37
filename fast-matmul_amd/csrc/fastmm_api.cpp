// fastmm_api.cpp -- C ABI of the fast-matmul engine: argument checks, plan
// cache, workspace arena, timing and registry.  Plans are built and run by
// plan.cpp; batches live in batch.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"
#include "scaling.hpp"

using namespace fmplan;

std::mutex& fm_api_mutex() {
  static std::mutex mu;
  return mu;
}

const FmAlgDesc* fm_find_alg(const char* name) {
  if (!name) return nullptr;
  for (int i = 0; i < fm_num_all_algorithms; ++i)
    if (strcmp(fm_all_algorithms[i]()->name, name) == 0) return fm_all_algorithms[i]();
  return nullptr;
}

namespace {

// Plan cache, bounded: the key holds the shape, leading dimensions, alpha,
// beta and x, so a caller that varies alpha from call to call would
// otherwise grow it without end.  The oldest plan goes first.
constexpr size_t kMaxPlans = 256;
std::map<std::string, std::unique_ptr<Plan>> g_plans;
std::deque<std::string> g_plan_order;
struct Arena {
  void* p = nullptr;
  size_t bytes = 0;
  uint64_t used = 0;  // last use, for eviction
};
// one cached workspace per (device, stream): calls on one stream are
// ordered by it, calls on different streams never share a workspace
std::map<std::pair<int, hipStream_t>, Arena> g_arenas;

std::string plan_key(int dev, const char* alg, int dt, int mode, int64_t m, int64_t k, int64_t n,
                     int64_t lda, int64_t ldb, int64_t ldc, int steps, double x, double alpha,
                     double beta, const fm_options& o) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%d|%s|%d|%d|%d|%d|%d|%d|%zu|%d|%lld|%lld|%lld|%lld|%lld|%lld|%d|%a|%a|%a",
           dev, alg ? alg : "<gemm>", dt, mode, mode == FM_MODE_OZAKI_INT8 ? o.ozaki_slices : 0,
           o.stage_fusion, o.split_remainder != 0, o.dfs_levels, o.workspace_limit, o.gemm_tile,
           (long long)m, (long long)k, (long long)n, (long long)lda, (long long)ldb,
           (long long)ldc, steps, x, alpha, beta);
  return buf;
}

}  // namespace

fm_status fmplan::check_args(fm_dtype dt, int m, int k, int n, const void* A, int64_t lda, const void* B,
                     int64_t ldb, const void* C, int64_t ldc) {
  if (dt != FM_F64 && dt != FM_F32) return set_err(FM_ERR_INVALID_ARG, "unknown dtype");
  if (m < 0 || k < 0 || n < 0) return set_err(FM_ERR_INVALID_ARG, "negative dimension");
  if (lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return set_err(FM_ERR_INVALID_ARG, "leading dimension too small");
  if ((m > 0 && n > 0) && (!C || (k > 0 && (!A || !B))))
    return set_err(FM_ERR_INVALID_ARG, "null matrix pointer");
  return FM_OK;
}

fm_status fmplan::check_mode(fm_dtype dt, fm_mode mode) {
  if (mode == FM_MODE_PLAIN) return FM_OK;
  if (mode == FM_MODE_SPLIT_BF16 && dt == FM_F32) return FM_OK;
  if (mode == FM_MODE_OZAKI_INT8 && dt == FM_F64) return FM_OK;
  return set_err(FM_ERR_UNSUPPORTED, mode == FM_MODE_SPLIT_BF16   ? "split_bf16 leaves need FP32 data"
                                     : mode == FM_MODE_OZAKI_INT8 ? "ozaki_int8 leaves need FP64 data"
                                                                  : "unknown leaf mode");
}

// One multiply's plan under options `o`.  Depth-first top levels: as many
// as asked, then more from the top while the plan's workspace exceeds the
// limit (the reference's DFS / HYBRID schedules, util/par_util.hpp:17-95,
// chosen per plan instead of at compile time).
fm_status fmplan::build_single(Plan& P, const FmAlgDesc* alg, fm_dtype dt, fm_mode mode, int64_t m,
                               int64_t k, int64_t n, int64_t lda, int64_t ldb, int64_t ldc,
                               int steps, double x, double alpha, double beta,
                               const fm_options& o) {
  // every launch numbers its work groups (tiles) in 32 bits
  if (((m + 63) / 64) * ((n + 63) / 64) > 2147483647LL)
    return set_err(FM_ERR_SHAPE, "product too large: more than 2^31 - 1 output tiles of 64 x 64");
  P.dtype = dt;
  P.mode = mode;
  P.slices = mode == FM_MODE_OZAKI_INT8 ? o.ozaki_slices : 0;
  P.tile = o.gemm_tile;
  P.split_rem = o.split_remainder != 0;
  Problem p;
  p.alg = alg;
  p.steps = steps;
  p.m = m;
  p.k = k;
  p.n = n;
  p.lda = lda;
  p.ldb = ldb;
  p.ldc = ldc;
  p.x = x;
  p.alpha = alpha;
  p.beta = beta;
  p.fuse = o.stage_fusion != 0;
  p.fuse_planes = o.stage_fusion == 1 && mode == FM_MODE_SPLIT_BF16 &&
                  (P.tile < 0 || fm_tile_planes(P.tile));
  p.mode = mode;
  p.slices = P.slices;
  p.tile = P.tile;
  p.split_rem = P.split_rem;
  size_t limit = o.workspace_limit;
  if (limit == 0 && o.dfs_levels < 0) {
    // automatic cut-off with no explicit limit: keep the workspace under a
    // quarter of the memory free now
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) limit = std::max<size_t>(free_b / 4, 1);
  }
  for (p.dfs = alg ? std::min(std::max(o.dfs_levels, 0), steps) : 0;; ++p.dfs) {
    Plan trial;
    trial.dtype = P.dtype;
    trial.mode = P.mode;
    trial.slices = P.slices;
    trial.tile = P.tile;
    trial.split_rem = P.split_rem;
    std::vector<SubPlan> subs(1);
    fm_status st = build_subplan(dt, p, 0, subs[0]);
    if (st != FM_OK) return st;
    st = assemble(trial, subs, subs[0].ws_elems);
    if (st != FM_OK) return st;
    const bool last = !alg || p.dfs >= steps;
    if (limit == 0 || trial.ws_bytes <= limit || last) {
      if (limit > 0 && trial.ws_bytes > limit)
        return set_err(FM_ERR_WORKSPACE,
                       "workspace limit below the depth-first plan's " +
                           std::to_string(trial.ws_bytes) + " bytes");
      P.launches = std::move(trial.launches);
      P.ws_bytes = trial.ws_bytes;
      P.dfs_levels = alg ? p.dfs : 0;
      return FM_OK;
    }
  }
}

namespace {

fm_status get_plan(const char* alg_name, const FmAlgDesc* alg, fm_dtype dt, fm_mode mode, int m,
                   int k, int n, int64_t lda, int64_t ldb, int64_t ldc, int steps, double x,
                   double alpha, double beta, const fm_options& o, Plan** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_err(e, "hipGetDevice");
  const std::string key =
      plan_key(dev, alg_name, dt, mode, m, k, n, lda, ldb, ldc, steps, x, alpha, beta, o);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    *out = it->second.get();
    return FM_OK;
  }
  std::unique_ptr<Plan> P(new Plan());
  fm_status st = build_single(*P, alg, dt, mode, m, k, n, lda, ldb, ldc, steps, x, alpha, beta, o);
  if (st != FM_OK) return st;
  st = P->upload();
  if (st != FM_OK) return st;
  if (g_plans.size() >= kMaxPlans) {
    // kernels already queued may still read the evicted plan's descriptors
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_err(e, "hipDeviceSynchronize");
    while (g_plans.size() >= kMaxPlans && !g_plan_order.empty()) {
      g_plans.erase(g_plan_order.front());
      g_plan_order.pop_front();
    }
  }
  *out = P.get();
  g_plans[key] = std::move(P);
  g_plan_order.push_back(key);
  return FM_OK;
}

// Cached arenas of one device are bounded together by half the device's
// memory: an arena that must grow past the bound first frees the least
// recently used arenas of other streams (after a device synchronisation,
// since queued kernels may still use them).
uint64_t g_arena_clock = 0;

fm_status arena_get(size_t bytes, hipStream_t stream, void** p) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_err(e, "hipGetDevice");
  Arena& a = g_arenas[std::make_pair(dev, stream)];
  a.used = ++g_arena_clock;
  if (a.bytes < bytes) {
    size_t free_b = 0, total_b = 0;
    e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) return hip_err(e, "hipMemGetInfo");
    const size_t cap = total_b / 2;
    size_t others = 0;
    for (auto& kv : g_arenas)
      if (kv.first.first == dev && &kv.second != &a) others += kv.second.bytes;
    const bool evict = others + bytes > cap;
    if (a.p || evict) {
      e = hipDeviceSynchronize();
      if (e != hipSuccess) return hip_err(e, "hipDeviceSynchronize");
    }
    if (a.p) {
      (void)hipFree(a.p);
      a.p = nullptr;
      a.bytes = 0;
    }
    while (evict && others + bytes > cap) {
      Arena* lru = nullptr;
      for (auto& kv : g_arenas)
        if (kv.first.first == dev && &kv.second != &a && kv.second.p &&
            (!lru || kv.second.used < lru->used))
          lru = &kv.second;
      if (!lru) break;
      others -= lru->bytes;
      (void)hipFree(lru->p);
      lru->p = nullptr;
      lru->bytes = 0;
    }
    e = hipMalloc(&a.p, bytes);
    if (e != hipSuccess) {
      a.p = nullptr;
      return hip_err(e, "hipMalloc(workspace)");
    }
    a.bytes = bytes;
  }
  *p = a.p;
  return FM_OK;
}

// Run `fn` (launches on stream s).  With ms_out, bracket it with events and
// wait for the second; the library lock is released before the wait, so a
// timed call does not hold other threads' calls back while the device works.
template <class F>
fm_status timed(hipStream_t s, float* ms_out, std::unique_lock<std::mutex>& lock, F&& fn) {
  if (!ms_out) return fn();
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return hip_err(e, "hipEventCreate");
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_err(e, "hipEventCreate");
  }
  (void)hipEventRecord(e0, s);
  fm_status st = fn();
  (void)hipEventRecord(e1, s);
  if (lock.owns_lock()) lock.unlock();
  if (st == FM_OK) {
    e = hipEventSynchronize(e1);
    if (e != hipSuccess)
      st = hip_err(e, "hipEventSynchronize");
    else
      (void)hipEventElapsedTime(ms_out, e0, e1);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

// Workspace for a plan: the caller's buffer if given, else the arena.
// Algorithmic HBM bytes of one launch: every source element read once,
// every output written once (the minimum any schedule of that launch needs).
double launch_bytes(const Launch& L, double es) {
  const FmAlgDesc* g = L.alg;
  const double nodes = L.st.nodes;
  switch (L.kind) {
    case LK_S:
      return es * nodes * L.st.rows * L.st.cols * (g->m * g->k + g->ns);
    case LK_T:
      return es * nodes * L.st.rows * L.st.cols * (g->k * g->n + g->nt);
    case LK_S2:
    case LK_T2: {
      const bool a = L.kind == LK_S2;
      const int nb = a ? g->m * g->k : g->k * g->n;
      int mat = 0;
      for (int r = 0; r < g->R; ++r) mat += a ? g->a_kind[r] : g->b_kind[r];
      const double outs = (double)g->R * (a ? g->ns : g->nt) + (double)mat * (a ? g->a_nviews : g->b_nviews);
      // outputs as split-bf16 slice planes: 6 bytes an element
      return (double)nodes * L.st.rows2 * L.st.cols2 * (es * nb * nb + (L.planes ? 6.0 : es) * outs);
    }
    case LK_W:
      return es * nodes * L.st.rows * L.st.cols *
             (g->R + g->m * g->n * (L.st.beta != 0.0 ? 2.0 : 1.0));
    case LK_W2: {
      const double mn = g->m * g->n;
      return es * nodes * L.st.rows2 * L.st.cols2 *
             ((double)g->R * g->R + mn * mn * (L.st.beta != 0.0 ? 2.0 : 1.0));
    }
    case LK_NONE:
      return 0.0;
    case LK_GEMM: {
      double b = 0;
      for (const FmGemmDesc& d : L.gdescs)
        b += es * ((double)d.M * d.K + (double)d.K * d.N + (double)d.M * d.N * (d.beta != 0.0 ? 2 : 1));
      return b;
    }
    case LK_SPLIT: {
      double b = 0;
      for (const FmSplitDesc& d : L.sdescs)
        b += es * ((double)d.M * d.K + (double)d.K * d.N + (double)d.M * d.N * (d.beta != 0.0 ? 2 : 1));
      return b;
    }
    case LK_OZAKI: {
      double b = 0;
      for (const FmOzakiDesc& d : L.odescs)
        b += es * ((double)d.M * d.K + (double)d.K * d.N + (double)d.M * d.N * (d.beta != 0.0 ? 2 : 1));
      return b;
    }
  }
  return 0.0;
}

fm_status plan_workspace(const Plan& P, void* workspace, size_t workspace_bytes, hipStream_t stream,
                         void** ws) {
  *ws = workspace;
  if (P.ws_bytes == 0) return FM_OK;
  if (workspace) {
    if (workspace_bytes < P.ws_bytes)
      return set_err(FM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(P.ws_bytes) + " bytes");
    return FM_OK;
  }
  return arena_get(P.ws_bytes, stream, ws);
}

}  // namespace

extern "C" {

void fm_options_default(fm_options* out) {
  if (!out) return;
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  *out = g_defaults;
}

fm_status fm_fastmatmul_ex(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                           const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                           int64_t ldc, int num_steps, double x, double alpha, double beta,
                           const fm_options* opts, void* stream, void* workspace,
                           size_t workspace_bytes, float* ms_out) {
  clear_error();
  if (ms_out) *ms_out = 0.f;
  const FmAlgDesc* alg = fm_find_alg(alg_name);
  if (!alg)
    return set_err(FM_ERR_UNKNOWN_ALG, std::string("unknown algorithm: ") + (alg_name ? alg_name : "(null)"));
  fm_status st = check_args(dtype, m, k, n, A, lda, B, ldb, C, ldc);
  if (st != FM_OK) return st;
  if (num_steps < 0) return set_err(FM_ERR_INVALID_ARG, "negative num_steps");
  st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  if (opts && (st = check_options(*opts)) != FM_OK) return st;
  if (m == 0 || n == 0) return FM_OK;
  if (k == 0) return set_err(FM_ERR_SHAPE, "k == 0");
  std::unique_lock<std::mutex> lock(fm_api_mutex());
  const fm_options o = opts ? *opts : g_defaults;
  Plan* P = nullptr;
  st = get_plan(alg_name, alg, dtype, mode, m, k, n, lda, ldb, ldc, num_steps, x, alpha, beta, o, &P);
  if (st != FM_OK) return st;
  void* ws = nullptr;
  const hipStream_t s = (hipStream_t)stream;
  st = plan_workspace(*P, workspace, workspace_bytes, s, &ws);
  if (st != FM_OK) return st;
  return timed(s, ms_out, lock, [&]() { return P->run(A, B, C, ws, s); });
}

fm_status fm_fastmatmul(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                        const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                        int64_t ldc, int num_steps, double x, double alpha, double beta,
                        void* stream, void* workspace, size_t workspace_bytes, float* ms_out) {
  return fm_fastmatmul_ex(alg_name, dtype, mode, m, k, n, A, lda, B, ldb, C, ldc, num_steps, x,
                          alpha, beta, nullptr, stream, workspace, workspace_bytes, ms_out);
}

}  // extern "C"

// ---- reusable plans -----------------------------------------------------------
struct fm_plan_s {
  Plan plan;
  int m = 0, k = 0, n = 0;
};

extern "C" {

fm_status fm_plan_create(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                         int64_t lda, int64_t ldb, int64_t ldc, int num_steps, double x,
                         double alpha, double beta, const fm_options* opts, fm_plan* out) {
  clear_error();
  if (!out) return set_err(FM_ERR_INVALID_ARG, "null output");
  *out = nullptr;
  const FmAlgDesc* alg = nullptr;
  if (alg_name) {
    alg = fm_find_alg(alg_name);
    if (!alg) return set_err(FM_ERR_UNKNOWN_ALG, std::string("unknown algorithm: ") + alg_name);
  }
  // shape checks as for a call with non-null matrices
  const char dummy = 0;
  fm_status st = check_args(dtype, m, k, n, &dummy, lda, &dummy, ldb, &dummy, ldc);
  if (st != FM_OK) return st;
  if (m == 0 || n == 0 || k == 0) return set_err(FM_ERR_SHAPE, "empty dimension");
  if (num_steps < 0) return set_err(FM_ERR_INVALID_ARG, "negative num_steps");
  st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  if (opts && (st = check_options(*opts)) != FM_OK) return st;
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  const fm_options o = opts ? *opts : g_defaults;
  std::unique_ptr<fm_plan_s> h(new fm_plan_s());
  st = build_single(h->plan, alg, dtype, mode, m, k, n, lda, ldb, ldc, alg ? num_steps : 0, x,
                    alpha, beta, o);
  if (st != FM_OK) return st;
  st = h->plan.upload();
  if (st != FM_OK) return st;
  h->m = m;
  h->k = k;
  h->n = n;
  *out = h.release();
  return FM_OK;
}

fm_status fm_plan_run(fm_plan h, const void* A, const void* B, void* C, void* stream,
                      void* workspace, size_t workspace_bytes, float* ms_out) {
  clear_error();
  if (ms_out) *ms_out = 0.f;
  if (!h) return set_err(FM_ERR_INVALID_ARG, "null plan");
  if (!A || !B || !C) return set_err(FM_ERR_INVALID_ARG, "null matrix pointer");
  const hipStream_t s = (hipStream_t)stream;
  std::unique_lock<std::mutex> lock(fm_api_mutex());
  void* ws = nullptr;
  fm_status st = plan_workspace(h->plan, workspace, workspace_bytes, s, &ws);
  if (st != FM_OK) return st;
  return timed(s, ms_out, lock, [&]() { return h->plan.run(A, B, C, ws, s); });
}

fm_status fm_plan_workspace(fm_plan h, size_t* bytes) {
  if (!h || !bytes) return set_err(FM_ERR_INVALID_ARG, "null argument");
  *bytes = h->plan.ws_bytes;
  return FM_OK;
}

void fm_plan_destroy(fm_plan h) {
  if (!h) return;
  // queued kernels may still read the plan's descriptors
  (void)hipDeviceSynchronize();
  delete h;
}

}  // extern "C"

namespace {
// Workspace layout of fm_fastmatmul_scaled: the plan's workspace, copies of
// A and B, the product (when beta != 0), the scaling vectors.
struct ScaledLayout {
  size_t o_a, o_b, o_t, o_r, o_s, o_f, o_f2, o_u1, o_u2, total;
};
ScaledLayout scaled_layout(size_t plan_ws, size_t es, int m, int k, int n, bool temp) {
  auto al = [](size_t v) { return (v + 255) / 256 * 256; };
  const size_t L = (size_t)std::max(m, std::max(k, n));
  ScaledLayout o;
  size_t off = al(plan_ws);
  o.o_a = off;
  off += al(es * (size_t)m * k);
  o.o_b = off;
  off += al(es * (size_t)k * n);
  o.o_t = off;
  if (temp) off += al(es * (size_t)m * n);
  o.o_r = off;
  off += al(es * m);
  o.o_s = off;
  off += al(es * n);
  o.o_f = off;
  off += al(es * L);
  o.o_f2 = off;
  off += al(es * L);
  o.o_u1 = off;
  off += al(8 * L);
  o.o_u2 = off;
  off += al(8 * L);
  o.total = off;
  return o;
}
}  // namespace

extern "C" {

fm_status fm_fastmatmul_scaled(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k,
                               int n, const void* A, int64_t lda, const void* B, int64_t ldb,
                               void* C, int64_t ldc, int num_steps, double x, double alpha,
                               double beta, int scaling_type, int scaling_steps, void* stream,
                               void* workspace, size_t workspace_bytes, float* ms_out) {
  clear_error();
  if (ms_out) *ms_out = 0.f;
  const FmAlgDesc* alg = fm_find_alg(alg_name);
  if (!alg)
    return set_err(FM_ERR_UNKNOWN_ALG, std::string("unknown algorithm: ") + (alg_name ? alg_name : "(null)"));
  fm_status st = check_args(dtype, m, k, n, A, lda, B, ldb, C, ldc);
  if (st != FM_OK) return st;
  if (num_steps < 0) return set_err(FM_ERR_INVALID_ARG, "negative num_steps");
  if (scaling_type < 0 || scaling_type > 3 || scaling_steps < 0)
    return set_err(FM_ERR_INVALID_ARG, "scaling_type must be 0..3 and scaling_steps >= 0");
  st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  if (m == 0 || n == 0) return FM_OK;
  if (k == 0) return set_err(FM_ERR_SHAPE, "k == 0");
  std::unique_lock<std::mutex> lock(fm_api_mutex());
  // the multiply of the scaled copies: contiguous operands, product into C
  // (beta = 0) or into a workspace temporary
  const bool temp = beta != 0.0;
  Plan* P = nullptr;
  st = get_plan(alg_name, alg, dtype, mode, m, k, n, m, k, temp ? m : ldc, num_steps, x, 1.0, 0.0,
                g_defaults, &P);
  if (st != FM_OK) return st;
  const ScaledLayout lay = scaled_layout(P->ws_bytes, (size_t)elem_size(dtype), m, k, n, temp);
  void* ws = workspace;
  if (workspace) {
    if (workspace_bytes < lay.total)
      return set_err(FM_ERR_WORKSPACE, "workspace too small: need " + std::to_string(lay.total) + " bytes");
  } else {
    st = arena_get(lay.total, (hipStream_t)stream, &ws);
    if (st != FM_OK) return st;
  }
  char* w = (char*)ws;
  hipStream_t s = (hipStream_t)stream;
  const int f64 = dtype == FM_F64;
  return timed(s, ms_out, lock, [&]() -> fm_status {
    hipError_t e = fm_scaling_copy(f64, A, lda, w + lay.o_a, m, m, k, s);
    if (e == hipSuccess) e = fm_scaling_copy(f64, B, ldb, w + lay.o_b, k, k, n, s);
    if (e != hipSuccess) return hip_err(e, "scaling copy");
    FmScaleArgs sa;
    sa.A = w + lay.o_a;
    sa.B = w + lay.o_b;
    sa.lda = m;
    sa.ldb = k;
    sa.m = m;
    sa.k = k;
    sa.n = n;
    sa.type = scaling_type;
    sa.steps = scaling_steps;
    sa.r = w + lay.o_r;
    sa.s = w + lay.o_s;
    sa.f = w + lay.o_f;
    sa.f2 = w + lay.o_f2;
    sa.u1 = w + lay.o_u1;
    sa.u2 = w + lay.o_u2;
    e = fm_scaling_run(f64, sa, s);
    if (e != hipSuccess) return hip_err(e, "scaling");
    void* prod = temp ? (void*)(w + lay.o_t) : C;
    const int64_t ldp = temp ? m : ldc;
    fm_status r = P->run(w + lay.o_a, w + lay.o_b, prod, w, s);
    if (r != FM_OK) return r;
    e = fm_scaling_post(f64, prod, ldp, C, ldc, m, n, w + lay.o_r, w + lay.o_s, alpha, beta, s);
    if (e != hipSuccess) return hip_err(e, "scaling post-process");
    return FM_OK;
  });
}

fm_status fm_scaled_workspace_query(const char* alg_name, fm_dtype dtype, fm_mode mode, int m,
                                    int k, int n, int num_steps, int beta_nonzero, size_t* bytes) {
  clear_error();
  if (!bytes) return set_err(FM_ERR_INVALID_ARG, "null output");
  const FmAlgDesc* alg = fm_find_alg(alg_name);
  if (!alg) return set_err(FM_ERR_UNKNOWN_ALG, "unknown algorithm");
  if (m <= 0 || k <= 0 || n <= 0 || num_steps < 0) return set_err(FM_ERR_INVALID_ARG, "bad shape");
  fm_status st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  const bool temp = beta_nonzero != 0;
  Plan P;
  // the same plan the scaled call builds (its product's C is contiguous
  // when beta != 0, else the caller's C with ldc >= m: the workspace does
  // not depend on ldc)
  st = build_single(P, alg, dtype, mode, m, k, n, m, k, m, num_steps, 1e-8, 1.0, 0.0, g_defaults);
  if (st != FM_OK) return st;
  *bytes = scaled_layout(P.ws_bytes, (size_t)elem_size(dtype), m, k, n, temp).total;
  return FM_OK;
}

fm_status fm_workspace_query(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k,
                             int n, int num_steps, size_t* bytes) {
  clear_error();
  if (!bytes) return set_err(FM_ERR_INVALID_ARG, "null output");
  const FmAlgDesc* alg = fm_find_alg(alg_name);
  if (!alg) return set_err(FM_ERR_UNKNOWN_ALG, "unknown algorithm");
  if (m <= 0 || k <= 0 || n <= 0 || num_steps < 0) return set_err(FM_ERR_INVALID_ARG, "bad shape");
  fm_status st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  Plan P;
  st = build_single(P, alg, dtype, mode, m, k, n, m, k, m, num_steps, 1e-8, 1.0, 0.0, g_defaults);
  if (st != FM_OK) return st;
  *bytes = P.ws_bytes;
  return FM_OK;
}

fm_status fm_plan_query(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                        int num_steps, fm_plan_info* info) {
  clear_error();
  if (!info) return set_err(FM_ERR_INVALID_ARG, "null output");
  const FmAlgDesc* alg = fm_find_alg(alg_name);
  if (!alg) return set_err(FM_ERR_UNKNOWN_ALG, "unknown algorithm");
  if (m <= 0 || k <= 0 || n <= 0 || num_steps < 0) return set_err(FM_ERR_INVALID_ARG, "bad shape");
  fm_status st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  Plan P;
  st = build_single(P, alg, dtype, mode, m, k, n, m, k, m, num_steps, 1e-8, 1.0, 0.0, g_defaults);
  if (st != FM_OK) return st;
  memset(info, 0, sizeof(*info));
  info->workspace_bytes = P.ws_bytes;
  info->dfs_levels = P.dfs_levels;
  int nleaf = 0;
  const double es = elem_size(dtype);
  for (const Launch& L : P.launches) {
    info->launches[L.stage] += 1;
    info->stage_bytes[L.stage] += launch_bytes(L, es);
    if (L.kind == LK_S2 || L.kind == LK_T2 || L.kind == LK_W2) info->two_level_stages += 1;
    if (L.kind == LK_SPLIT)
      for (const FmSplitPrep& r : L.spreps)
        info->prep_bytes += es * (double)r.rows * r.K + 6.0 * r.rows_pad * r.Kp;
    if (L.kind == LK_OZAKI)
      for (const FmOzakiDesc& d : L.odescs)
        info->prep_bytes += es * ((double)d.M * d.K + (double)d.K * d.N) +
                            (double)L.oz_slices * ((double)d.rows_pad + d.cols_pad) * d.Kp +
                            4.0 * (d.rows_pad + d.cols_pad);
    if (L.stage == FM_STAGE_LEAF) {
      const int tiles = L.kind == LK_OZAKI   ? L.oz.total_tiles
                        : L.kind == LK_SPLIT ? L.sp.total_tiles
                                             : L.gm.total_tiles;
      const int prods = L.kind == LK_OZAKI ? L.oz.ndesc : L.kind == LK_SPLIT ? L.sp.ndesc : L.gm.ndesc;
      if (nleaf < 2) {
        info->leaf_cfg[nleaf] = L.kind == LK_OZAKI ? -1 : L.cfg;
        info->leaf_products[nleaf] = prods;
        info->leaf_tiles[nleaf] = tiles;
      }
      ++nleaf;
    }
  }
  return FM_OK;
}

fm_status fm_gemm(fm_dtype dtype, fm_mode mode, int m, int k, int n, const void* A, int64_t lda,
                  const void* B, int64_t ldb, void* C, int64_t ldc, double alpha, double beta,
                  void* stream) {
  clear_error();
  fm_status st = check_args(dtype, m, k, n, A, lda, B, ldb, C, ldc);
  if (st != FM_OK) return st;
  st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  if (m == 0 || n == 0) return FM_OK;
  if (k == 0) return set_err(FM_ERR_SHAPE, "k == 0");
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  Plan* P = nullptr;
  st = get_plan(nullptr, nullptr, dtype, mode, m, k, n, lda, ldb, ldc, 0, 0.0, alpha, beta, g_defaults,
                &P);
  if (st != FM_OK) return st;
  void* ws = nullptr;
  st = plan_workspace(*P, nullptr, 0, (hipStream_t)stream, &ws);
  if (st != FM_OK) return st;
  return P->run(A, B, C, ws, (hipStream_t)stream);
}

int fm_num_algorithms(void) { return fm_num_all_algorithms; }

const char* fm_algorithm_name(int index) {
  if (index < 0 || index >= fm_num_all_algorithms) return nullptr;
  return fm_all_algorithms[index]()->name;
}

fm_status fm_algorithm_info(const char* alg_name, int* m, int* k, int* n, int* rank) {
  const FmAlgDesc* alg = fm_find_alg(alg_name);
  if (!alg) return set_err(FM_ERR_UNKNOWN_ALG, "unknown algorithm");
  if (m) *m = alg->m;
  if (k) *k = alg->k;
  if (n) *n = alg->n;
  if (rank) *rank = alg->R;
  return FM_OK;
}

const char* fm_status_string(fm_status s) {
  switch (s) {
    case FM_OK: return "ok";
    case FM_ERR_INVALID_ARG: return "invalid argument";
    case FM_ERR_UNKNOWN_ALG: return "unknown algorithm";
    case FM_ERR_SHAPE: return "shape not supported at this depth";
    case FM_ERR_WORKSPACE: return "workspace too small";
    case FM_ERR_HIP: return "HIP runtime error";
    case FM_ERR_UNSUPPORTED: return "unsupported dtype/mode";
  }
  return "unknown status";
}

const char* fm_last_error(void) { return last_error().c_str(); }

// The process defaults of fm_options.  Plans are cached under their
// options, so changing a default only affects calls made afterwards.
fm_status fm_set_ozaki_slices(int slices) {
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  if (slices < 6 || slices > 8) return set_err(FM_ERR_INVALID_ARG, "slices must be 6, 7 or 8");
  g_defaults.ozaki_slices = slices;
  return FM_OK;
}

fm_status fm_set_gemm_tile(int cfg) {
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  if (cfg < -1 || cfg >= FM_TILE_NUM) return set_err(FM_ERR_INVALID_ARG, "unknown tile config");
  g_defaults.gemm_tile = cfg;
  return FM_OK;
}

fm_status fm_set_schedule(int dfs_levels, size_t workspace_limit, int split_remainder) {
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  if (dfs_levels < -1) return set_err(FM_ERR_INVALID_ARG, "dfs_levels must be >= -1");
  g_defaults.dfs_levels = dfs_levels;
  g_defaults.workspace_limit = workspace_limit;
  g_defaults.split_remainder = split_remainder != 0;
  return FM_OK;
}

void fm_set_stage_fusion(int on) {
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  g_defaults.stage_fusion = on == 2 ? 2 : on != 0;
}

void fm_profile_enable(int on) {
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  g_prof_on = on != 0;
}

fm_status fm_profile_read(fm_profile* out) {
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  return profile_read(out);
}

void fm_release_caches(void) {
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  (void)hipDeviceSynchronize();
  g_plans.clear();
  g_plan_order.clear();
  for (auto& kv : g_arenas)
    if (kv.second.p) (void)hipFree(kv.second.p);
  g_arenas.clear();
}

void fm_release_workspace(void* stream) {
  std::lock_guard<std::mutex> lock(fm_api_mutex());
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  auto it = g_arenas.find(std::make_pair(dev, (hipStream_t)stream));
  if (it == g_arenas.end()) return;
  (void)hipDeviceSynchronize();
  if (it->second.p) (void)hipFree(it->second.p);
  g_arenas.erase(it);
}

}  // extern "C"
