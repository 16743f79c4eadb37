// fastmm_api.cpp -- C ABI, plan builder and executor of the fast-matmul engine.
//
// A plan is the reference's recursion (algorithms/strassen.hpp:378-659)
// unrolled level by level.  Level l holds R^l nodes, each an (A, B, C)
// operand triple plus the node's multiplier.  Because every node of a level
// has the same block shape (the floor split of linalg/linalg.hpp:179-184),
// each level becomes a handful of launches over all of its nodes:
//
//   for l = 0 .. L-1:   S stage (all nodes)  T stage (all nodes)
//   leaves:             one grouped MFMA GEMM over the R^L products
//   for l = L-1 .. 0:   W stage (C_ij = sum W M (+ beta C_ij))   peeling GEMMs
//
// which is the reference's BFS schedule (util/par_util.hpp:17-95,
// util/MemoryManager.hpp:65-110) with every task of a level in one launch.
// Multipliers follow the reference exactly: the root carries alpha, a product
// whose S (T) column has a single coefficient c multiplies its node's
// multiplier by c (codegen/gen.py:614-621), the leaf GEMM uses the node's
// multiplier as alpha (linalg/linalg.hpp:266) and so do its peeling products
// (linalg/linalg.hpp:201-244).  beta reaches only the root's W stage and
// peeling; children use 0 (algorithms/strassen.hpp:459).
//
// Addresses in a plan are (role, offset) pairs resolved at launch time, so a
// plan is built and uploaded once per shape and replayed for any buffers.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fastmm.h"
#include "alg_desc.hpp"
#include "gemm_mfma.hpp"

namespace {

thread_local std::string g_err;

fm_status set_err(fm_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

fm_status hip_err(hipError_t e, const char* what) {
  return set_err(FM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

struct Dims {
  int64_t m, k, n;
};

struct Node {
  FmOp a, b, c;
  double mu;
};

enum LaunchKind { LK_S, LK_T, LK_W, LK_GEMM };

struct Launch {
  LaunchKind kind;
  int stage = 0;  // fm_stage_kind for profiling
  FmStageArgs st;
  FmGemmArgs gm;
  int cfg = 0;
  int vec_ok = 1;
  size_t dev_off = 0;   // byte offset of this launch's ops/descs in the plan buffer
  size_t count = 0;     // number of ops/descs
};

struct Plan {
  const FmAlgDesc* alg = nullptr;
  fm_dtype dtype = FM_F64;
  fm_mode mode = FM_MODE_PLAIN;
  std::vector<Launch> launches;
  std::vector<char> host_buf;
  void* dbuf = nullptr;
  size_t ws_bytes = 0;
  ~Plan() {
    if (dbuf) (void)hipFree(dbuf);
  }
};

const int64_t kLdAlign = 16;    // elements; 128 B for FP64
const int64_t kSlabAlign = 64;  // elements; keeps every buffer 256 B aligned

int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

FmOp op_make(int role, int64_t off, int64_t ld) {
  FmOp o;
  o.role = role;
  o.pad = 0;
  o.off = off;
  o.ld = ld;
  return o;
}

FmOp op_sub(const FmOp& o, int64_t i, int64_t j) {
  return op_make(o.role, o.off + i + j * o.ld, o.ld);
}

double round_scalar(double v, fm_dtype dt) { return dt == FM_F32 ? (double)(float)v : v; }

int elem_size(fm_dtype dt) { return dt == FM_F64 ? 8 : 4; }

int g_tile_override = -1;  // fm_set_gemm_tile

// Choose the GEMM tile for `count` products of M x N: the smallest estimated
// time over the configurations, where a configuration's throughput per tile
// (eff, measured on MI355X at 8192^3) and its resident work groups per CU
// set how many rounds of tiles the launch takes.
int pick_tile(fm_dtype dt, fm_mode mode, int64_t M, int64_t N, int64_t count) {
  if (g_tile_override >= 0) return g_tile_override;
  if (mode == FM_MODE_SPLIT_BF16) {
    // 128x128 (4 waves) unless the launch has too few tiles to fill the chip
    const int64_t big = ((M + 127) / 128) * ((N + 127) / 128) * count;
    return big >= 512 ? FM_TILE_128x128 : FM_TILE_64x64;
  }
  struct C {
    int cfg;
    double eff;
    int per_cu;
  };
  static const C f64c[] = {{FM_TILE_128x128_W8, 1.00, 2}, {FM_TILE_256x128_W8, 0.94, 1},
                           {FM_TILE_128x256_W8, 0.94, 1}, {FM_TILE_64x64, 0.80, 4},
                           {FM_TILE_128x64, 0.77, 2},     {FM_TILE_64x128, 0.77, 2}};
  static const C f32c[] = {{FM_TILE_256x256_W16, 1.00, 1}, {FM_TILE_256x128_W8, 0.90, 2},
                           {FM_TILE_128x256_W8, 0.90, 2},  {FM_TILE_128x128_W8, 0.82, 4},
                           {FM_TILE_64x128, 0.77, 4},      {FM_TILE_128x64, 0.74, 4},
                           {FM_TILE_64x64, 0.69, 8}};
  const C* cands = dt == FM_F64 ? f64c : f32c;
  const int nc = dt == FM_F64 ? 6 : 7;
  int best = FM_TILE_64x64;
  double best_t = 1e300;
  for (int i = 0; i < nc; ++i) {
    int bm, bn;
    fm_tile_dims(cands[i].cfg, &bm, &bn);
    const int64_t tiles = ((M + bm - 1) / bm) * ((N + bn - 1) / bn) * count;
    const int64_t slots = 256LL * cands[i].per_cu;
    const double rounds = (double)((tiles + slots - 1) / slots);
    const double t = rounds * bm * bn * cands[i].per_cu / cands[i].eff;
    if (t < best_t * 0.999) {
      best_t = t;
      best = cands[i].cfg;
    }
  }
  return best;
}

void add_gemm_launch(Plan& P, std::vector<FmGemmDesc>& descs, int cfg, bool uniform) {
  int bm, bn;
  fm_tile_dims(cfg, &bm, &bn);
  int32_t t = 0;
  for (FmGemmDesc& d : descs) {
    d.tiles_m = (int32_t)((d.M + bm - 1) / bm);
    d.tiles_n = (int32_t)((d.N + bn - 1) / bn);
    d.tile_begin = t;
    t += d.tiles_m * d.tiles_n;
  }
  Launch L;
  L.kind = LK_GEMM;
  memset(&L.gm, 0, sizeof(L.gm));
  L.gm.ndesc = (int32_t)descs.size();
  L.gm.total_tiles = t;
  L.gm.uniform = uniform ? 1 : 0;
  // 16-byte loads need every A/B column start aligned: even (FP64) or
  // multiple-of-4 (FP32) offsets and leading dimensions; base pointers are
  // checked at launch
  const int vl = P.dtype == FM_F64 ? 2 : 4;
  bool vec = true;
  for (const FmGemmDesc& d : descs)
    vec = vec && d.a.off % vl == 0 && d.a.ld % vl == 0 && d.b.off % vl == 0 && d.b.ld % vl == 0;
  L.vec_ok = vec ? 1 : 0;
  L.stage = uniform ? FM_STAGE_LEAF : FM_STAGE_PEEL;
  L.cfg = cfg;
  L.count = descs.size();
  L.dev_off = align_up((int64_t)P.host_buf.size(), 256);
  P.host_buf.resize(L.dev_off + descs.size() * sizeof(FmGemmDesc));
  memcpy(P.host_buf.data() + L.dev_off, descs.data(), descs.size() * sizeof(FmGemmDesc));
  P.launches.push_back(L);
}

void add_stage_launch(Plan& P, LaunchKind kind, const std::vector<FmOp>& ops, int64_t rows,
                      int64_t cols, int64_t out_off, int64_t out_ld, int64_t out_slab, double x,
                      double beta) {
  Launch L;
  L.kind = kind;
  memset(&L.st, 0, sizeof(L.st));
  L.st.rows = (int32_t)rows;
  L.st.cols = (int32_t)cols;
  L.st.nodes = (int32_t)ops.size();
  L.st.out_off = out_off;
  L.st.out_ld = out_ld;
  L.st.out_slab = out_slab;
  L.st.x = x;
  L.st.beta = beta;
  L.stage = kind == LK_S ? FM_STAGE_S : kind == LK_T ? FM_STAGE_T : FM_STAGE_W;
  const int vmax = P.dtype == FM_F64 ? 2 : 4;
  bool ok = rows % vmax == 0 && out_ld % vmax == 0 && out_slab % vmax == 0 &&
            out_off % vmax == 0;
  for (const FmOp& o : ops) ok = ok && o.off % vmax == 0 && o.ld % vmax == 0;
  L.vec_ok = ok ? vmax : 1;
  L.count = ops.size();
  L.dev_off = align_up((int64_t)P.host_buf.size(), 256);
  P.host_buf.resize(L.dev_off + ops.size() * sizeof(FmOp));
  memcpy(P.host_buf.data() + L.dev_off, ops.data(), ops.size() * sizeof(FmOp));
  P.launches.push_back(L);
}

// Peeling products of one node (linalg/linalg.hpp:201-244).
void peel_descs(const Node& nd, const Dims& D, int am, int ak, int an, double beta,
                std::vector<FmGemmDesc>& out) {
  const int64_t erA = D.m % am, ecA = D.k % ak, ecB = D.n % an;
  auto mk = [&](FmOp a, FmOp b, FmOp c, int64_t M, int64_t N, int64_t K, double bta) {
    if (M <= 0 || N <= 0) return;
    FmGemmDesc d;
    memset(&d, 0, sizeof(d));
    d.a = a;
    d.b = b;
    d.c = c;
    d.M = (int32_t)M;
    d.N = (int32_t)N;
    d.K = (int32_t)K;
    d.alpha = nd.mu;
    d.beta = bta;
    out.push_back(d);
  };
  if (ecA > 0) {
    const int64_t row_dim = D.m - erA, col_dim = D.n - ecB, kk = D.k - ecA;
    mk(op_sub(nd.a, 0, kk), op_sub(nd.b, kk, 0), nd.c, row_dim, col_dim, ecA, 1.0);
  }
  if (ecB > 0) {
    const int64_t s = D.n - ecB;
    mk(nd.a, op_sub(nd.b, 0, s), op_sub(nd.c, 0, s), D.m, ecB, D.k, beta);
  }
  if (erA > 0) {
    const int64_t s = D.m - erA, nc = D.n - ecB;
    mk(op_sub(nd.a, s, 0), nd.b, op_sub(nd.c, s, 0), erA, nc, D.k, beta);
  }
}

fm_status build_plan(Plan& P, const FmAlgDesc* alg, fm_dtype dt, fm_mode mode, int64_t M,
                     int64_t K, int64_t N, int64_t lda, int64_t ldb, int64_t ldc, int steps,
                     double x, double alpha, double beta) {
  P.alg = alg;
  P.dtype = dt;
  P.mode = mode;
  const int L = steps;
  std::vector<Dims> dims(L + 1);
  dims[0] = {M, K, N};
  for (int l = 0; l < L; ++l) {
    dims[l + 1] = {dims[l].m / alg->m, dims[l].k / alg->k, dims[l].n / alg->n};
    if (dims[l + 1].m <= 0 || dims[l + 1].k <= 0 || dims[l + 1].n <= 0)
      return set_err(FM_ERR_SHAPE, "num_steps too deep for this shape: a block dimension "
                                   "reaches 0 at level " + std::to_string(l + 1));
  }
  std::vector<std::vector<Node>> lv(L + 1);
  Node root;
  root.a = op_make(FM_ROLE_A, 0, lda);
  root.b = op_make(FM_ROLE_B, 0, ldb);
  root.c = op_make(FM_ROLE_C, 0, ldc);
  root.mu = round_scalar(alpha, dt);
  lv[0].push_back(root);

  struct LevelBuf {
    int64_t m_off, m_ld, m_slab;
  };
  std::vector<LevelBuf> lb(L);
  int64_t ws = 0;  // elements
  const int R = alg->R, NS = alg->ns, NT = alg->nt;
  for (int l = 0; l < L; ++l) {
    const Dims c = dims[l + 1];
    const int64_t Q = (int64_t)lv[l].size();
    const int64_t s_ld = align_up(c.m, kLdAlign), s_slab = align_up(s_ld * c.k, kSlabAlign);
    const int64_t t_ld = align_up(c.k, kLdAlign), t_slab = align_up(t_ld * c.n, kSlabAlign);
    const int64_t m_ld = align_up(c.m, kLdAlign), m_slab = align_up(m_ld * c.n, kSlabAlign);
    const int64_t s_off = ws;
    ws += Q * NS * s_slab;
    const int64_t t_off = ws;
    ws += Q * NT * t_slab;
    const int64_t m_off = ws;
    ws += Q * R * m_slab;
    lb[l] = {m_off, m_ld, m_slab};
    lv[l + 1].reserve(Q * R);
    for (int64_t q = 0; q < Q; ++q) {
      const Node& nd = lv[l][q];
      for (int r = 0; r < R; ++r) {
        Node ch;
        if (alg->a_kind[r])
          ch.a = op_make(FM_ROLE_WS, s_off + (q * NS + alg->a_index[r]) * s_slab, s_ld);
        else
          ch.a = op_sub(nd.a, (alg->a_index[r] / alg->k) * c.m, (alg->a_index[r] % alg->k) * c.k);
        if (alg->b_kind[r])
          ch.b = op_make(FM_ROLE_WS, t_off + (q * NT + alg->b_index[r]) * t_slab, t_ld);
        else
          ch.b = op_sub(nd.b, (alg->b_index[r] / alg->n) * c.k, (alg->b_index[r] % alg->n) * c.n);
        ch.c = op_make(FM_ROLE_WS, m_off + (q * R + r) * m_slab, m_ld);
        double mu = round_scalar(nd.mu * round_scalar(alg->a_coef(r, x), dt), dt);
        mu = round_scalar(mu * round_scalar(alg->b_coef(r, x), dt), dt);
        ch.mu = mu;
        lv[l + 1].push_back(ch);
      }
    }
    std::vector<FmOp> ops;
    if (NS > 0) {
      ops.clear();
      for (const Node& nd : lv[l]) ops.push_back(nd.a);
      add_stage_launch(P, LK_S, ops, c.m, c.k, s_off, s_ld, s_slab, x, 0.0);
    }
    if (NT > 0) {
      ops.clear();
      for (const Node& nd : lv[l]) ops.push_back(nd.b);
      add_stage_launch(P, LK_T, ops, c.k, c.n, t_off, t_ld, t_slab, x, 0.0);
    }
  }
  // leaves: one grouped product launch (the reference's base case MatMul with
  // beta 0, algorithms/strassen.hpp:385-388)
  {
    const Dims d = dims[L];
    std::vector<FmGemmDesc> descs;
    descs.reserve(lv[L].size());
    for (const Node& nd : lv[L]) {
      FmGemmDesc g;
      memset(&g, 0, sizeof(g));
      g.a = nd.a;
      g.b = nd.b;
      g.c = nd.c;
      g.M = (int32_t)d.m;
      g.N = (int32_t)d.n;
      g.K = (int32_t)d.k;
      g.alpha = nd.mu;
      g.beta = 0.0;
      descs.push_back(g);
    }
    add_gemm_launch(P, descs, pick_tile(dt, P.mode, d.m, d.n, (int64_t)descs.size()), true);
  }
  for (int l = L - 1; l >= 0; --l) {
    const Dims c = dims[l + 1];
    const double bl = l == 0 ? round_scalar(beta, dt) : 0.0;
    std::vector<FmOp> ops;
    for (const Node& nd : lv[l]) ops.push_back(nd.c);
    add_stage_launch(P, LK_W, ops, c.m, c.n, lb[l].m_off, lb[l].m_ld, lb[l].m_slab, x, bl);
    std::vector<FmGemmDesc> peel;
    for (const Node& nd : lv[l]) peel_descs(nd, dims[l], alg->m, alg->k, alg->n, bl, peel);
    if (!peel.empty()) add_gemm_launch(P, peel, FM_TILE_64x64, false);  // split mode: 64x64 too
  }
  P.ws_bytes = (size_t)ws * elem_size(dt);
  return FM_OK;
}

fm_status build_gemm_plan(Plan& P, fm_dtype dt, fm_mode mode, int64_t M, int64_t K, int64_t N,
                          int64_t lda, int64_t ldb, int64_t ldc, double alpha, double beta) {
  P.alg = nullptr;
  P.dtype = dt;
  P.mode = mode;
  FmGemmDesc g;
  memset(&g, 0, sizeof(g));
  g.a = op_make(FM_ROLE_A, 0, lda);
  g.b = op_make(FM_ROLE_B, 0, ldb);
  g.c = op_make(FM_ROLE_C, 0, ldc);
  g.M = (int32_t)M;
  g.N = (int32_t)N;
  g.K = (int32_t)K;
  g.alpha = round_scalar(alpha, dt);
  g.beta = round_scalar(beta, dt);
  std::vector<FmGemmDesc> descs(1, g);
  add_gemm_launch(P, descs, pick_tile(dt, P.mode, M, N, 1), true);
  P.ws_bytes = 0;
  return FM_OK;
}

fm_status upload_plan(Plan& P) {
  if (P.host_buf.empty()) return FM_OK;
  hipError_t e = hipMalloc(&P.dbuf, P.host_buf.size());
  if (e != hipSuccess) return hip_err(e, "hipMalloc(plan)");
  e = hipMemcpy(P.dbuf, P.host_buf.data(), P.host_buf.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_err(e, "hipMemcpy(plan)");
  for (Launch& L : P.launches) {
    if (L.kind == LK_GEMM)
      L.gm.descs = reinterpret_cast<const FmGemmDesc*>((char*)P.dbuf + L.dev_off);
    else
      L.st.ops = reinterpret_cast<const FmOp*>((char*)P.dbuf + L.dev_off);
  }
  return FM_OK;
}

// ---- profiling -------------------------------------------------------------
struct ProfRec {
  int stage;
  hipEvent_t e0, e1;
};
bool g_prof_on = false;
std::vector<ProfRec> g_prof_pending;
std::vector<hipEvent_t> g_event_pool;
fm_profile g_prof_acc;

hipEvent_t pool_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

fm_status run_plan(const Plan& P, const void* A, const void* B, void* C, void* ws,
                   hipStream_t s) {
  FmBases bases;
  bases.p[FM_ROLE_A] = const_cast<void*>(A);
  bases.p[FM_ROLE_B] = const_cast<void*>(B);
  bases.p[FM_ROLE_C] = C;
  bases.p[FM_ROLE_WS] = ws;
  bool aligned = true;
  for (int i = 0; i < 4; ++i) aligned = aligned && ((uintptr_t)bases.p[i] % 16 == 0);
  for (const Launch& L : P.launches) {
    hipError_t e = hipSuccess;
    ProfRec pr{L.stage, nullptr, nullptr};
    if (g_prof_on) {
      pr.e0 = pool_event();
      pr.e1 = pool_event();
      if (pr.e0) (void)hipEventRecord(pr.e0, s);
    }
    if (L.kind == LK_GEMM) {
      FmGemmArgs g = L.gm;
      g.bases = bases;
      g.vec = (L.vec_ok && aligned) ? 1 : 0;
      if (P.mode == FM_MODE_SPLIT_BF16)
        e = fm_gemm_launch_bf16x3(L.cfg, g, s);
      else
        e = P.dtype == FM_F64 ? fm_gemm_launch_f64(L.cfg, g, s) : fm_gemm_launch_f32(L.cfg, g, s);
    } else {
      FmStageArgs a = L.st;
      a.bases = bases;
      a.vec = aligned ? L.vec_ok : 1;
      const FmAlgDesc* g = P.alg;
      FmStageLaunch fn;
      if (P.dtype == FM_F64)
        fn = L.kind == LK_S ? g->s_d : L.kind == LK_T ? g->t_d : g->w_d;
      else
        fn = L.kind == LK_S ? g->s_f : L.kind == LK_T ? g->t_f : g->w_f;
      e = fn(a, s);
    }
    if (g_prof_on && pr.e0 && pr.e1) {
      (void)hipEventRecord(pr.e1, s);
      g_prof_pending.push_back(pr);
    }
    if (e != hipSuccess) return hip_err(e, "kernel launch");
  }
  return FM_OK;
}

// ---- caches ---------------------------------------------------------------
std::mutex g_mu;
std::map<std::string, std::unique_ptr<Plan>> g_plans;
struct Arena {
  void* p = nullptr;
  size_t bytes = 0;
};
std::map<int, Arena> g_arenas;

std::string plan_key(int dev, const char* alg, int dt, int mode, int64_t m, int64_t k, int64_t n,
                     int64_t lda, int64_t ldb, int64_t ldc, int steps, double x, double alpha,
                     double beta) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%d|%s|%d|%d|%lld|%lld|%lld|%lld|%lld|%lld|%d|%a|%a|%a", dev,
           alg ? alg : "<gemm>", dt, mode, (long long)m, (long long)k, (long long)n,
           (long long)lda, (long long)ldb, (long long)ldc, steps, x, alpha, beta);
  return buf;
}

const FmAlgDesc* find_alg(const char* name) {
  if (!name) return nullptr;
  for (int i = 0; i < fm_num_all_algorithms; ++i)
    if (strcmp(fm_all_algorithms[i]()->name, name) == 0) return fm_all_algorithms[i]();
  return nullptr;
}

fm_status check_args(fm_dtype dt, int m, int k, int n, const void* A, int64_t lda,
                     const void* B, int64_t ldb, const void* C, int64_t ldc) {
  if (dt != FM_F64 && dt != FM_F32) return set_err(FM_ERR_INVALID_ARG, "unknown dtype");
  if (m < 0 || k < 0 || n < 0) return set_err(FM_ERR_INVALID_ARG, "negative dimension");
  if (lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
    return set_err(FM_ERR_INVALID_ARG, "leading dimension too small");
  if ((m > 0 && n > 0) && (!C || (k > 0 && (!A || !B))))
    return set_err(FM_ERR_INVALID_ARG, "null matrix pointer");
  return FM_OK;
}

fm_status check_mode(fm_dtype dt, fm_mode mode) {
  if (mode == FM_MODE_PLAIN) return FM_OK;
  if (mode == FM_MODE_SPLIT_BF16 && dt == FM_F32) return FM_OK;
  return set_err(FM_ERR_UNSUPPORTED, mode == FM_MODE_SPLIT_BF16
                                         ? "split_bf16 leaves need FP32 data"
                                         : "this leaf mode is not available for this dtype");
}

// Find or build the plan; returns it through *out.
fm_status get_plan(const char* alg_name, const FmAlgDesc* alg, fm_dtype dt, fm_mode mode, int m,
                   int k, int n, int64_t lda, int64_t ldb, int64_t ldc, int steps, double x,
                   double alpha, double beta, Plan** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_err(e, "hipGetDevice");
  const std::string key =
      plan_key(dev, alg_name, dt, mode, m, k, n, lda, ldb, ldc, steps, x, alpha, beta);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    *out = it->second.get();
    return FM_OK;
  }
  std::unique_ptr<Plan> P(new Plan());
  fm_status st = alg ? build_plan(*P, alg, dt, mode, m, k, n, lda, ldb, ldc, steps, x, alpha, beta)
                     : build_gemm_plan(*P, dt, mode, m, k, n, lda, ldb, ldc, alpha, beta);
  if (st != FM_OK) return st;
  st = upload_plan(*P);
  if (st != FM_OK) return st;
  *out = P.get();
  g_plans[key] = std::move(P);
  return FM_OK;
}

fm_status arena_get(size_t bytes, void** p) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_err(e, "hipGetDevice");
  Arena& a = g_arenas[dev];
  if (a.bytes < bytes) {
    if (a.p) {
      e = hipDeviceSynchronize();
      if (e != hipSuccess) return hip_err(e, "hipDeviceSynchronize");
      (void)hipFree(a.p);
      a.p = nullptr;
      a.bytes = 0;
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

fm_status timed_run(const Plan& P, const void* A, const void* B, void* C, void* ws,
                    hipStream_t s, float* ms_out) {
  if (!ms_out) return run_plan(P, A, B, C, ws, s);
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return hip_err(e, "hipEventCreate");
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_err(e, "hipEventCreate");
  }
  (void)hipEventRecord(e0, s);
  fm_status st = run_plan(P, A, B, C, ws, s);
  (void)hipEventRecord(e1, s);
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

}  // namespace

extern "C" {

fm_status fm_fastmatmul(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                        const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                        int64_t ldc, int num_steps, double x, double alpha, double beta,
                        void* stream, void* workspace, size_t workspace_bytes, float* ms_out) {
  g_err.clear();
  if (ms_out) *ms_out = 0.f;
  const FmAlgDesc* alg = find_alg(alg_name);
  if (!alg) return set_err(FM_ERR_UNKNOWN_ALG, std::string("unknown algorithm: ") +
                                                   (alg_name ? alg_name : "(null)"));
  fm_status st = check_args(dtype, m, k, n, A, lda, B, ldb, C, ldc);
  if (st != FM_OK) return st;
  if (num_steps < 0) return set_err(FM_ERR_INVALID_ARG, "negative num_steps");
  st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  if (m == 0 || n == 0) return FM_OK;
  if (k == 0) return set_err(FM_ERR_SHAPE, "k == 0");
  std::lock_guard<std::mutex> lock(g_mu);
  Plan* P = nullptr;
  st = get_plan(alg_name, alg, dtype, mode, m, k, n, lda, ldb, ldc, num_steps, x, alpha, beta, &P);
  if (st != FM_OK) return st;
  void* ws = workspace;
  if (P->ws_bytes > 0) {
    if (ws) {
      if (workspace_bytes < P->ws_bytes)
        return set_err(FM_ERR_WORKSPACE, "workspace too small: need " +
                                             std::to_string(P->ws_bytes) + " bytes");
    } else {
      st = arena_get(P->ws_bytes, &ws);
      if (st != FM_OK) return st;
    }
  }
  return timed_run(*P, A, B, C, ws, (hipStream_t)stream, ms_out);
}

fm_status fm_workspace_query(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k,
                             int n, int num_steps, size_t* bytes) {
  g_err.clear();
  if (!bytes) return set_err(FM_ERR_INVALID_ARG, "null output");
  const FmAlgDesc* alg = find_alg(alg_name);
  if (!alg) return set_err(FM_ERR_UNKNOWN_ALG, "unknown algorithm");
  if (m <= 0 || k <= 0 || n <= 0 || num_steps < 0)
    return set_err(FM_ERR_INVALID_ARG, "bad shape");
  fm_status st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  Plan P;
  st = build_plan(P, alg, dtype, mode, m, k, n, m, k, m, num_steps, 1e-8, 1.0, 0.0);
  if (st != FM_OK) return st;
  *bytes = P.ws_bytes;
  return FM_OK;
}

fm_status fm_gemm(fm_dtype dtype, fm_mode mode, int m, int k, int n, const void* A, int64_t lda,
                  const void* B, int64_t ldb, void* C, int64_t ldc, double alpha, double beta,
                  void* stream) {
  g_err.clear();
  fm_status st = check_args(dtype, m, k, n, A, lda, B, ldb, C, ldc);
  if (st != FM_OK) return st;
  st = check_mode(dtype, mode);
  if (st != FM_OK) return st;
  if (m == 0 || n == 0) return FM_OK;
  if (k == 0) return set_err(FM_ERR_SHAPE, "k == 0");
  std::lock_guard<std::mutex> lock(g_mu);
  Plan* P = nullptr;
  st = get_plan(nullptr, nullptr, dtype, mode, m, k, n, lda, ldb, ldc, 0, 0.0, alpha, beta, &P);
  if (st != FM_OK) return st;
  return run_plan(*P, A, B, C, nullptr, (hipStream_t)stream);
}

int fm_num_algorithms(void) { return fm_num_all_algorithms; }

const char* fm_algorithm_name(int index) {
  if (index < 0 || index >= fm_num_all_algorithms) return nullptr;
  return fm_all_algorithms[index]()->name;
}

fm_status fm_algorithm_info(const char* alg_name, int* m, int* k, int* n, int* rank) {
  const FmAlgDesc* alg = find_alg(alg_name);
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

const char* fm_last_error(void) { return g_err.c_str(); }

fm_status fm_set_gemm_tile(int cfg) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (cfg < -1 || cfg >= FM_TILE_NUM) return set_err(FM_ERR_INVALID_ARG, "unknown tile config");
  g_tile_override = cfg;
  g_plans.clear();  // plans embed the tile choice
  return FM_OK;
}

void fm_profile_enable(int on) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_prof_on = on != 0;
}

fm_status fm_profile_read(fm_profile* out) {
  std::lock_guard<std::mutex> lock(g_mu);
  fm_status st = FM_OK;
  for (ProfRec& r : g_prof_pending) {
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(r.e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.e0, r.e1);
    if (e != hipSuccess && st == FM_OK) st = hip_err(e, "profile event");
    g_prof_acc.ms[r.stage] += ms;
    g_prof_acc.launches[r.stage] += 1;
    g_event_pool.push_back(r.e0);
    g_event_pool.push_back(r.e1);
  }
  g_prof_pending.clear();
  if (out) *out = g_prof_acc;
  memset(&g_prof_acc, 0, sizeof(g_prof_acc));
  return st;
}

void fm_release_caches(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  (void)hipDeviceSynchronize();
  g_plans.clear();
  for (auto& kv : g_arenas)
    if (kv.second.p) (void)hipFree(kv.second.p);
  g_arenas.clear();
}

}  // extern "C"
