// plan.cpp -- plan builder and executor.
//
// A sub-plan is the reference's recursion (algorithms/strassen.hpp:378-659)
// unrolled level by level.  Level l holds R^l nodes, each an (A, B, C)
// operand triple plus the node's multiplier.  Every node of a level has the
// same block shape (the floor split of linalg/linalg.hpp:179-184), so each
// level is one S and one T launch over all of its nodes; the base-case
// products of every node become one grouped GEMM launch; then, deepest level
// first, one W launch and one grouped launch of the peeling products.  This
// is the reference's BFS schedule (util/par_util.hpp:17-95) with every task
// of a level in one launch.
//
// Multipliers follow the reference exactly: the root carries alpha, a
// product whose S (T) column has a single coefficient c multiplies its
// node's multiplier by c (codegen/gen.py:614-621), base-case products and
// peeling products use the node's multiplier as alpha (linalg/linalg.hpp:266,
// 201-244).  beta reaches only the root's W stage and peeling; children use
// 0 (algorithms/strassen.hpp:459).  With num_steps = 0 the root is itself
// the base case, MatMul(A, B, C) with beta 0 (algorithms/strassen.hpp:385-388).
//
// Addresses are (role, offset) pairs resolved at launch time, so a plan is
// built and uploaded once per shape and replayed for any buffers.
#include "plan.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

namespace fmplan {

namespace {
thread_local std::string t_err;
const int64_t kLdAlign = 16;    // elements; 128 B for FP64
const int64_t kSlabAlign = 64;  // elements; keeps every buffer 256 B aligned

struct Node {
  FmOp a, b, c;
  double mu;
  // operand produced by a two-level stage and never stored: the workspace
  // offset of its view sub-blocks (-1 when the operand is stored)
  int64_t a_views = -1, b_views = -1;
};

FmOp op_make(int role, int64_t off, int64_t ld) {
  FmOp o;
  o.role = role;
  o.pad = 0;
  o.off = off;
  o.ld = ld;
  return o;
}

FmOp op_sub(const FmOp& o, int64_t i, int64_t j) { return op_make(o.role, o.off + i + j * o.ld, o.ld); }

FmGemmDesc gdesc(FmOp a, FmOp b, FmOp c, int64_t M, int64_t N, int64_t K, double alpha,
                 double beta) {
  FmGemmDesc d;
  memset(&d, 0, sizeof(d));
  d.a = a;
  d.b = b;
  d.c = c;
  d.M = (int32_t)M;
  d.N = (int32_t)N;
  d.K = (int32_t)K;
  d.alpha = alpha;
  d.beta = beta;
  return d;
}

// Two-level S/T stage over the nodes whose operands are `ops` (blocks
// rows x cols, sub-blocks rows2 x cols2).
Launch stage2_launch(LaunchKind kind, const FmAlgDesc* alg, std::vector<FmOp> ops, int64_t rows,
                     int64_t cols, int64_t rows2, int64_t cols2, int64_t out_off, int64_t out_ld,
                     int64_t out_slab, int64_t view_off, double x) {
  Launch L;
  L.kind = kind;
  L.alg = alg;
  L.stage = kind == LK_S2 ? FM_STAGE_S : kind == LK_T2 ? FM_STAGE_T : FM_STAGE_W;
  L.st.rows = (int32_t)rows;
  L.st.cols = (int32_t)cols;
  L.st.rows2 = (int32_t)rows2;
  L.st.cols2 = (int32_t)cols2;
  L.st.nodes = (int32_t)ops.size();
  L.st.out_off = out_off;
  L.st.out_ld = out_ld;
  L.st.out_slab = out_slab;
  L.st.view_off = view_off;
  L.st.x = x;
  L.st.beta = 0.0;
  L.vec_ok = 1;
  L.ops = std::move(ops);
  return L;
}

Launch stage_launch(fm_dtype dt, LaunchKind kind, const FmAlgDesc* alg, std::vector<FmOp> ops,
                    int64_t rows, int64_t cols, int64_t out_off, int64_t out_ld, int64_t out_slab,
                    double x, double beta) {
  Launch L;
  L.kind = kind;
  L.alg = alg;
  L.stage = kind == LK_S ? FM_STAGE_S : kind == LK_T ? FM_STAGE_T : FM_STAGE_W;
  L.st.rows = (int32_t)rows;
  L.st.cols = (int32_t)cols;
  L.st.nodes = (int32_t)ops.size();
  L.st.out_off = out_off;
  L.st.out_ld = out_ld;
  L.st.out_slab = out_slab;
  L.st.x = x;
  L.st.beta = beta;
  const int vmax = dt == FM_F64 ? 2 : 4;
  bool ok = rows % vmax == 0 && out_ld % vmax == 0 && out_slab % vmax == 0 && out_off % vmax == 0;
  for (const FmOp& o : ops) ok = ok && o.off % vmax == 0 && o.ld % vmax == 0;
  L.vec_ok = ok ? vmax : 1;
  L.ops = std::move(ops);
  return L;
}

// Peeling products of one node (linalg/linalg.hpp:201-244).
void peel_descs(const Node& nd, const Dims& D, int am, int ak, int an, double beta,
                std::vector<FmGemmDesc>& out) {
  const int64_t erA = D.m % am, ecA = D.k % ak, ecB = D.n % an;
  auto mk = [&](FmOp a, FmOp b, FmOp c, int64_t M, int64_t N, int64_t K, double bta) {
    if (M > 0 && N > 0) out.push_back(gdesc(a, b, c, M, N, K, nd.mu, bta));
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

Launch gemm_launch(std::vector<FmGemmDesc> descs, int cfg, int stage, fm_dtype dt) {
  int bm, bn;
  fm_tile_dims(cfg, &bm, &bn);
  int32_t t = 0;
  bool uniform = true;
  for (FmGemmDesc& d : descs) {
    d.tiles_m = (int32_t)((d.M + bm - 1) / bm);
    d.tiles_n = (int32_t)((d.N + bn - 1) / bn);
    d.tile_begin = t;
    t += d.tiles_m * d.tiles_n;
    uniform = uniform && d.M == descs[0].M && d.N == descs[0].N;
  }
  Launch L;
  L.kind = LK_GEMM;
  L.stage = stage;
  L.gm.ndesc = (int32_t)descs.size();
  L.gm.total_tiles = t;
  L.gm.uniform = uniform ? 1 : 0;
  // 16-byte loads need A/B column starts aligned: even (FP64) or
  // multiple-of-4 (FP32) offsets and leading dimensions.  The kernels check
  // each product's own; the base pointers are checked at launch.  The
  // unpredicated whole-tile kernel needs every product whole and aligned.
  const int vl = dt == FM_F64 ? 2 : 4;
  const int bk = fm_tile_bk(cfg);
  bool vec = true, full = true, even = true;
  for (const FmGemmDesc& d : descs) {
    vec = vec && d.a.off % vl == 0 && d.a.ld % vl == 0 && d.b.off % vl == 0 && d.b.ld % vl == 0;
    full = full && d.M % bm == 0 && d.N % bn == 0 && d.K % bk == 0 && d.K > 0;
    even = even && d.M % vl == 0 && d.K % vl == 0 && d.K > 0;
  }
  L.vec_ok = 1;
  L.gm.full = (full && vec) ? 1 : 0;
  L.gm.even = (even && vec) ? 1 : 0;
  L.cfg = cfg;
  L.gdescs = std::move(descs);
  return L;
}

// Leaf products on int8 digits: descriptors plus digit-plane workspace,
// appended at byte offset *ws_bytes (advanced).
fm_status ozaki_launch(int S, int tile, const std::vector<FmGemmDesc>& prods, size_t* ws_bytes,
                       Launch& L) {
  int32_t t = 0;
  int max_rows = 0, max_cols = 0;
  size_t off = (size_t)align_up((int64_t)*ws_bytes, 256);
  // kernel: FmTileCfg 64x64 forced selects the 16 x 16 x 64 one, 128x64
  // the 64 x 64-tile 32 x 32 x 32 one; otherwise 64 x 128 tiles of
  // 32 x 32 x 32 products (B digit planes padded to 128 columns)
  const int variant = tile == FM_TILE_64x64 ? 0 : tile == FM_TILE_128x64 ? 1 : 2;
  const int bn = variant == 2 ? 128 : 64;
  for (const FmGemmDesc& g : prods) {
    FmOzakiDesc d;
    memset(&d, 0, sizeof(d));
    d.a = g.a;
    d.b = g.b;
    d.c = g.c;
    d.M = g.M;
    d.N = g.N;
    d.K = g.K;
    d.Kp = (int32_t)align_up(g.K, 64);
    if (d.Kp > fm_ozaki_max_kp(S))
      return set_err(FM_ERR_SHAPE, "ozaki_int8: leaf K too large for exact int32 sums");
    d.rows_pad = (int32_t)align_up(g.M, 64);
    d.cols_pad = (int32_t)align_up(g.N, bn);
    d.tiles_m = d.rows_pad / 64;
    d.tiles_n = d.cols_pad / bn;
    d.tile_begin = t;
    t += d.tiles_m * d.tiles_n;
    d.alpha = g.alpha;
    d.beta = g.beta;
    d.a_slc = (int64_t)off;
    off += (size_t)S * d.rows_pad * d.Kp;
    d.b_slc = (int64_t)off;
    off += (size_t)S * d.cols_pad * d.Kp;
    d.a_exp = (int64_t)off;
    off += align_up(4 * (int64_t)d.rows_pad, 256);
    d.b_exp = (int64_t)off;
    off += align_up(4 * (int64_t)d.cols_pad, 256);
    max_rows = std::max(max_rows, d.rows_pad);
    max_cols = std::max(max_cols, d.cols_pad);
    L.odescs.push_back(d);
  }
  L.kind = LK_OZAKI;
  L.stage = FM_STAGE_LEAF;
  L.oz.ndesc = (int32_t)L.odescs.size();
  L.oz.total_tiles = t;
  L.oz.variant = variant;
  L.oz_slices = S;
  L.oz_rows = max_rows;
  L.oz_cols = max_cols;
  *ws_bytes = off;
  return FM_OK;
}

// Split-bf16 leaf products on slice planes: one preparation record per
// distinct operand (products that share an operand share its planes), the
// planes appended to the workspace at byte offset *ws_bytes (advanced).
Launch split_launch(std::vector<FmGemmDesc> prods, int cfg, size_t* ws_bytes) {
  int bm, bn;
  fm_tile_dims(cfg, &bm, &bn);
  Launch L;
  L.kind = LK_SPLIT;
  L.stage = FM_STAGE_LEAF;
  L.cfg = cfg;
  size_t off = (size_t)align_up((int64_t)*ws_bytes, 256);
  std::map<std::tuple<int, int64_t, int64_t, int, int, int>, int64_t> seen;
  int32_t prep_wgs = 0;
  auto planes = [&](const FmOp& src, int rows, int K, int pad_to, int kcontig) {
    if (src.role == FM_ROLE_PLANES) return src.off;  // written by the S/T stage
    const auto key = std::make_tuple(src.role, src.off, src.ld, rows, K, kcontig);
    auto it = seen.find(key);
    if (it != seen.end()) return it->second;
    FmSplitPrep r;
    memset(&r, 0, sizeof(r));
    r.src = src;
    r.rows = rows;
    r.K = K;
    r.rows_pad = (int32_t)align_up(rows, pad_to);
    r.Kp = (int32_t)align_up(K, 32);
    r.kcontig = kcontig;
    r.dst = (int64_t)off;
    off += (size_t)align_up((int64_t)r.rows_pad * r.Kp * 6, 256);
    r.wg_begin = prep_wgs;
    prep_wgs += fm_prep_wgs(r);
    L.spreps.push_back(r);
    seen.emplace(key, r.dst);
    return r.dst;
  };
  int32_t t = 0;
  bool uniform = true;
  for (const FmGemmDesc& g : prods) {
    FmSplitDesc d;
    memset(&d, 0, sizeof(d));
    d.c = g.c;
    d.M = g.M;
    d.N = g.N;
    d.K = g.K;
    d.Kp = (int32_t)align_up(g.K, 32);
    d.tiles_m = (g.M + bm - 1) / bm;
    d.tiles_n = (g.N + bn - 1) / bn;
    d.tile_begin = t;
    t += d.tiles_m * d.tiles_n;
    d.alpha = g.alpha;
    d.beta = g.beta;
    d.a_pl = planes(g.a, g.M, g.K, bm, 0);
    d.b_pl = planes(g.b, g.N, g.K, bn, 1);
    uniform = uniform && g.M == prods[0].M && g.N == prods[0].N;
    L.sdescs.push_back(d);
  }
  L.sp.ndesc = (int32_t)L.sdescs.size();
  L.sp.total_tiles = t;
  L.sp.uniform = uniform ? 1 : 0;
  L.sp.nprep = (int32_t)L.spreps.size();
  L.sp.prep_wgs = prep_wgs;
  L.sp.even_k = 1;
  for (const FmSplitDesc& d : L.sdescs)
    if ((d.Kp / 32) % 2 != 0) L.sp.even_k = 0;
  *ws_bytes = off;
  return L;
}

}  // namespace

fm_options g_defaults = {-1, 0, 0, 1, 1, 8};

fm_status check_options(const fm_options& o) {
  if (o.gemm_tile < -1 || o.gemm_tile >= FM_TILE_NUM) return set_err(FM_ERR_INVALID_ARG, "unknown tile config");
  if (o.dfs_levels < -1) return set_err(FM_ERR_INVALID_ARG, "dfs_levels must be >= -1");
  if (o.stage_fusion < 0 || o.stage_fusion > 2)
    return set_err(FM_ERR_INVALID_ARG, "stage_fusion must be 0, 1 or 2");
  if (o.ozaki_slices < 6 || o.ozaki_slices > 8)
    return set_err(FM_ERR_INVALID_ARG, "ozaki_slices must be 6, 7 or 8");
  return FM_OK;
}

fm_status set_err(fm_status s, const std::string& msg) {
  t_err = msg;
  return s;
}
fm_status hip_err(hipError_t e, const char* what) {
  return set_err(FM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
const std::string& last_error() { return t_err; }
void clear_error() { t_err.clear(); }

int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }
double round_scalar(double v, fm_dtype dt) { return dt == FM_F32 ? (double)(float)v : v; }
int elem_size(fm_dtype dt) { return dt == FM_F64 ? 8 : 4; }

// Tile of a grouped launch: smallest estimated time over the configurations,
// where a configuration's throughput per tile (eff, measured on MI355X at
// 8192^3) and its resident work groups per CU set how many rounds of tiles
// the launch takes.
namespace {
// A tile shape with the kernel used when every product of the launch is a
// whole number of tiles and K slices with 16-byte aligned columns
// (cfg_full), and the kernel for launches with edges (cfg_edge; -1: the
// shape is not offered, its edge kernel spills registers).
struct TileCand {
  int cfg_full, cfg_edge;
  double eff;           // throughput per tile area, relative
  int per_cu;           // resident work groups per CU
  bool spread = false;  // cfg_full also takes edge tiles when every product has whole K
                        // slices and aligned columns (its EVEN variant)
};
// Relative throughput: whole-tile rates on MI355X (8192^3 and the 676-leaf
// product of 512 x 2048 x 512), best = 1.  Work groups per CU: the lesser of
// what the register allocation (waves per SIMD x 4 / waves per group) and
// the 160 KB of LDS allow.
// (8192^3: FP64 72.6 TFLOP/s with the spread 128 x 128 tile, 69.0 with
// the same tile loading each slice at once; FP32 141.9 with the spread
// 128 x 256 tile, 136 with the 256 x 256 one)
const TileCand kF64Tiles[] = {
    {FM_TILE_PS_128x128_O2, FM_TILE_128x128_O2_U1, 1.00, 2},
    {FM_TILE_PS_128x256_W8, FM_TILE_128x256_W8_U1, 0.97, 1},
    {FM_TILE_128x256_W8, FM_TILE_128x256_W8_U1, 0.93, 1},
    {FM_TILE_256x128_W8, FM_TILE_256x128_W8_U1, 0.92, 1},
    {FM_TILE_128x128_W16, FM_TILE_128x128_W16, 0.86, 1},
    {FM_TILE_128x64_W8, FM_TILE_128x64_W8, 0.84, 2},
    {FM_TILE_PS_128x96_O2, FM_TILE_PI_128x96_O2, 0.87, 2, true},
    {FM_TILE_64x64, FM_TILE_64x64, 0.75, 3}};
const TileCand kF32Tiles[] = {
    {FM_TILE_PS_128x256_W8, FM_TILE_PP_128x256_W8, 1.00, 2},
    {FM_TILE_PS_128x128_O2, FM_TILE_128x128_W8, 0.985, 3},
    {FM_TILE_256x256_W16, -1, 0.96, 1},
    {FM_TILE_256x128_W8, FM_TILE_256x128_W8, 0.89, 1},
    {FM_TILE_128x128_W8, FM_TILE_128x128_W8, 0.76, 2},
    {FM_TILE_64x128, FM_TILE_64x128, 0.71, 3},
    {FM_TILE_128x64, FM_TILE_128x64, 0.68, 3},
    {FM_TILE_64x64, FM_TILE_64x64, 0.63, 4}};
// split-bf16: kernels on prepared slice planes (any shape: the planes are
// padded to whole tiles)
// (Strassen, two levels, 16384^3 FP32, i.e. 49 products of 4096^3, whole
// multiply: 251 TFLOP/s with B straight to registers, 232 and 229 with the
// LDS-staged 128 x 256 and 256 x 128 kernels)
const TileCand kSplitTiles[] = {{FM_TILE_SP_128x256_BD, FM_TILE_SP_128x256_BD, 1.00, 1},
                                {FM_TILE_SP_256x128_IL, FM_TILE_SP_256x128_IL, 0.92, 1},
                                {FM_TILE_SP_128x256_IL, FM_TILE_SP_128x256_IL, 0.92, 1},
                                {FM_TILE_SP_128x128, FM_TILE_SP_128x128, 0.88, 2},
                                {FM_TILE_SP_128x128_DB, FM_TILE_SP_128x128_DB, 0.83, 1}};
// edge kernels run slower than whole-tile ones (predicated loads).  The
// spread kernels' edge variant is offered for the 128 x 96 tile only: on
// 128 x 128 and 128 x 256 tiles it measured 6% slower per tile than the
// whole-tile kernel (classical 131072 x 2400 x 2400 FP64 66.8 TFLOP/s as
// one launch, 71.8 cut into a whole-tile core and a strip; the C3 FP32
// batch 119.2 against 124.4), while as the tall-skinny leaves' 88-column
// strip kernel it took the one-level multiply from 68.7 to 71.0
const double kEdgePenalty = 0.93;
const double kSpreadEdgePenalty = 0.93;

const TileCand* tile_cands(fm_dtype dt, fm_mode mode, int* n) {
  if (mode == FM_MODE_SPLIT_BF16) {
    *n = (int)(sizeof(kSplitTiles) / sizeof(kSplitTiles[0]));
    return kSplitTiles;
  }
  if (dt == FM_F64) {
    *n = (int)(sizeof(kF64Tiles) / sizeof(kF64Tiles[0]));
    return kF64Tiles;
  }
  *n = (int)(sizeof(kF32Tiles) / sizeof(kF32Tiles[0]));
  return kF32Tiles;
}

int64_t count_tiles(int cfg, const std::vector<FmGemmDesc>& prods, size_t from = 0) {
  int bm, bn;
  fm_tile_dims(cfg, &bm, &bn);
  int64_t tiles = 0;
  for (size_t i = from; i < prods.size(); ++i)
    tiles += ((prods[i].M + bm - 1) / bm) * ((prods[i].N + bn - 1) / bn);
  return tiles;
}

// Every product a whole number of tiles and K slices, columns 16-byte aligned.
bool whole_tiles(fm_dtype dt, int cfg, const std::vector<FmGemmDesc>& prods) {
  if (fm_tile_planes(cfg)) return true;
  int bm, bn;
  fm_tile_dims(cfg, &bm, &bn);
  const int bk = fm_tile_bk(cfg), vl = dt == FM_F64 ? 2 : 4;
  for (const FmGemmDesc& d : prods)
    if (d.M % bm || d.N % bn || d.K % bk || d.K == 0 || d.a.off % vl || d.a.ld % vl ||
        d.b.off % vl || d.b.ld % vl)
      return false;
  return true;
}

// Every product has M and K multiples of the vector width and 16-byte
// aligned columns (the spread kernels' EVEN launch).
bool even_slices(fm_dtype dt, int cfg, const std::vector<FmGemmDesc>& prods) {
  (void)cfg;
  const int vl = dt == FM_F64 ? 2 : 4;
  for (const FmGemmDesc& d : prods)
    if (d.M % vl || d.K % vl || d.K == 0 || d.a.off % vl || d.a.ld % vl || d.b.off % vl ||
        d.b.ld % vl)
      return false;
  return true;
}

// How a candidate takes a product list: the kernel (-1: not offered) and
// its form, 2 whole tiles, 1 the spread kernel's edge variant, 0 the
// predicated edge kernel.
int cand_kernel(fm_dtype dt, const TileCand& c, const std::vector<FmGemmDesc>& prods, int* form) {
  if (whole_tiles(dt, c.cfg_full, prods)) {
    *form = 2;
    return c.cfg_full;
  }
  if (c.spread && even_slices(dt, c.cfg_full, prods)) {
    *form = 1;
    return c.cfg_full;
  }
  *form = 0;
  return c.cfg_edge;
}

// Estimated time of one launch of `tiles` tiles: rounds of per_cu work
// groups on each of the 256 CUs.
double tile_time(const TileCand& c, int64_t tiles, int form) {
  int bm, bn;
  fm_tile_dims(c.cfg_full, &bm, &bn);
  const int64_t slots = 256LL * c.per_cu;
  const double rounds = (double)((tiles + slots - 1) / slots);
  const double pen = form == 2 ? 1.0 : form == 1 ? kSpreadEdgePenalty : kEdgePenalty;
  return rounds * bm * bn * c.per_cu / (c.eff * pen);
}

const TileCand* find_cand(fm_dtype dt, fm_mode mode, int cfg) {
  int n;
  const TileCand* c = tile_cands(dt, mode, &n);
  for (int i = 0; i < n; ++i)
    if (c[i].cfg_full == cfg || c[i].cfg_edge == cfg) return &c[i];
  return nullptr;
}

// Best candidate for a product list: its kernel and estimated time.
int best_tile(fm_dtype dt, fm_mode mode, const std::vector<FmGemmDesc>& prods, double* t_out) {
  int nc;
  const TileCand* cands = tile_cands(dt, mode, &nc);
  int best = FM_TILE_64x64;
  double best_t = 1e300;
  for (int i = 0; i < nc; ++i) {
    int form;
    const int cfg = cand_kernel(dt, cands[i], prods, &form);
    if (cfg < 0) continue;
    const double t = tile_time(cands[i], count_tiles(cfg, prods), form);
    if (t < best_t * 0.999) {
      best_t = t;
      best = cfg;
    }
  }
  if (t_out) *t_out = best_t;
  return best;
}
}  // namespace

namespace {
// Whole-tile carve-out: a launch whose products are not all whole tiles runs
// every tile on the predicated edge kernel.  Cutting each ragged product
// into its whole-tile core (the unpredicated kernel) and the row / column
// strips left over (a second launch, its own tile) is worth it when the
// strips are a small share; every output element is still computed by the
// same K-ordered MFMA sequence.  Returns false (prods unchanged) when it
// does not pay or does not apply.
FmGemmDesc gsub(const FmGemmDesc& d, int64_t i0, int64_t j0, int64_t m, int64_t n) {
  return gdesc(op_sub(d.a, i0, 0), op_sub(d.b, 0, j0), op_sub(d.c, i0, j0), m, n, d.K, d.alpha,
               d.beta);
}

bool carve_full(const Plan& P, int cfg, std::vector<FmGemmDesc>& prods, std::vector<FmGemmDesc>& strips,
                std::vector<FmGemmDesc>& odd, int* cfg_core, int* cfg_strips, int* cfg_odd) {
  if (P.tile >= 0 || P.mode != FM_MODE_PLAIN) return false;
  const TileCand* c = find_cand(P.dtype, P.mode, cfg);
  if (!c || whole_tiles(P.dtype, c->cfg_full, prods)) return false;
  int bm, bn;
  fm_tile_dims(c->cfg_full, &bm, &bn);
  const int bk = fm_tile_bk(c->cfg_full), vl = P.dtype == FM_F64 ? 2 : 4;
  std::vector<FmGemmDesc> core, st, od;
  for (const FmGemmDesc& d : prods) {
    if (d.K % bk || d.K == 0 || d.a.off % vl || d.a.ld % vl || d.b.off % vl || d.b.ld % vl) {
      // a product the whole-tile kernel cannot take (ragged K, unaligned
      // columns) runs whole in a launch of its own, with its own tile, so
      // that it neither sends every other product of a grouped launch to
      // the edge kernel nor takes the thin strips' small tile
      od.push_back(d);
      continue;
    }
    const int64_t Mf = d.M / bm * bm, Nf = d.N / bn * bn;
    if (Mf > 0 && Nf > 0) core.push_back(gsub(d, 0, 0, Mf, Nf));
    if (d.N > Nf) st.push_back(gsub(d, 0, Nf, d.M, d.N - Nf));          // right strip, all rows
    if (d.M > Mf && Nf > 0) st.push_back(gsub(d, Mf, 0, d.M - Mf, Nf));  // bottom strip
  }
  if (core.empty() || (st.empty() && od.empty())) return false;
  double t_strips = 0, t_odd = 0;
  const int cs = st.empty() ? cfg : best_tile(P.dtype, P.mode, st, &t_strips);
  const int co = od.empty() ? cfg : best_tile(P.dtype, P.mode, od, &t_odd);
  int form_all;
  (void)cand_kernel(P.dtype, *c, prods, &form_all);
  const double t_core = tile_time(*c, count_tiles(c->cfg_full, core), 2);
  const double t_all = tile_time(*c, count_tiles(cfg, prods), form_all);
  if (t_core + t_strips + t_odd >= 0.97 * t_all) return false;
  prods = std::move(core);
  strips = std::move(st);
  odd = std::move(od);
  *cfg_core = c->cfg_full;
  *cfg_strips = cs;
  *cfg_odd = co;
  return true;
}
}  // namespace

int pick_tile(fm_dtype dt, fm_mode mode, int tile, const std::vector<FmGemmDesc>& prods) {
  if (tile >= 0) return tile;
  return best_tile(dt, mode, prods, nullptr);
}

// The GPU form of the reference's HYBRID remainder rule (util/par_util.hpp:
// 65-95: leaf tasks that do not fill the last round of threads run one at a
// time on all threads): when the products' tiles leave the last round of
// work groups part-empty, the products of that round move to a second
// launch with the tile that finishes them soonest.  Products stay whole, so
// every output element is computed by the same instructions either way.
void split_remainder(const Plan& P, int cfg, std::vector<FmGemmDesc>& prods,
                     std::vector<FmGemmDesc>& rest, int* cfg_rest, int* cfg_head) {
  if (P.tile >= 0 || P.mode == FM_MODE_OZAKI_INT8) return;
  const TileCand* c = find_cand(P.dtype, P.mode, cfg);
  if (!c) return;
  const int64_t slots = 256LL * c->per_cu;
  const int64_t total = count_tiles(cfg, prods);
  const int64_t full = total / slots * slots;
  if (full == 0 || full == total) return;
  // the longest prefix of products whose tiles fill whole rounds
  int bm, bn;
  fm_tile_dims(cfg, &bm, &bn);
  int64_t acc = 0;
  size_t cut = 0;
  while (cut < prods.size()) {
    const int64_t t = ((prods[cut].M + bm - 1) / bm) * ((prods[cut].N + bn - 1) / bn);
    if (acc + t > full) break;
    acc += t;
    ++cut;
  }
  if (acc != full || cut == prods.size()) return;
  std::vector<FmGemmDesc> tail(prods.begin() + (std::ptrdiff_t)cut, prods.end());
  std::vector<FmGemmDesc> head(prods.begin(), prods.begin() + (std::ptrdiff_t)cut);
  double t_tail = 0;
  const int cr = best_tile(P.dtype, P.mode, tail, &t_tail);
  int form_one, form_head;
  (void)cand_kernel(P.dtype, *c, prods, &form_one);
  const int ch = cand_kernel(P.dtype, *c, head, &form_head);
  const double t_one = tile_time(*c, total, form_one);
  const double t_two = tile_time(*c, full, form_head) + t_tail;
  if (t_two >= t_one * 0.99) return;
  prods.resize(cut);
  rest = std::move(tail);
  *cfg_rest = cr;
  // the full rounds may be whole tiles even where the remainder was not
  if (ch >= 0) *cfg_head = ch;
}

// Depth-first top level (the reference's DFS order, algorithms/strassen.hpp:
// 404-463): the level-0 S/T stages, then each child's whole sub-plan in turn
// (its own stages, leaf launch and W stages), every child reusing one
// workspace region, then the level-0 W stage and peeling.  Arithmetic is
// that of the breadth-first plan; only the launch order and the workspace
// differ.
fm_status build_dfs(fm_dtype dt, const Problem& p, const std::vector<Dims>& dims, int64_t ws_base,
                    SubPlan& out) {
  const FmAlgDesc* alg = p.alg;
  const int R = alg->R, NS = alg->ns, NT = alg->nt;
  const FmOp A0 = op_make(p.role_a, p.off_a, p.lda), B0 = op_make(p.role_b, p.off_b, p.ldb),
             C0 = op_make(p.role_c, p.off_c, p.ldc);
  const Dims c = dims[1];
  const int64_t s_ld = align_up(c.m, kLdAlign), s_slab = align_up(s_ld * c.k, kSlabAlign);
  const int64_t t_ld = align_up(c.k, kLdAlign), t_slab = align_up(t_ld * c.n, kSlabAlign);
  const int64_t m_ld = align_up(c.m, kLdAlign), m_slab = align_up(m_ld * c.n, kSlabAlign);
  int64_t ws = ws_base;
  const int64_t s_off = ws;
  ws += NS * s_slab;
  const int64_t t_off = ws;
  ws += NT * t_slab;
  const int64_t m_off = ws;
  ws += R * m_slab;
  if (NS > 0)
    out.pre.push_back(stage_launch(dt, LK_S, alg, {A0}, c.m, c.k, s_off, s_ld, s_slab, p.x, 0.0));
  if (NT > 0)
    out.pre.push_back(stage_launch(dt, LK_T, alg, {B0}, c.k, c.n, t_off, t_ld, t_slab, p.x, 0.0));
  const int64_t child_base = ws;
  const size_t es = (size_t)elem_size(dt);
  size_t end_bytes = (size_t)ws * es;
  const double mu0 = round_scalar(p.alpha, dt);
  for (int r = 0; r < R; ++r) {
    const int ai = alg->a_index[r], bi = alg->b_index[r];
    const FmOp a = alg->a_kind[r] ? op_make(FM_ROLE_WS, s_off + ai * s_slab, s_ld)
                                  : op_sub(A0, (ai / alg->k) * c.m, (ai % alg->k) * c.k);
    const FmOp b = alg->b_kind[r] ? op_make(FM_ROLE_WS, t_off + bi * t_slab, t_ld)
                                  : op_sub(B0, (bi / alg->n) * c.k, (bi % alg->n) * c.n);
    Problem q = p;
    q.steps = p.steps - 1;
    q.dfs = p.dfs - 1;
    q.fuse_planes = false;
    q.m = c.m;
    q.k = c.k;
    q.n = c.n;
    q.role_a = a.role;
    q.off_a = a.off;
    q.lda = a.ld;
    q.role_b = b.role;
    q.off_b = b.off;
    q.ldb = b.ld;
    q.role_c = FM_ROLE_WS;
    q.off_c = m_off + r * m_slab;
    q.ldc = m_ld;
    const double mu = round_scalar(mu0 * round_scalar(alg->a_coef(r, p.x), dt), dt);
    q.alpha = round_scalar(mu * round_scalar(alg->b_coef(r, p.x), dt), dt);
    q.beta = 0.0;
    std::vector<SubPlan> child(1);
    fm_status st = build_subplan(dt, q, child_base, child[0]);
    if (st != FM_OK) return st;
    Plan seq;
    seq.dtype = dt;
    seq.mode = p.mode;
    seq.slices = p.slices;
    seq.tile = p.tile;
    seq.split_rem = p.split_rem;
    st = assemble(seq, child, child_base + child[0].ws_elems);
    if (st != FM_OK) return st;
    for (Launch& L : seq.launches) out.pre.push_back(std::move(L));
    end_bytes = std::max(end_bytes, seq.ws_bytes);
  }
  const double bl = round_scalar(p.beta, dt);
  out.w.push_back(stage_launch(dt, LK_W, alg, {C0}, c.m, c.n, m_off, m_ld, m_slab, p.x, bl));
  out.peels.resize(1);
  peel_descs(Node{A0, B0, C0, mu0}, dims[0], alg->m, alg->k, alg->n, bl, out.peels[0]);
  out.depth = 1;
  out.leaf = dims.back();
  out.ws_elems = (int64_t)((end_bytes + es - 1) / es) - ws_base;
  return FM_OK;
}

fm_status build_subplan(fm_dtype dt, const Problem& p, int64_t ws_base, SubPlan& out) {
  out = SubPlan();
  const FmOp A0 = op_make(p.role_a, p.off_a, p.lda), B0 = op_make(p.role_b, p.off_b, p.ldb),
             C0 = op_make(p.role_c, p.off_c, p.ldc);
  if (!p.alg) {  // classical product, beta honoured (MatMul, linalg/linalg.hpp:261-269)
    out.leaves.push_back(gdesc(A0, B0, C0, p.m, p.n, p.k, round_scalar(p.alpha, dt),
                               round_scalar(p.beta, dt)));
    out.leaf = {p.m, p.k, p.n};
    return FM_OK;
  }
  const FmAlgDesc* alg = p.alg;
  const int L = p.steps;
  std::vector<Dims> dims(L + 1);
  dims[0] = {p.m, p.k, p.n};
  for (int l = 0; l < L; ++l) {
    dims[l + 1] = {dims[l].m / alg->m, dims[l].k / alg->k, dims[l].n / alg->n};
    if (dims[l + 1].m <= 0 || dims[l + 1].k <= 0 || dims[l + 1].n <= 0)
      return set_err(FM_ERR_SHAPE, "num_steps too deep for this shape: a block dimension "
                                   "reaches 0 at level " + std::to_string(l + 1));
  }
  if (p.dfs > 0 && L > 0) return build_dfs(dt, p, dims, ws_base, out);
  std::vector<std::vector<Node>> lv(L + 1);
  lv[0].push_back(Node{A0, B0, C0, round_scalar(p.alpha, dt)});
  std::vector<int64_t> m_off(L), m_ld(L), m_slab(L);
  int64_t ws = ws_base;
  const int R = alg->R, NS = alg->ns, NT = alg->nt;
  // Two-level stages: s2[l] (t2[l]) merges the S (T) stage of level l into
  // that of level l + 1.  Pairs are taken from the deepest level up (deeper
  // levels move more data); a pair needs level l + 1 free of peeling, which
  // would read the level-(l+1) operands that a two-level stage never stores.
  // w2[l]: the W stages of levels l + 1 and l run as one launch, and the
  // level-(l+1) products are never stored.
  std::vector<char> s2(L, 0), t2(L, 0), w2(L, 0);
  if (p.fuse) {
    const bool fs = (dt == FM_F64 ? alg->s2_d : alg->s2_f) != nullptr;
    const bool ft = (dt == FM_F64 ? alg->t2_d : alg->t2_f) != nullptr;
    const bool fw = (dt == FM_F64 ? alg->w2_d : alg->w2_f) != nullptr;
    for (int l = L - 2; l >= 0;) {
      const Dims c = dims[l + 1];
      if (c.m % alg->m == 0 && c.k % alg->k == 0 && c.n % alg->n == 0 && (fs || ft || fw)) {
        s2[l] = fs;
        t2[l] = ft;
        w2[l] = fw;
        l -= 2;
      } else {
        l -= 1;
      }
    }
  }
  // Split-bf16 leaves on slice planes: the two-level S (T) stage of the last
  // two levels writes the leaf operands' planes itself, and the preparation
  // pass covers only leaf operands that are sub-blocks of A (B).  Taken when
  // every split kernel's tiles divide the leaves (rows multiples of 256) and
  // K is whole 32-deep slices, so the planes need no padding, and when the
  // leaves' preparation grid would fit one launch (else they split on the
  // fly, on the FP32 operands).
  bool s_pl = false, t_pl = false;
  if (p.fuse_planes && dt == FM_F32 && L >= 2) {
    const Dims d = dims[L];
    double leaves = 1.0;
    for (int l = 0; l < L; ++l) leaves *= R;
    const double wgs = leaves * (double)((d.m + d.n + 126) / 64) * (double)((d.k + 127) / 128);
    const bool ok = d.m % 256 == 0 && d.n % 256 == 0 && d.k % 32 == 0 && wgs <= 2147483647.0;
    s_pl = ok && s2[L - 2] && alg->s2p_f != nullptr;
    t_pl = ok && t2[L - 2] && alg->t2p_f != nullptr;
  }
  const FmOp virt = op_make(-1, 0, 0);
  for (int l = 0; l < L; ++l) {
    const Dims c = dims[l + 1];
    const int64_t Q = (int64_t)lv[l].size();
    const int64_t s_ld = align_up(c.m, kLdAlign), s_slab = align_up(s_ld * c.k, kSlabAlign);
    const int64_t t_ld = align_up(c.k, kLdAlign), t_slab = align_up(t_ld * c.n, kSlabAlign);
    // the last level's S (T) outputs and views as slice planes: bytes per
    // operand (c.m x c.k, 6 bytes an element), in FP32 workspace elements
    const bool spl = s_pl && l == L - 1, tpl = t_pl && l == L - 1;
    const int64_t s_pb = align_up(c.m * c.k * 6, 256), t_pb = align_up(c.k * c.n * 6, 256);
    const int64_t s_step = spl ? s_pb / 4 : s_slab, t_step = tpl ? t_pb / 4 : t_slab;
    m_ld[l] = align_up(c.m, kLdAlign);
    m_slab[l] = align_up(m_ld[l] * c.n, kSlabAlign);
    const bool s_first = s2[l], s_second = l > 0 && s2[l - 1];
    const bool t_first = t2[l], t_second = l > 0 && t2[l - 1];
    const int64_t s_off = ws;
    if (!s_first) ws += Q * NS * s_step;
    const int64_t t_off = ws;
    if (!t_first) ws += Q * NT * t_step;
    const int64_t sv_off = ws;
    if (s_second) ws += Q * alg->a_nviews * s_step;
    const int64_t tv_off = ws;
    if (t_second) ws += Q * alg->b_nviews * t_step;
    m_off[l] = ws;
    if (!w2[l]) ws += Q * R * m_slab[l];
    lv[l + 1].reserve(Q * R);
    for (int64_t q = 0; q < Q; ++q) {
      Node& nd = lv[l][q];
      if (s_second && nd.a.role < 0) nd.a_views = sv_off + q * alg->a_nviews * s_step;
      if (t_second && nd.b.role < 0) nd.b_views = tv_off + q * alg->b_nviews * t_step;
      // a workspace operand, or (last level, planes) the byte offset of its planes
      auto s_op = [&](int64_t off) { return spl ? op_make(FM_ROLE_PLANES, off * 4, 0) : op_make(FM_ROLE_WS, off, s_ld); };
      auto t_op = [&](int64_t off) { return tpl ? op_make(FM_ROLE_PLANES, off * 4, 0) : op_make(FM_ROLE_WS, off, t_ld); };
      for (int r = 0; r < R; ++r) {
        Node ch;
        const int ai = alg->a_index[r], bi = alg->b_index[r];
        if (alg->a_kind[r])
          ch.a = s_first ? virt : s_op(s_off + (q * NS + ai) * s_step);
        else if (nd.a_views >= 0)
          ch.a = s_op(nd.a_views + alg->a_view_slot[ai] * s_step);
        else
          ch.a = op_sub(nd.a, (ai / alg->k) * c.m, (ai % alg->k) * c.k);
        if (alg->b_kind[r])
          ch.b = t_first ? virt : t_op(t_off + (q * NT + bi) * t_step);
        else if (nd.b_views >= 0)
          ch.b = t_op(nd.b_views + alg->b_view_slot[bi] * t_step);
        else
          ch.b = op_sub(nd.b, (bi / alg->n) * c.k, (bi % alg->n) * c.n);
        ch.c = w2[l] ? virt : op_make(FM_ROLE_WS, m_off[l] + (q * R + r) * m_slab[l], m_ld[l]);
        double mu = round_scalar(nd.mu * round_scalar(alg->a_coef(r, p.x), dt), dt);
        ch.mu = round_scalar(mu * round_scalar(alg->b_coef(r, p.x), dt), dt);
        lv[l + 1].push_back(ch);
      }
    }
    if (NS > 0 && s_second) {
      const Dims pc = dims[l];
      std::vector<FmOp> ops;
      for (const Node& nd : lv[l - 1]) ops.push_back(nd.a);
      if (spl) {
        Launch S = stage2_launch(LK_S2, alg, std::move(ops), pc.m, pc.k, c.m, c.k, s_off * 4, 0,
                                 s_pb, sv_off * 4, p.x);
        S.planes = 1;
        out.pre.push_back(std::move(S));
      } else {
        out.pre.push_back(stage2_launch(LK_S2, alg, std::move(ops), pc.m, pc.k, c.m, c.k, s_off,
                                        s_ld, s_slab, sv_off, p.x));
      }
    } else if (NS > 0 && !s_first) {
      std::vector<FmOp> ops;
      for (const Node& nd : lv[l]) ops.push_back(nd.a);
      out.pre.push_back(stage_launch(dt, LK_S, alg, std::move(ops), c.m, c.k, s_off, s_ld, s_slab,
                                     p.x, 0.0));
    }
    if (NT > 0 && t_second) {
      const Dims pc = dims[l];
      std::vector<FmOp> ops;
      for (const Node& nd : lv[l - 1]) ops.push_back(nd.b);
      if (tpl) {
        Launch T = stage2_launch(LK_T2, alg, std::move(ops), pc.k, pc.n, c.k, c.n, t_off * 4, 0,
                                 t_pb, tv_off * 4, p.x);
        T.planes = 1;
        out.pre.push_back(std::move(T));
      } else {
        out.pre.push_back(stage2_launch(LK_T2, alg, std::move(ops), pc.k, pc.n, c.k, c.n, t_off,
                                        t_ld, t_slab, tv_off, p.x));
      }
    } else if (NT > 0 && !t_first) {
      std::vector<FmOp> ops;
      for (const Node& nd : lv[l]) ops.push_back(nd.b);
      out.pre.push_back(stage_launch(dt, LK_T, alg, std::move(ops), c.k, c.n, t_off, t_ld, t_slab,
                                     p.x, 0.0));
    }
  }
  const Dims d = dims[L];
  for (const Node& nd : lv[L]) out.leaves.push_back(gdesc(nd.a, nd.b, nd.c, d.m, d.n, d.k, nd.mu, 0.0));
  out.leaf = d;
  out.w.resize(L);
  out.peels.resize(L);
  for (int l = L - 1; l >= 0; --l) {
    const Dims c = dims[l + 1];
    const double bl = l == 0 ? round_scalar(p.beta, dt) : 0.0;
    std::vector<FmOp> ops;
    for (const Node& nd : lv[l]) ops.push_back(nd.c);
    if (w2[l]) {
      // levels l + 1 and l in one launch; the level-(l+1) launch is dropped
      const Dims c2 = dims[l + 2];
      out.w[l + 1] = Launch();
      out.w[l + 1].kind = LK_NONE;
      Launch W = stage2_launch(LK_W2, alg, std::move(ops), c.m, c.n, c2.m, c2.n, m_off[l + 1],
                               m_ld[l + 1], m_slab[l + 1], 0, p.x);
      W.st.beta = bl;
      out.w[l] = std::move(W);
    } else {
      out.w[l] = stage_launch(dt, LK_W, alg, std::move(ops), c.m, c.n, m_off[l], m_ld[l],
                              m_slab[l], p.x, bl);
    }
    for (const Node& nd : lv[l]) peel_descs(nd, dims[l], alg->m, alg->k, alg->n, bl, out.peels[l]);
  }
  out.depth = L;
  out.ws_elems = ws - ws_base;
  return FM_OK;
}

namespace {
// The stage kernels walk a block's elements with 32-bit counters in a
// grid-stride loop; blocks of 2^31 elements or more are refused (only
// reachable with matrices near the device's memory size).
fm_status check_stage_size(const Launch& L) {
  if (L.kind == LK_GEMM || L.kind == LK_OZAKI || L.kind == LK_SPLIT || L.kind == LK_NONE) return FM_OK;
  const bool two = L.kind == LK_S2 || L.kind == LK_T2 || L.kind == LK_W2;
  const int64_t elems = two ? (int64_t)L.st.rows2 * L.st.cols2 : (int64_t)L.st.rows * L.st.cols;
  if (elems >= (int64_t(1) << 31))
    return set_err(FM_ERR_SHAPE, "a recursion block of " + std::to_string(elems) +
                                     " elements exceeds the stage kernels' 2^31 limit; use more steps");
  return FM_OK;
}
}  // namespace

namespace {
// Output slabs per node of a stage launch's layout (see FmNodeView).
void stage_slabs(const Launch& L, int* nout, int* nview) {
  const FmAlgDesc* g = L.alg;
  *nview = 0;
  switch (L.kind) {
    case LK_S: *nout = g->ns; break;
    case LK_T: *nout = g->nt; break;
    case LK_W: *nout = g->R; break;
    case LK_S2: *nout = g->R * g->ns; *nview = g->R * g->a_nviews; break;
    case LK_T2: *nout = g->R * g->nt; *nview = g->R * g->b_nviews; break;
    case LK_W2: *nout = g->R * g->R; break;
    default: *nout = 0; break;
  }
}

// One launch for stage launches of the same kind and algorithm from several
// problems (their block shapes may differ): per-node descriptors, the
// largest shape for the grid, 16-byte accesses only if every node allows
// them.  The kernels compute every node exactly as its own launch would.
Launch group_stages(std::vector<Launch>& ls, fm_dtype dt) {
  Launch G;
  const Launch& f = ls[0];
  G.kind = f.kind;
  G.alg = f.alg;
  G.stage = f.stage;
  G.st = f.st;
  G.st.rows = G.st.cols = G.st.rows2 = G.st.cols2 = 0;
  int vec = dt == FM_F64 ? 2 : 4;
  for (const Launch& L : ls) {
    int nout, nview;
    stage_slabs(L, &nout, &nview);
    vec = std::min(vec, L.vec_ok);
    G.st.rows = std::max(G.st.rows, L.st.rows);
    G.st.cols = std::max(G.st.cols, L.st.cols);
    G.st.rows2 = std::max(G.st.rows2, L.st.rows2);
    G.st.cols2 = std::max(G.st.cols2, L.st.cols2);
    for (size_t q = 0; q < L.ops.size(); ++q) {
      FmStageNode n;
      memset(&n, 0, sizeof(n));
      n.op = L.ops[q];
      n.rows = L.st.rows;
      n.cols = L.st.cols;
      n.rows2 = L.st.rows2;
      n.cols2 = L.st.cols2;
      n.out_off = L.st.out_off + (int64_t)q * nout * L.st.out_slab;
      n.out_ld = L.st.out_ld;
      n.out_slab = L.st.out_slab;
      n.view_off = L.st.view_off + (int64_t)q * nview * L.st.out_slab;
      n.x = L.st.x;
      n.beta = L.st.beta;
      G.snodes.push_back(n);
    }
  }
  G.st.nodes = (int32_t)G.snodes.size();
  G.vec_ok = vec;
  return G;
}

// Merge the stage launches of several problems: launches at the same
// position of their problems' sequences with the same kind, algorithm and
// (for W) beta class become one grouped launch; positions keep their order,
// so every problem's own stage order is kept.
void group_sequence(std::vector<std::vector<Launch>*> seqs, std::vector<Launch>& out, fm_dtype dt) {
  size_t depth = 0;
  for (auto* sq : seqs) depth = std::max(depth, sq->size());
  for (size_t i = 0; i < depth; ++i) {
    std::vector<std::pair<std::tuple<int, const FmAlgDesc*, int>, std::vector<Launch>>> groups;
    for (auto* sq : seqs) {
      if (i >= sq->size() || (*sq)[i].kind == LK_NONE) continue;
      Launch& L = (*sq)[i];
      const bool stage = L.kind == LK_S || L.kind == LK_T || L.kind == LK_W || L.kind == LK_S2 ||
                         L.kind == LK_T2 || L.kind == LK_W2;
      if (!stage) {  // a depth-first child's leaf launch: kept as it is
        out.push_back(std::move(L));
        continue;
      }
      const auto key = std::make_tuple((int)L.kind, L.alg, L.st.beta != 0.0 ? 1 : 0);
      auto it = std::find_if(groups.begin(), groups.end(), [&](const auto& g) { return g.first == key; });
      if (it == groups.end()) {
        groups.emplace_back(key, std::vector<Launch>());
        it = groups.end() - 1;
      }
      it->second.push_back(std::move(L));
    }
    for (auto& g : groups) {
      if (g.second.size() == 1)
        out.push_back(std::move(g.second[0]));
      else
        out.push_back(group_stages(g.second, dt));
    }
  }
}
}  // namespace

fm_status assemble(Plan& P, std::vector<SubPlan>& subs, int64_t ws_elems, int persistent_grid) {
  for (SubPlan& s : subs) {
    for (const Launch& L : s.pre) {
      const fm_status st = check_stage_size(L);
      if (st != FM_OK) return st;
    }
    for (const Launch& L : s.w) {
      const fm_status st = check_stage_size(L);
      if (st != FM_OK) return st;
    }
  }
  if (subs.size() > 1) {
    // several problems (a batch): one launch per stage position and kind
    std::vector<std::vector<Launch>*> seqs;
    for (SubPlan& s : subs) seqs.push_back(&s.pre);
    group_sequence(seqs, P.launches, P.dtype);
  } else {
    for (SubPlan& s : subs)
      for (Launch& L : s.pre) P.launches.push_back(std::move(L));
  }
  std::vector<FmGemmDesc> leaves;
  for (SubPlan& s : subs) leaves.insert(leaves.end(), s.leaves.begin(), s.leaves.end());
  size_t wsb = (size_t)ws_elems * elem_size(P.dtype);
  if (!leaves.empty()) {
    if (P.mode == FM_MODE_OZAKI_INT8) {
      Launch L;
      fm_status st = ozaki_launch(P.slices, P.tile, leaves, &wsb, L);
      if (st != FM_OK) return st;
      P.launches.push_back(std::move(L));
    } else {
      const int cfg = pick_tile(P.dtype, P.mode, P.tile, leaves);
      if (fm_tile_planes(cfg) && P.mode != FM_MODE_SPLIT_BF16)
        return set_err(FM_ERR_INVALID_ARG, "tile config " + std::to_string(cfg) + " is for split_bf16 leaves");
      // split_bf16 leaves off the slice planes run the in-loop kernel, which
      // has the 128 x 128, 64 x 64, 128 x 64, 64 x 128 and 128 x 128 (8-wave) tiles
      if (P.mode == FM_MODE_SPLIT_BF16 && !fm_tile_planes(cfg) && cfg != FM_TILE_128x128 &&
          cfg != FM_TILE_64x64 && cfg != FM_TILE_128x64 && cfg != FM_TILE_64x128 &&
          cfg != FM_TILE_128x128_W8)
        return set_err(FM_ERR_INVALID_ARG, "tile config " + std::to_string(cfg) +
                                               " has no split_bf16 kernel");
      std::vector<FmGemmDesc> rest, strips, odd;
      int cfg_rest = cfg, cfg_head = cfg, cfg_main = cfg, cfg_strips = cfg, cfg_odd = cfg;
      if (persistent_grid == 0 && P.split_rem &&
          carve_full(P, cfg, leaves, strips, odd, &cfg_main, &cfg_strips, &cfg_odd))
        cfg_rest = cfg_head = cfg_main;
      if (persistent_grid == 0 && P.split_rem)
        split_remainder(P, cfg_main, leaves, rest, &cfg_rest, &cfg_head);
      auto leaf = [&](std::vector<FmGemmDesc> prods, int c) {
        // the preparation launch is one flat grid of at most 2^31 - 1 work
        // groups (64 rows x 4 K slices each, per operand; an upper bound
        // here): beyond that the leaves split on the fly instead
        int64_t wgs = 0;
        for (const FmGemmDesc& d : prods)
          wgs += ((int64_t)d.M + d.N + 126) / 64 * ((d.K + 127) / 128);
        if (fm_tile_planes(c) && wgs > 2147483647LL) c = FM_TILE_128x128;
        if (fm_tile_planes(c)) return split_launch(std::move(prods), c, &wsb);
        return gemm_launch(std::move(prods), c, FM_STAGE_LEAF, P.dtype);
      };
      P.launches.push_back(leaf(std::move(leaves), cfg_head));
      P.launches.back().gm.grid = persistent_grid;
      P.launches.back().sp.grid = persistent_grid;
      if (!rest.empty()) P.launches.push_back(leaf(std::move(rest), cfg_rest));
      if (!strips.empty()) P.launches.push_back(leaf(std::move(strips), cfg_strips));
      if (!odd.empty()) P.launches.push_back(leaf(std::move(odd), cfg_odd));
    }
  }
  int maxd = 0;
  for (const SubPlan& s : subs) maxd = std::max(maxd, s.depth);
  for (int l = maxd - 1; l >= 0; --l) {
    std::vector<FmGemmDesc> peel;
    std::vector<Launch> wl;
    for (SubPlan& s : subs) {
      if (s.depth <= l) continue;
      if (s.w[l].kind != LK_NONE) wl.push_back(std::move(s.w[l]));
      peel.insert(peel.end(), s.peels[l].begin(), s.peels[l].end());
    }
    if (wl.size() > 1) {
      // this level's W stages of every problem: grouped by kind, algorithm, beta class
      std::vector<std::vector<Launch>> one(wl.size());
      std::vector<std::vector<Launch>*> seqs;
      for (size_t i = 0; i < wl.size(); ++i) {
        one[i].push_back(std::move(wl[i]));
        seqs.push_back(&one[i]);
      }
      group_sequence(seqs, P.launches, P.dtype);
    } else {
      for (Launch& L : wl) P.launches.push_back(std::move(L));
    }
    // peeling products stay on the plain kernel in OZAKI mode (FP64) and
    // take the split kernel in SPLIT_BF16 mode
    if (!peel.empty()) P.launches.push_back(gemm_launch(std::move(peel), FM_TILE_64x64, FM_STAGE_PEEL, P.dtype));
  }
  P.ws_bytes = wsb;
  return FM_OK;
}

Plan::~Plan() {
  if (dbuf) (void)hipFree(dbuf);
}

// Descriptors and the role table, packed into one host buffer; each
// launch records its offsets in it.
void Plan::pack(std::vector<char>& buf) {
  buf.clear();
  auto put = [&](const void* src, size_t bytes) {
    const size_t off = (size_t)align_up((int64_t)buf.size(), 256);
    buf.resize(off + bytes);
    if (bytes) memcpy(buf.data() + off, src, bytes);
    return off;
  };
  for (Launch& L : launches) {
    if (L.kind == LK_GEMM)
      L.dev_off = put(L.gdescs.data(), L.gdescs.size() * sizeof(FmGemmDesc));
    else if (L.kind == LK_OZAKI)
      L.dev_off = put(L.odescs.data(), L.odescs.size() * sizeof(FmOzakiDesc));
    else if (L.kind == LK_SPLIT) {
      L.dev_off = put(L.sdescs.data(), L.sdescs.size() * sizeof(FmSplitDesc));
      L.dev_off2 = put(L.spreps.data(), L.spreps.size() * sizeof(FmSplitPrep));
    } else if (!L.snodes.empty())
      L.dev_off = put(L.snodes.data(), L.snodes.size() * sizeof(FmStageNode));
    else
      L.dev_off = put(L.ops.data(), L.ops.size() * sizeof(FmOp));
  }
  ext_off = put(ext.data(), ext.size() * sizeof(void*));
}

// Point every launch at its descriptors in a device copy of the packed buffer.
void Plan::bind(void* dev) {
  desc_base = dev;
  for (Launch& L : launches) {
    char* p = (char*)dev + L.dev_off;
    if (L.kind == LK_GEMM)
      L.gm.descs = reinterpret_cast<const FmGemmDesc*>(p);
    else if (L.kind == LK_OZAKI)
      L.oz.descs = reinterpret_cast<const FmOzakiDesc*>(p);
    else if (L.kind == LK_SPLIT) {
      L.sp.descs = reinterpret_cast<const FmSplitDesc*>(p);
      L.sp.preps = reinterpret_cast<const FmSplitPrep*>((char*)dev + L.dev_off2);
    } else if (!L.snodes.empty()) {
      L.st.node = reinterpret_cast<const FmStageNode*>(p);
      L.st.ops = nullptr;
    } else {
      L.st.ops = reinterpret_cast<const FmOp*>(p);
      L.st.node = nullptr;
    }
  }
}

fm_status Plan::upload() {
  std::vector<char> buf;
  pack(buf);
  if (buf.empty()) return FM_OK;
  hipError_t e = hipMalloc(&dbuf, buf.size());
  if (e != hipSuccess) {
    dbuf = nullptr;
    return hip_err(e, "hipMalloc(plan)");
  }
  e = hipMemcpy(dbuf, buf.data(), buf.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_err(e, "hipMemcpy(plan)");
  bind(dbuf);
  return FM_OK;
}

// ---- profiling ---------------------------------------------------------------
namespace {
struct ProfRec {
  int stage;
  hipEvent_t e0, e1;
};
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
}  // namespace

bool g_prof_on = false;

fm_status profile_read(fm_profile* out) {
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

fm_status Plan::run(const void* A, const void* B, void* C, void* ws, hipStream_t s) const {
  FmBases bases;
  bases.p[FM_ROLE_A] = const_cast<void*>(A);
  bases.p[FM_ROLE_B] = const_cast<void*>(B);
  bases.p[FM_ROLE_C] = C;
  bases.p[FM_ROLE_WS] = ws;
  bases.ext = ext.empty() ? nullptr : reinterpret_cast<void* const*>((char*)desc_base + ext_off);
  bool aligned = true;
  for (int i = 0; i < 4; ++i) aligned = aligned && ((uintptr_t)bases.p[i] % 16 == 0);
  for (void* p : ext) aligned = aligned && ((uintptr_t)p % 16 == 0);
  for (const Launch& L : launches) {
    hipError_t e = hipSuccess;
    ProfRec pr{L.stage, nullptr, nullptr};
    if (g_prof_on) {
      pr.e0 = pool_event();
      pr.e1 = pool_event();
      if (pr.e0) (void)hipEventRecord(pr.e0, s);
    }
    if (L.kind == LK_SPLIT && ((uintptr_t)ws % 16 != 0))
      return set_err(FM_ERR_INVALID_ARG, "split_bf16 leaves need a 16-byte aligned workspace");
    if (L.kind == LK_OZAKI || L.kind == LK_SPLIT) {
      // profiled: the plane preparation and the products are timed apart
      // (preparation as FM_STAGE_PREP)
      auto go = [&](int phases) {
        if (L.kind == LK_OZAKI) {
          FmOzakiArgs o = L.oz;
          o.bases = bases;
          return fm_ozaki_launch(L.oz_slices, o, L.oz_rows, L.oz_cols, phases, s);
        }
        FmSplitArgs a = L.sp;
        a.bases = bases;
        return fm_split_launch(L.cfg, a, phases, s);
      };
      if (g_prof_on && pr.e0 && pr.e1) {
        e = go(FM_PHASE_PREP);
        (void)hipEventRecord(pr.e1, s);
        g_prof_pending.push_back(ProfRec{FM_STAGE_PREP, pr.e0, pr.e1});
        pr.e0 = pool_event();
        pr.e1 = pool_event();
        if (pr.e0) (void)hipEventRecord(pr.e0, s);
        if (e == hipSuccess) e = go(FM_PHASE_PRODUCT);
      } else {
        e = go(FM_PHASE_ALL);
      }
    } else if (L.kind == LK_GEMM) {
      FmGemmArgs g = L.gm;
      g.bases = bases;
      g.vec = (L.vec_ok && aligned) ? 1 : 0;
      if (mode == FM_MODE_SPLIT_BF16)
        e = fm_gemm_launch_bf16x3(L.cfg, g, s);
      else
        e = dtype == FM_F64 ? fm_gemm_launch_f64(L.cfg, g, s) : fm_gemm_launch_f32(L.cfg, g, s);
    } else {
      FmStageArgs a = L.st;
      a.bases = bases;
      a.vec = aligned ? L.vec_ok : 1;
      const FmAlgDesc* g = L.alg;
      FmStageLaunch fn;
      const bool d = dtype == FM_F64;
      switch (L.kind) {
        case LK_S: fn = d ? g->s_d : g->s_f; break;
        case LK_T: fn = d ? g->t_d : g->t_f; break;
        case LK_S2: fn = d ? g->s2_d : L.planes ? g->s2p_f : g->s2_f; break;
        case LK_T2: fn = d ? g->t2_d : L.planes ? g->t2p_f : g->t2_f; break;
        case LK_W2: fn = d ? g->w2_d : g->w2_f; break;
        default: fn = d ? g->w_d : g->w_f; break;
      }
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

}  // namespace fmplan
