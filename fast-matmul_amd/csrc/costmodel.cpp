// costmodel.cpp -- the autotuner's cost model (see costmodel.hpp).
//
// Table format, one entry per line ('#' starts a comment):
//   leaf <f64|f32> <plain|split_bf16|ozaki_int8> <size> <tflops>
//   stage_gbps <f64|f32> <GB/s>
//   launch_us <us>
//   level_us <f64|f32> <us>
//   tile_k_us <f64|f32> <us per unit of K>
#include "costmodel.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

namespace fmplan {

namespace {

CostTable builtin() {
  CostTable t;
  // flat rates measured on MI355X before calibration: FP64 leaves 60, FP32
  // 100, split-bf16 220, int8 digits 65 TFLOP/s (FP64-equivalent)
  t.leaf[0][0] = {{1024.0, 60.0}};
  t.leaf[1][0] = {{1024.0, 100.0}};
  t.leaf[1][1] = {{1024.0, 220.0}};
  t.leaf[0][2] = {{1024.0, 65.0}};
  return t;
}

int dtype_index(const std::string& s) { return s == "f64" ? 0 : s == "f32" ? 1 : -1; }
int mode_index(const std::string& s) {
  return s == "plain" ? 0 : s == "split_bf16" ? 1 : s == "ozaki_int8" ? 2 : -1;
}

bool parse(std::istream& in, CostTable& t, std::string* err) {
  CostTable r;
  std::string line;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ss(line);
    std::string key;
    if (!(ss >> key)) continue;
    bool ok = true;
    if (key == "leaf") {
      std::string dt, mode;
      double size = 0, tf = 0;
      ok = (bool)(ss >> dt >> mode >> size >> tf);
      const int d = dtype_index(dt), m = mode_index(mode);
      ok = ok && d >= 0 && m >= 0 && size > 0 && tf > 0;
      if (ok) r.leaf[d][m].emplace_back(size, tf);
    } else if (key == "stage_gbps" || key == "level_us" || key == "tile_k_us") {
      std::string dt;
      double v = 0;
      ok = (bool)(ss >> dt >> v);
      const int d = dtype_index(dt);
      ok = ok && d >= 0 && v >= 0 && (key != "stage_gbps" || v > 0);
      if (ok) (key == "stage_gbps" ? r.stage_gbps : key == "level_us" ? r.level_us : r.tile_k_us)[d] = v;
    } else if (key == "launch_us") {
      ok = (bool)(ss >> r.launch_us) && r.launch_us >= 0;
    } else {
      ok = false;
    }
    if (!ok) {
      if (err) *err = "line " + std::to_string(lineno) + ": cannot read \"" + line + "\"";
      return false;
    }
  }
  // modes the table does not measure keep the built-in rates
  const CostTable b = builtin();
  for (int d = 0; d < 2; ++d)
    for (int m = 0; m < 3; ++m) {
      if (r.leaf[d][m].empty()) r.leaf[d][m] = b.leaf[d][m];
      std::sort(r.leaf[d][m].begin(), r.leaf[d][m].end());
    }
  t = r;
  return true;
}

std::mutex g_table_mu;
CostTable* g_table = nullptr;

std::string default_path() {
  if (const char* env = std::getenv("FM_TUNING_FILE")) return env;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&default_path), &info) && info.dli_fname) {
    std::string so = info.dli_fname;
    const size_t slash = so.rfind('/');
    if (slash != std::string::npos) return so.substr(0, slash) + "/tuning/mi355x.txt";
  }
  return "";
}

// Leaf rate at `size`, interpolated linearly in log(size) between measured
// sizes and held constant beyond them.
double leaf_rate(const CostTable& t, fm_dtype dt, fm_mode mode, double size) {
  const auto& pts = t.leaf[dt == FM_F64 ? 0 : 1][(int)mode];
  if (pts.empty()) return dt == FM_F64 ? 60.0 : 100.0;
  if (size <= pts.front().first) return pts.front().second;
  if (size >= pts.back().first) return pts.back().second;
  for (size_t i = 1; i < pts.size(); ++i)
    if (size <= pts[i].first) {
      const double x0 = std::log(pts[i - 1].first), x1 = std::log(pts[i].first);
      const double w = (std::log(size) - x0) / (x1 - x0);
      return pts[i - 1].second + w * (pts[i].second - pts[i - 1].second);
    }
  return pts.back().second;
}

}  // namespace

const CostTable& cost_table() {
  std::lock_guard<std::mutex> lock(g_table_mu);
  if (!g_table) {
    g_table = new CostTable(builtin());
    const std::string path = default_path();
    std::ifstream f(path);
    CostTable t;
    if (f && parse(f, t, nullptr)) {
      t.source = path;
      *g_table = t;
    }
  }
  return *g_table;
}

fm_status cost_table_load(const char* path) {
  if (!path) return set_err(FM_ERR_INVALID_ARG, "null path");
  std::ifstream f(path);
  if (!f) return set_err(FM_ERR_INVALID_ARG, std::string("cannot open ") + path);
  CostTable t;
  std::string err;
  if (!parse(f, t, &err)) return set_err(FM_ERR_INVALID_ARG, std::string(path) + ": " + err);
  t.source = path;
  (void)cost_table();  // make sure the default is loaded first, then replace it
  std::lock_guard<std::mutex> lock(g_table_mu);
  *g_table = t;
  return FM_OK;
}

double estimate_us(fm_dtype dt, fm_mode mode, const FmAlgDesc* alg, int64_t m, int64_t k, int64_t n,
                   int steps) {
  const CostTable& t = cost_table();
  const int di = dt == FM_F64 ? 0 : 1;
  const double es = elem_size(dt);
  auto padded = [](int64_t v) { return (double)align_up(v, 64); };
  auto leaf_us = [&](double nodes, int64_t lm, int64_t lk, int64_t ln) {
    const double size = std::cbrt((double)lm * lk * ln);
    return nodes * 2.0 * padded(lm) * padded(ln) * lk / (leaf_rate(t, dt, mode, size) * 1e6);
  };
  // int8-digit leaves keep exact int32 sums only up to a K bound: a
  // candidate whose leaves are deeper cannot run in that mode
  auto too_deep = [&](int64_t lk) {
    return mode == FM_MODE_OZAKI_INT8 && align_up(lk, 64) > fm_ozaki_max_kp(g_defaults.ozaki_slices);
  };
  if (!alg || steps == 0) return too_deep(k) ? INFINITY : leaf_us(1.0, m, k, n) + t.launch_us;
  double us = 0.0;
  int64_t rm = m, rk = k, rn = n;
  double nodes = 1.0;
  for (int l = 0; l < steps; ++l) {
    const int64_t bm = rm / alg->m, bk = rk / alg->k, bn = rn / alg->n;
    if (bm <= 0 || bk <= 0 || bn <= 0) return INFINITY;
    const double blkA = (double)bm * bk, blkB = (double)bk * bn, blkC = (double)bm * bn;
    // S: read m*k blocks, write NS; T likewise; W: read R, write m*n
    const double bytes = es * nodes *
                         (blkA * (alg->m * alg->k + alg->ns) + blkB * (alg->k * alg->n + alg->nt) +
                          blkC * (alg->R + alg->m * alg->n));
    us += bytes / (t.stage_gbps[di] * 1e3);
    // peeling products of this level's nodes, at the rate of their size;
    // their launch lasts at least as long as one tile takes over its K
    const int64_t erA = rm % alg->m, ecA = rk % alg->k, ecB = rn % alg->n;
    double peel = 0.0;
    int64_t peel_k = 0;
    if (ecA > 0) peel += leaf_us(nodes, rm - erA, ecA, rn - ecB), peel_k = std::max(peel_k, ecA);
    if (ecB > 0) peel += leaf_us(nodes, rm, rk, ecB), peel_k = std::max(peel_k, rk);
    if (erA > 0) peel += leaf_us(nodes, erA, rk, rn - ecB), peel_k = std::max(peel_k, rk);
    us += std::max(peel, (double)peel_k * t.tile_k_us[di]);
    us += t.level_us[di];
    nodes *= alg->R;
    rm = bm;
    rk = bk;
    rn = bn;
  }
  if (too_deep(rk)) return INFINITY;
  return us + leaf_us(nodes, rm, rk, rn) + t.launch_us;
}

}  // namespace fmplan
