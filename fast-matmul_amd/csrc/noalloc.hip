// noalloc.hip -- fm_fastmatmul_noalloc: fm_fastmatmul for callers that may
// not allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or
// allocating library routine: the plan's workspace (S/T operands, products,
// slice and digit planes) and its launch descriptors live in `scratch`, a
// fixed array in device memory that every device the library runs on holds
// until the process ends.
//
// Host code never uses the array's name as an address: on the host it is a
// stand-in, and a kernel handed it would fault.  The device address of each
// device's copy comes from hipGetSymbolAddress, looked up once per device
// (with that device current) on the first call there, and is what the
// launches receive.
//
// All calls on one device share its array, which keeps the last call's
// contents; nothing here needs it zeroed (every kernel writes what it later
// reads).  Calls that overlap on one device -- on two streams, or from two
// threads without ordering -- race on it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"

using namespace fmplan;

extern std::mutex& fm_api_mutex();
extern const FmAlgDesc* fm_find_alg(const char* name);

// 256 MiB per device: descriptors plus workspace (e.g. a Strassen level of
// 2048^3 FP64 needs 136 MiB of S, T and products).
static __device__ __align__(256) unsigned char scratch[FM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up once per device).
fm_status scratch_ptr(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return set_err(FM_ERR_INVALID_ARG, "device index out of range");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_err(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_err(FM_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FM_OK;
}

// Plans of this entry point, with their packed descriptors kept on the host
// (copied into the scratch array by every call, since other calls overwrite
// it).  Bounded like the main plan cache.
struct NoallocPlan {
  Plan plan;
  std::vector<char> desc;
  size_t ws_off = 0;  // workspace offset in the scratch array (after the descriptors)
};
constexpr size_t kMaxNoallocPlans = 64;
std::map<std::string, std::unique_ptr<NoallocPlan>> g_na_plans;
std::deque<std::string> g_na_order;

}  // namespace

extern "C" {

fm_status fm_fastmatmul_noalloc(const char* alg_name, fm_dtype dtype, fm_mode mode, int m, int k,
                                int n, const void* A, int64_t lda, const void* B, int64_t ldb,
                                void* C, int64_t ldc, int num_steps, double x, double alpha,
                                double beta, void* stream, float* ms_out) {
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
  if (m == 0 || n == 0) return FM_OK;
  if (k == 0) return set_err(FM_ERR_SHAPE, "k == 0");
  std::unique_lock<std::mutex> lock(fm_api_mutex());
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_err(e, "hipGetDevice");
  unsigned char* base = nullptr;
  st = scratch_ptr(dev, &base);
  if (st != FM_OK) return st;
  const fm_options o = g_defaults;
  char key[512];
  snprintf(key, sizeof(key), "%d|%s|%d|%d|%d|%d|%d|%d|%zu|%d|%d|%d|%d|%lld|%lld|%lld|%d|%a|%a|%a",
           dev, alg_name, (int)dtype, (int)mode, o.ozaki_slices, o.stage_fusion != 0,
           o.split_remainder != 0, o.dfs_levels, o.workspace_limit, o.gemm_tile, m, k, n,
           (long long)lda, (long long)ldb, (long long)ldc, num_steps, x, alpha, beta);
  NoallocPlan* np = nullptr;
  auto it = g_na_plans.find(key);
  if (it != g_na_plans.end()) {
    np = it->second.get();
  } else {
    std::unique_ptr<NoallocPlan> fresh(new NoallocPlan());
    st = build_single(fresh->plan, alg, dtype, mode, m, k, n, lda, ldb, ldc, num_steps, x, alpha,
                      beta, o);
    if (st != FM_OK) return st;
    fresh->plan.pack(fresh->desc);
    fresh->ws_off = (size_t)align_up((int64_t)fresh->desc.size(), 256);
    const size_t need = fresh->ws_off + fresh->plan.ws_bytes;
    if (need > (size_t)FM_NOALLOC_SCRATCH_BYTES)
      return set_err(FM_ERR_WORKSPACE,
                     "fm_fastmatmul_noalloc: this multiply needs " + std::to_string(need) +
                         " bytes of scratch (descriptors and workspace); the fixed scratch holds " +
                         std::to_string((size_t)FM_NOALLOC_SCRATCH_BYTES) +
                         " (use fm_fastmatmul with a workspace, or fewer levels)");
    if (g_na_plans.size() >= kMaxNoallocPlans) {
      // queued copies may still read an evicted plan's host descriptors
      e = hipDeviceSynchronize();
      if (e != hipSuccess) return hip_err(e, "hipDeviceSynchronize");
      while (g_na_plans.size() >= kMaxNoallocPlans && !g_na_order.empty()) {
        g_na_plans.erase(g_na_order.front());
        g_na_order.pop_front();
      }
    }
    np = fresh.get();
    g_na_plans[key] = std::move(fresh);
    g_na_order.push_back(key);
  }
  const hipStream_t s = (hipStream_t)stream;
  // descriptors first in the array, the workspace after them
  np->plan.bind(base);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms_out) {
    if (hipEventCreate(&e0) != hipSuccess) return set_err(FM_ERR_HIP, "hipEventCreate");
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      return set_err(FM_ERR_HIP, "hipEventCreate");
    }
    (void)hipEventRecord(e0, s);
  }
  if (!np->desc.empty()) {
    e = hipMemcpyAsync(base, np->desc.data(), np->desc.size(), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) st = hip_err(e, "hipMemcpyAsync(descriptors)");
  }
  if (st == FM_OK) st = np->plan.run(A, B, C, base + np->ws_off, s);
  if (ms_out) {
    (void)hipEventRecord(e1, s);
    lock.unlock();
    if (st == FM_OK) {
      e = hipEventSynchronize(e1);
      if (e != hipSuccess)
        st = hip_err(e, "hipEventSynchronize");
      else
        (void)hipEventElapsedTime(ms_out, e0, e1);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return st;
}

}  // extern "C"
