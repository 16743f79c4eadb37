// probe.hip -- measured peaks on this GPU, for the roofline figures:
// * the matrix units: every SIMD issues back-to-back MFMAs on register
//   operands (no memory traffic) for a fixed count, timed with HIP events.
//   The rate this gives is the one the chip sustains at the clock it runs
//   under that load, which can sit below the published figure (quoted at
//   the peak engine clock);
// * HBM under a stage kernel's access pattern: each lane reads one element
//   position of `reads` slabs and writes it to `writes` slabs, as the S, T
//   and W stages do (with hundreds of output slabs the rate sits well below
//   a single write stream's).
#include <hip/hip_runtime.h>

#include <string>

#include "plan.hpp"

namespace {

typedef double fm_f64x4 __attribute__((ext_vector_type(4)));
typedef float fm_f32x4p __attribute__((ext_vector_type(4)));
typedef __bf16 fm_bf16x8p __attribute__((ext_vector_type(8)));
typedef int fm_i32x4p __attribute__((ext_vector_type(4)));
typedef int fm_i32x16p __attribute__((ext_vector_type(16)));

constexpr int kChains = 8;  // independent accumulators per wave: covers the MFMA latency
// MFMA rounds per loop trip: the compiler moves the accumulators between
// register files once per trip, which this many rounds make negligible
constexpr int kRounds = 16;

// kind 0: v_mfma_f64_16x16x4f64, 1: v_mfma_f32_16x16x4f32,
// 2: v_mfma_f32_16x16x32_bf16, 3: v_mfma_i32_32x32x32_i8; `iters` trips of
// kRounds x kChains MFMAs per wave
template <int KIND>
__global__ void __launch_bounds__(256) fm_mfma_probe(int iters, float* out) {
  const int t = threadIdx.x;
  float sink = 0.f;
  if constexpr (KIND == 0) {
    const double a = 1.0 + t * 1e-3, b = 1.0 - t * 1e-3;
    fm_f64x4 acc[kChains];
    for (int c = 0; c < kChains; ++c) acc[c] = fm_f64x4{0, 0, 0, 0};
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int u = 0; u < kRounds * kChains; ++u) acc[u % kChains] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[u % kChains], 0, 0, 0);
    for (int c = 0; c < kChains; ++c) sink += (float)(acc[c][0] + acc[c][3]);
  } else if constexpr (KIND == 1) {
    const float a = 1.f + t * 1e-3f, b = 1.f - t * 1e-3f;
    fm_f32x4p acc[kChains];
    for (int c = 0; c < kChains; ++c) acc[c] = fm_f32x4p{0, 0, 0, 0};
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int u = 0; u < kRounds * kChains; ++u) acc[u % kChains] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[u % kChains], 0, 0, 0);
    for (int c = 0; c < kChains; ++c) sink += acc[c][0] + acc[c][3];
  } else if constexpr (KIND == 2) {
    fm_bf16x8p a, b;
    for (int i = 0; i < 8; ++i) {
      a[i] = (__bf16)(1.f + (t + i) * 1e-3f);
      b[i] = (__bf16)(1.f - (t + i) * 1e-3f);
    }
    fm_f32x4p acc[kChains];
    for (int c = 0; c < kChains; ++c) acc[c] = fm_f32x4p{0, 0, 0, 0};
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int u = 0; u < kRounds * kChains; ++u) acc[u % kChains] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[u % kChains], 0, 0, 0);
    for (int c = 0; c < kChains; ++c) sink += acc[c][0] + acc[c][3];
  } else {
    const fm_i32x4p a = fm_i32x4p{t & 7, 1, 2, 3}, b = fm_i32x4p{1, (t >> 3) & 7, 3, 1};
    fm_i32x16p acc[kChains];
    for (int c = 0; c < kChains; ++c)
      for (int i = 0; i < 16; ++i) acc[c][i] = 0;
    for (int it = 0; it < iters; ++it)
#pragma unroll
      for (int u = 0; u < kRounds * kChains; ++u) acc[u % kChains] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[u % kChains], 0, 0, 0);
    for (int c = 0; c < kChains; ++c) sink += (float)(acc[c][0] + acc[c][15]);
  }
  out[blockIdx.x * 256 + t] = sink;
}

// operations (2 per multiply-add) of one MFMA of each kind
constexpr double kOps[4] = {2.0 * 16 * 16 * 4, 2.0 * 16 * 16 * 4, 2.0 * 16 * 16 * 32,
                            2.0 * 32 * 32 * 32};

template <int KIND>
hipError_t launch(int blocks, int iters, float* out, hipStream_t s) {
  hipLaunchKernelGGL((fm_mfma_probe<KIND>), dim3(blocks), dim3(256), 0, s, iters, out);
  return hipGetLastError();
}

hipError_t launch_kind(int kind, int blocks, int iters, float* out, hipStream_t s) {
  switch (kind) {
    case 0: return launch<0>(blocks, iters, out, s);
    case 1: return launch<1>(blocks, iters, out, s);
    case 2: return launch<2>(blocks, iters, out, s);
    default: return launch<3>(blocks, iters, out, s);
  }
}

// Stage-pattern stream: position e of `reads` input slabs and `writes`
// output slabs (slab t at t * elems), one lane per position.  Sources are
// loaded 64 at a time, all in flight together as in the stage kernels (a
// last partial batch reloads the last slab, which the cache serves).
constexpr int kProbeBatch = 64;

__global__ void __launch_bounds__(256) fm_stream_probe(const double* __restrict__ in,
                                                       double* __restrict__ out, int reads,
                                                       int writes, int64_t elems) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < elems;
       e += (int64_t)gridDim.x * 256) {
    double acc = 0.0;
    for (int b0 = 0; b0 < reads; b0 += kProbeBatch) {
      double v[kProbeBatch];
#pragma unroll
      for (int i = 0; i < kProbeBatch; ++i)
        v[i] = in[(int64_t)min(b0 + i, reads - 1) * elems + e];
#pragma unroll
      for (int i = 0; i < kProbeBatch; ++i) acc += b0 + i < reads ? v[i] : 0.0;
    }
    double* op = out + e;
#pragma unroll 16
    for (int t = 0; t < writes; ++t) {
      *op = acc + t;
      op += elems;
    }
  }
}

}  // namespace

using namespace fmplan;

extern "C" fm_status fm_probe_stream(int reads, int writes, int64_t elems, double* gb_per_s) {
  clear_error();
  if (!gb_per_s || reads < 0 || writes < 1 || reads > 4096 || writes > 4096 || elems < 1 ||
      elems > (1LL << 30) || (double)(reads + writes) * elems * 8 > 32e9)
    return set_err(FM_ERR_INVALID_ARG,
                   "reads 0..4096, writes 1..4096, 1 <= elems <= 2^30, at most 32 GB in all");
  *gb_per_s = 0.0;
  double *in = nullptr, *out = nullptr;
  hipError_t e = hipSuccess;
  if (reads > 0) e = hipMalloc(&in, sizeof(double) * (size_t)reads * (size_t)elems);
  if (e == hipSuccess) e = hipMalloc(&out, sizeof(double) * (size_t)writes * (size_t)elems);
  if (e == hipSuccess && reads > 0)
    e = hipMemset(in, 0, sizeof(double) * (size_t)reads * (size_t)elems);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  fm_status st = FM_OK;
  if (e != hipSuccess) {
    st = hip_err(e, "hipMalloc(stream probe)");
  } else if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    st = set_err(FM_ERR_HIP, "hipEventCreate");
  } else {
    const int64_t wgs = (elems + 255) / 256;
    const dim3 grid((unsigned)(wgs < 65536 ? wgs : 65536));
    const int reps = 5;
    hipLaunchKernelGGL(fm_stream_probe, grid, dim3(256), 0, nullptr, in, out, reads, writes, elems);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      hipLaunchKernelGGL(fm_stream_probe, grid, dim3(256), 0, nullptr, in, out, reads, writes,
                         elems);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess)
      st = hip_err(e, "stream probe launch");
    else if (ms > 0.f)
      *gb_per_s = (double)(reads + writes) * elems * 8 * reps / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (in) (void)hipFree(in);
  if (out) (void)hipFree(out);
  return st;
}

extern "C" fm_status fm_probe_mfma_peak(int kind, int iters, double* tera_ops_per_s) {
  // iters counts single MFMA rounds; the kernel runs them kRounds per trip
  iters = (iters + kRounds - 1) / kRounds;
  clear_error();
  if (kind < 0 || kind > 3 || iters <= 0 || !tera_ops_per_s)
    return set_err(FM_ERR_INVALID_ARG, "kind must be 0..3 and iters > 0");
  *tera_ops_per_s = 0.0;
  hipDeviceProp_t prop;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return hip_err(e, "hipGetDeviceProperties");
  // eight waves per SIMD: every matrix unit busy, latency hidden
  const int blocks = prop.multiProcessorCount * 8;
  float* out = nullptr;
  e = hipMalloc(&out, sizeof(float) * 256 * (size_t)blocks);
  if (e != hipSuccess) return hip_err(e, "hipMalloc(probe)");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  fm_status st = FM_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    st = set_err(FM_ERR_HIP, "hipEventCreate");
  } else {
    e = launch_kind(kind, blocks, iters / 10 + 1, out, nullptr);  // warm-up: clocks settle
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_kind(kind, blocks, iters, out, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess)
      st = hip_err(e, "probe launch");
    else if (ms > 0.f)
      *tera_ops_per_s = kOps[kind] * kRounds * kChains * (double)iters * blocks * 4 / (ms * 1e-3) / 1e12;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return st;
}
