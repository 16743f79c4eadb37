// gemm_dv.hip -- grouped leaf GEMM with the B operand loaded straight into
// registers (FP64 and FP32 MFMA).
//
// Same contract and tile scheduling as gemm_mfma.hip (C = alpha A B + beta C,
// column-major, every product of a launch found from the concatenated tile
// list; the reference's leaf dgemm, linalg/linalg.hpp:261-269).  What
// differs is the data path:
//  * Within a 16-deep K slice, MFMA step s (0..3) takes k = 4g + s from lane
//    group g = lane / 16 (instead of k = 4s + g).  Every lane then needs four
//    consecutive k of one B column per slice: one contiguous 32-byte (FP64)
//    or 16-byte (FP32) run of column-major B, loaded with 16-byte loads
//    directly into the MFMA operand registers, one slice ahead.  B never
//    passes through LDS.
//  * Only the A tile is staged through LDS (As[k][m], double-buffered, with a
//    register prefetch of the next slice), so the LDS footprint is BM x 16
//    elements per buffer and two work groups share a CU.
// Each output element still accumulates every k exactly once in FP64 (FP32)
// MFMA arithmetic; only the order of k within a slice differs from
// gemm_mfma.hip.
#include <hip/hip_runtime.h>

#include "gemm_mfma.hpp"

namespace {

template <typename T>
struct DvMfma;

template <>
struct DvMfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ __forceinline__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output column (within a 16-column block) of accumulator item v, lane group g
  __device__ __forceinline__ static int out_col(int g, int v) { return g + 4 * v; }
};

template <>
struct DvMfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ __forceinline__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int out_col(int g, int v) { return 4 * g + v; }
};

__device__ __forceinline__ int dv_xcd_remap(int wg, int nwg) {
  const int full = nwg & ~7;
  if (wg < full) return (wg & 7) * (full >> 3) + (wg >> 3);
  return wg;
}

// four consecutive elements p[0..3] (16-byte aligned when vec)
template <typename T>
struct Quad {
  T v[4];
};

template <typename T>
__device__ __forceinline__ Quad<T> load_quad(const T* p, bool vec, int kleft, bool col_ok) {
  Quad<T> q;
  if (vec) {
    if constexpr (sizeof(T) == 8) {
      typedef double d2 __attribute__((ext_vector_type(2)));
      const d2 a = *reinterpret_cast<const d2*>(p);
      const d2 b = *reinterpret_cast<const d2*>(p + 2);
      q.v[0] = a[0];
      q.v[1] = a[1];
      q.v[2] = b[0];
      q.v[3] = b[1];
    } else {
      typedef float f4 __attribute__((ext_vector_type(4)));
      const f4 a = *reinterpret_cast<const f4*>(p);
      q.v[0] = a[0];
      q.v[1] = a[1];
      q.v[2] = a[2];
      q.v[3] = a[3];
    }
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) q.v[e] = (col_ok && e < kleft) ? p[e] : T(0);
  }
  return q;
}

}  // namespace

template <typename T, int BM, int BN, int WM, int WN, int MINW>
__global__ void __launch_bounds__(64 * WM * WN, MINW) fm_gemm_dv_kernel(FmGemmArgs args) {
  using Op = DvMfma<T>;
  using acc_t = typename Op::acc_t;
  constexpr int BK = 16;
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int VL = 16 / sizeof(T);            // elements per 16-byte load
  constexpr int A_PER = BM * BK / NT;            // A elements staged per thread
  constexpr int AV_PER = A_PER / VL;
  constexpr int AV_KSTEP = NT * VL / BM;         // k columns per pass of the vector mapping
  constexpr int A_KSTEP = NT / BM;               // k columns per pass of the element mapping
  static_assert(TM >= 1 && TN >= 1, "bad tile");
  static_assert(A_PER * NT == BM * BK && AV_PER * VL == A_PER, "tile/thread mismatch");
  static_assert(NT % BM == 0 && (NT * VL) % BM == 0 && BM % VL == 0, "tile/thread mismatch");
  typedef T vec_t __attribute__((ext_vector_type(VL)));

  __shared__ __align__(16) T As[2][BK][BM];

  for (int tile = blockIdx.x; tile < args.total_tiles; tile += gridDim.x) {
    const int wg = dv_xcd_remap(tile, args.total_tiles);
    int g;
    if (args.uniform) {
      g = wg / (args.descs[0].tiles_m * args.descs[0].tiles_n);
    } else {
      int lo = 0, hi = args.ndesc - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (args.descs[mid].tile_begin <= wg)
          lo = mid;
        else
          hi = mid - 1;
      }
      g = lo;
    }
    const FmGemmDesc d = args.descs[g];
    const int local = wg - d.tile_begin;
    const int m0 = (local % d.tiles_m) * BM, n0 = (local / d.tiles_m) * BN;
    const int M = d.M, N = d.N, K = d.K;
    const T* __restrict__ A = reinterpret_cast<const T*>(fm_base(args.bases, d.a.role)) + d.a.off;
    const T* __restrict__ B = reinterpret_cast<const T*>(fm_base(args.bases, d.b.role)) + d.b.off;
    T* __restrict__ C = reinterpret_cast<T*>(fm_base(args.bases, d.c.role)) + d.c.off;
    const int64_t lda = d.a.ld, ldb = d.b.ld, ldc = d.c.ld;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int lg = lane >> 4, l16 = lane & 15;
    const bool interior = args.vec && (m0 + BM <= M) && (n0 + BN <= N);

    // A staging coordinates (vector and element mappings)
    const int av_row = (tid % (BM / VL)) * VL, av_k = tid / (BM / VL);
    const T* av_ptr = A + (m0 + av_row) + (int64_t)av_k * lda;
    const int a_row = tid % BM, a_k = tid / BM;
    const bool a_row_ok = m0 + a_row < M;
    const T* a_ptr = A + (m0 + a_row) + (int64_t)a_k * lda;

    // B columns of this lane: one per 16-column block of the wave tile
    const T* b_ptr[TN];
    bool b_ok[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int n = n0 + wn * WTN + j * 16 + l16;
      b_ok[j] = n < N;
      b_ptr[j] = B + (int64_t)(b_ok[j] ? n : 0) * ldb + 4 * lg;
    }

    acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    T ra[A_PER];
    Quad<T> bcur[TN], bnext[TN];

    auto gload_a = [&](int k0, bool vec) {
      if (vec) {
#pragma unroll
        for (int i = 0; i < AV_PER; ++i) {
          const vec_t v = *reinterpret_cast<const vec_t*>(av_ptr + (int64_t)(k0 + i * AV_KSTEP) * lda);
#pragma unroll
          for (int e = 0; e < VL; ++e) ra[i * VL + e] = v[e];
        }
      } else {
#pragma unroll
        for (int i = 0; i < A_PER; ++i) {
          const int gk = k0 + a_k + i * A_KSTEP;
          ra[i] = (a_row_ok && gk < K) ? a_ptr[(int64_t)(k0 + i * A_KSTEP) * lda] : T(0);
        }
      }
    };
    auto sstore_a = [&](int buf, bool vec) {
      if (vec) {
#pragma unroll
        for (int i = 0; i < AV_PER; ++i) {
          vec_t v;
#pragma unroll
          for (int e = 0; e < VL; ++e) v[e] = ra[i * VL + e];
          *reinterpret_cast<vec_t*>(&As[buf][av_k + i * AV_KSTEP][av_row]) = v;
        }
      } else {
#pragma unroll
        for (int i = 0; i < A_PER; ++i) As[buf][a_k + i * A_KSTEP][a_row] = ra[i];
      }
    };
    auto load_b = [&](Quad<T> (&dst)[TN], int k0, bool vec) {
      const int kleft = K - k0 - 4 * lg;
#pragma unroll
      for (int j = 0; j < TN; ++j) dst[j] = load_quad<T>(b_ptr[j] + k0, vec, kleft, b_ok[j]);
    };

    const int nk = (K + BK - 1) / BK;
    // a slice takes the unpredicated path when the tile is interior and the
    // slice lies wholly inside K
    auto full = [&](int kt) { return interior && (kt + 1) * BK <= K; };

    gload_a(0, full(0));
    load_b(bcur, 0, full(0));
    sstore_a(0, full(0));
    if (nk > 1) {
      gload_a(BK, full(1));
      load_b(bnext, BK, full(1));
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        T af[TM];
#pragma unroll
        for (int i = 0; i < TM; ++i) af[i] = As[cur][4 * lg + s][wm * WTM + i * 16 + l16];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[i][j] = Op::mma(bcur[j].v[s], af[i], acc[i][j]);
      }
      if (kt + 1 < nk) {
        sstore_a(cur ^ 1, full(kt + 1));
#pragma unroll
        for (int j = 0; j < TN; ++j) bcur[j] = bnext[j];
        if (kt + 2 < nk) {
          gload_a((kt + 2) * BK, full(kt + 2));
          load_b(bnext, (kt + 2) * BK, full(kt + 2));
        }
      }
      __syncthreads();
    }

    const T alpha = (T)d.alpha, beta = (T)d.beta;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int row = m0 + wm * WTM + i * 16 + l16;
      if (row >= M) continue;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int col = n0 + wn * WTN + j * 16 + Op::out_col(lg, v);
          if (col < N) {
            T* p = C + row + (int64_t)col * ldc;
            T r = alpha * acc[i][j][v];
            if (beta != T(0)) r += beta * *p;
            *p = r;
          }
        }
      }
    }
  }  // tile loop
}

template <typename T, int BM, int BN, int WM, int WN, int MINW>
static void fm_gemm_dv_go(const FmGemmArgs& a, hipStream_t s) {
  const int grid = a.grid > 0 && a.grid < a.total_tiles ? a.grid : a.total_tiles;
  hipLaunchKernelGGL((fm_gemm_dv_kernel<T, BM, BN, WM, WN, MINW>), dim3(grid), dim3(64 * WM * WN),
                     0, s, a);
}

template <typename T>
static hipError_t fm_gemm_dv_launch(int cfg, const FmGemmArgs& a, hipStream_t s) {
  if (a.total_tiles <= 0) return hipSuccess;
  switch (cfg) {
    case FM_TILE_DV_128x128: fm_gemm_dv_go<T, 128, 128, 2, 2, 2>(a, s); break;
    case FM_TILE_DV_128x128_W8: fm_gemm_dv_go<T, 128, 128, 2, 4, 2>(a, s); break;
    case FM_TILE_DV_256x128_W16: fm_gemm_dv_go<T, 256, 128, 4, 4, 1>(a, s); break;
    case FM_TILE_DV_128x64: fm_gemm_dv_go<T, 128, 64, 2, 2, 2>(a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t fm_gemm_dv_launch_f64(int cfg, const FmGemmArgs& a, hipStream_t s) {
  return fm_gemm_dv_launch<double>(cfg, a, s);
}

hipError_t fm_gemm_dv_launch_f32(int cfg, const FmGemmArgs& a, hipStream_t s) {
  return fm_gemm_dv_launch<float>(cfg, a, s);
}
