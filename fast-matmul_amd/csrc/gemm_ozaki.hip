// gemm_ozaki.hip -- FP64-accurate leaf products from int8 slices (Ozaki scheme).
//
// Row i of A is scaled by 2^-e_i (|A[i,:]| < 2^e_i) and cut into S signed
// 7-bit digits by truncation: a = 2^e_i * sum_{s=1..S} 2^-7s alpha_s + rho,
// |alpha_s| <= 127, |rho| < 2^(e_i - 7S).  Columns of B likewise (2^f_j,
// beta_t).  Then
//   (A B)[i,j] ~= 2^(e_i + f_j) * sum_{u=2..S+1} 2^-7u * sum_{s+t=u} (alpha_s beta_t)[i,j]
// where each alpha_s beta_t is an int8 product with exact int32 accumulation
// (v_mfma_i32_16x16x64_i8).  The per-u integer sums stay exact in int32 as
// long as K * (u-1) * 127^2 < 2^31 (checked on the host), the FP64
// reconstruction adds the S groups from the smallest scale up, and the
// dropped digit pairs (s + t > S + 1) bound the error by about
// (S+1) 2^-7S relative to |A[i,:]|max |B[:,j]|max per term.
//
// Two preparation kernels write the digit planes once per operand into the
// workspace (A: K contiguous per row, B: K contiguous per column, K padded
// to 64), then one grouped kernel runs every leaf product of the launch.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gemm_ozaki.hpp"

typedef int fm_i32x4 __attribute__((ext_vector_type(4)));

// exponent e with |x| < 2^e for the largest |x| (e = -1074 for all zeros)
__device__ __forceinline__ int fm_exp_bound(double mx) {
  if (!(mx > 0.0)) return -1074;
  int e;
  (void)frexp(mx, &e);  // mx = f 2^e, 0.5 <= f < 1  =>  mx < 2^e
  return e;
}

// Digits of y (|y| < 1): S truncated 7-bit slices, each in [-127, 127].
template <int S>
__device__ __forceinline__ void fm_digits(double y, int8_t (&d)[S]) {
#pragma unroll
  for (int s = 0; s < S; ++s) {
    y = y * 128.0;               // exact: power-of-two scaling
    const double t = trunc(y);   // |t| <= 127 since |y| < 128
    d[s] = (int8_t)(int)t;
    y = y - t;                   // exact: t holds y's leading bits
  }
}

// ---- A side: one work group per 64 rows of one operand ---------------------
// Row exponents to ws + exp_off (int32), digits to ws + slc_off:
// plane s, row r, byte k at slc_off + (s * rows_pad + r) * kp + k.
template <int S>
__global__ void __launch_bounds__(256) fm_ozaki_prep_rows(FmOzakiArgs args) {
  const FmOzakiDesc d = args.descs[blockIdx.y];
  const int r0 = blockIdx.x * 64;
  if (r0 >= d.M) return;
  const double* A = reinterpret_cast<const double*>(fm_base(args.bases, d.a.role)) + d.a.off;
  const int64_t lda = d.a.ld;
  char* ws = reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]);
  int32_t* ex = reinterpret_cast<int32_t*>(ws + d.a_exp);
  int8_t* sl = reinterpret_cast<int8_t*>(ws + d.a_slc);
  const int tid = threadIdx.x, r = tid & 63, kq = tid >> 6;  // 4 threads per row
  const int row = r0 + r;
  const bool rok = row < d.M;
  __shared__ double red[4][64];
  double mx = 0.0;
  if (rok)
    for (int k = kq; k < d.K; k += 4) mx = fmax(mx, fabs(A[row + (int64_t)k * lda]));
  red[kq][r] = mx;
  __syncthreads();
  mx = fmax(fmax(red[0][r], red[1][r]), fmax(red[2][r], red[3][r]));
  const int e = fm_exp_bound(mx);
  if (kq == 0 && rok) ex[row] = e;
  // digits: each thread 16 consecutive k of its row per 64-wide chunk
  for (int k0 = 0; k0 < d.Kp; k0 += 64) {
    int8_t dig[S][16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = k0 + 16 * kq + q;
      const double v = (rok && k < d.K) ? ldexp(A[row + (int64_t)k * lda], -e) : 0.0;
      int8_t t[S];
      fm_digits<S>(v, t);
#pragma unroll
      for (int s = 0; s < S; ++s) dig[s][q] = t[s];
    }
    if (row < d.rows_pad) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        int8_t* p = sl + ((int64_t)s * d.rows_pad + row) * d.Kp + k0 + 16 * kq;
        *reinterpret_cast<int4*>(p) = *reinterpret_cast<const int4*>(&dig[s][0]);
      }
    }
  }
}

// ---- B side: one wave per column ---------------------------------------------
// plane t, column c, byte k at b_slc + (t * cols_pad + c) * kp + k.
template <int S>
__global__ void __launch_bounds__(256) fm_ozaki_prep_cols(FmOzakiArgs args) {
  const FmOzakiDesc d = args.descs[blockIdx.y];
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (col >= d.cols_pad) return;
  const double* B = reinterpret_cast<const double*>(fm_base(args.bases, d.b.role)) + d.b.off;
  char* ws = reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]);
  int32_t* ex = reinterpret_cast<int32_t*>(ws + d.b_exp);
  int8_t* sl = reinterpret_cast<int8_t*>(ws + d.b_slc);
  const bool cok = col < d.N;
  const double* bc = B + (int64_t)col * d.b.ld;
  double mx = 0.0;
  if (cok)
    for (int k = lane; k < d.K; k += 64) mx = fmax(mx, fabs(bc[k]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  const int e = fm_exp_bound(mx);
  if (lane == 0 && cok) ex[col] = e;
  for (int k0 = 16 * lane; k0 < d.Kp; k0 += 16 * 64) {
    int8_t dig[S][16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int k = k0 + q;
      const double v = (cok && k < d.K) ? ldexp(bc[k], -e) : 0.0;
      int8_t t[S];
      fm_digits<S>(v, t);
#pragma unroll
      for (int s = 0; s < S; ++s) dig[s][q] = t[s];
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      int8_t* p = sl + ((int64_t)s * d.cols_pad + col) * d.Kp + k0;
      *reinterpret_cast<int4*>(p) = *reinterpret_cast<const int4*>(&dig[s][0]);
    }
  }
}

// ---- grouped digit product + FP64 reconstruction -------------------------------
// Work group: 2 x 2 waves, each a 32 x 32 block of 16 x 16 MFMA tiles
// (64 x 64 output tile); K in 64-deep slices through LDS (all S planes).
template <int S>
__global__ void __launch_bounds__(256) fm_ozaki_gemm(FmOzakiArgs args) {
  constexpr int BM = 64, BN = 64, BK = 64, KP = BK + 16;  // padded LDS row (bytes)
  constexpr int TM = 2, TN = 2;
  __shared__ __align__(16) int8_t As[S][BM][KP];
  __shared__ __align__(16) int8_t Bs[S][BN][KP];

  int wg = blockIdx.x;
  {
    const int nwg = gridDim.x, full = nwg & ~7;
    if (wg < full) wg = (wg & 7) * (full >> 3) + (wg >> 3);
  }
  int lo = 0, hi = args.ndesc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (args.descs[mid].tile_begin <= wg)
      lo = mid;
    else
      hi = mid - 1;
  }
  const FmOzakiDesc d = args.descs[lo];
  const int local = wg - d.tile_begin;
  const int m0 = (local % d.tiles_m) * BM, n0 = (local / d.tiles_m) * BN;
  char* ws = reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]);
  const int8_t* sa = reinterpret_cast<const int8_t*>(ws + d.a_slc);
  const int8_t* sb = reinterpret_cast<const int8_t*>(ws + d.b_slc);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  fm_i32x4 acc[S][TM][TN];  // one exact integer sum per scale group u = s + t + 2
#pragma unroll
  for (int u = 0; u < S; ++u)
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) acc[u][i][j] = fm_i32x4{0, 0, 0, 0};

  // tile loads: S planes x 64 rows x 64 bytes per side = S * 256 int4 chunks
  const int lr = tid >> 2, lc = (tid & 3) * 16;
  for (int k0 = 0; k0 < d.Kp; k0 += BK) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int4 va = *reinterpret_cast<const int4*>(sa + ((int64_t)s * d.rows_pad + m0 + lr) * d.Kp + k0 + lc);
      const int4 vb = *reinterpret_cast<const int4*>(sb + ((int64_t)s * d.cols_pad + n0 + lr) * d.Kp + k0 + lc);
      *reinterpret_cast<int4*>(&As[s][lr][lc]) = va;
      *reinterpret_cast<int4*>(&Bs[s][lr][lc]) = vb;
    }
    __syncthreads();
    const int kofs = 16 * (lane >> 4);
    fm_i32x4 af[S][TM], bf[S][TN];
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
        af[s][i] = *reinterpret_cast<const fm_i32x4*>(&As[s][wm * 32 + i * 16 + (lane & 15)][kofs]);
#pragma unroll
      for (int j = 0; j < TN; ++j)
        bf[s][j] = *reinterpret_cast<const fm_i32x4*>(&Bs[s][wn * 32 + j * 16 + (lane & 15)][kofs]);
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int t = 0; t + s < S; ++t)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[s + t][i][j] =
                __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[t][j], af[s][i], acc[s + t][i][j], 0, 0, 0);
    __syncthreads();
  }

  const int32_t* ea = reinterpret_cast<const int32_t*>(ws + d.a_exp);
  const int32_t* eb = reinterpret_cast<const int32_t*>(ws + d.b_exp);
  double* C = reinterpret_cast<double*>(fm_base(args.bases, d.c.role)) + d.c.off;
  const int64_t ldc = d.c.ld;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int row = m0 + wm * 32 + i * 16 + (lane & 15);
    if (row >= d.M) continue;
    const int er = ea[row];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = n0 + wn * 32 + j * 16 + 4 * (lane >> 4) + v;
        if (col >= d.N) continue;
        // smallest scale first: group u = s + t + 2 carries 2^-7u
        // (every step is exact or one explicitly rounded FP64 operation)
        double sum = 0.0;
#pragma unroll
        for (int u = S - 1; u >= 0; --u)
          sum = __dadd_rn(sum, ldexp((double)acc[u][i][j][v], -7 * (u + 2)));
        double r = __dmul_rn(ldexp(sum, er + eb[col]), d.alpha);
        double* p = C + row + (int64_t)col * ldc;
        if (d.beta != 0.0) r = __dadd_rn(r, __dmul_rn(d.beta, *p));
        *p = r;
      }
    }
  }
}

template <int S>
static hipError_t fm_ozaki_run(const FmOzakiArgs& a, int max_rows, int max_cols, hipStream_t s) {
  hipLaunchKernelGGL((fm_ozaki_prep_rows<S>), dim3((max_rows + 63) / 64, a.ndesc), dim3(256), 0,
                     s, a);
  hipLaunchKernelGGL((fm_ozaki_prep_cols<S>), dim3((max_cols + 3) / 4, a.ndesc), dim3(256), 0, s,
                     a);
  hipLaunchKernelGGL((fm_ozaki_gemm<S>), dim3(a.total_tiles), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t fm_ozaki_launch(int slices, const FmOzakiArgs& a, int max_rows, int max_cols,
                           hipStream_t s) {
  if (a.total_tiles <= 0) return hipSuccess;
  switch (slices) {
    case 6: return fm_ozaki_run<6>(a, max_rows, max_cols, s);
    case 7: return fm_ozaki_run<7>(a, max_rows, max_cols, s);
    case 8: return fm_ozaki_run<8>(a, max_rows, max_cols, s);
    default: return hipErrorInvalidValue;
  }
}
