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
// workspace, then one grouped kernel runs every leaf product of the launch.
// Digit layout of an operand with rows_pad rows (A: rows, B: columns), K
// padded to 64: digit s of (row r, k) is the byte at
//   slc + ((k / 32 * 2 S + 2 s + (k / 16) % 2) * rows_pad + r) * 16 + k % 16,
// i.e. for every 16-deep half of a 32-deep K slice and every plane, the
// 16-byte digit groups of all rows are consecutive: the preparation writes
// and the product kernel's loads and LDS stores all move 1 KB runs per wave.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gemm_ozaki.hpp"
#include "gmem.hpp"

typedef int fm_i32x4 __attribute__((ext_vector_type(4)));

// exponent e with |x| < 2^e for the largest |x| (e = -1074 for all zeros)
__device__ __forceinline__ int fm_exp_bound(double mx) {
  if (!(mx > 0.0)) return -1074;
  int e;
  (void)frexp(mx, &e);  // mx = f 2^e, 0.5 <= f < 1  =>  mx < 2^e
  return e;
}

// Digits of y (|y| < 1): S truncated 7-bit slices, each in [-127, 127]:
// d_s = sign(y) * (floor(|y| 2^(7 s)) mod 128), s = 1..S (S <= 8), the
// digits that repeated truncation (y <- 128 y - trunc(128 y)) would give.
// They are read as bit fields of h = floor(|y| 2^28) and
// l = floor((|y| 2^28 - h) 2^28), both exact (power-of-two scalings and
// the difference of a value and its integer part).
template <int S>
__device__ __forceinline__ void fm_digits(double y, int8_t (&d)[S]) {
  static_assert(S <= 8, "at most 56 digit bits");
  const double a = fabs(y) * 268435456.0;
  const double h = trunc(a);
  const uint32_t hi = (uint32_t)h;
  const uint32_t lo = S > 4 ? (uint32_t)trunc((a - h) * 268435456.0) : 0u;
  const bool neg = y < 0.0;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const uint32_t v = s < 4 ? (hi >> (21 - 7 * s)) & 127u : (lo >> (21 - 7 * (s - 4))) & 127u;
    d[s] = (int8_t)(neg ? -(int)v : (int)v);
  }
}

// byte offset of the 16-byte digit group of (row r, k .. k + 15), plane s
// (k a multiple of 16)
template <int S>
__host__ __device__ __forceinline__ int64_t fm_oz_chunk(int k, int s, int rows_pad, int r) {
  return ((int64_t)((k >> 5) * 2 * S + 2 * s + ((k >> 4) & 1)) * rows_pad + r) * 16;
}

// ---- A side: one work group per 64 rows of one operand ---------------------
// Row exponents to ws + a_exp (int32), digits to ws + a_slc.
template <int S>
__global__ void __launch_bounds__(256) fm_ozaki_prep_rows(FmOzakiArgs args) {
  const FmOzakiDesc d = args.descs[blockIdx.y];
  const int r0 = blockIdx.x * 64;
  if (r0 >= d.M) return;
  const fm_gptr<const double> A = fm_global(reinterpret_cast<const double*>(fm_base(args.bases, d.a.role)) + d.a.off);
  const int64_t lda = d.a.ld;
  const fm_gptr<char> ws = fm_global(reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]));
  const fm_gptr<int32_t> ex = (fm_gptr<int32_t>)(ws + d.a_exp);
  const fm_gptr<int8_t> sl = (fm_gptr<int8_t>)(ws + d.a_slc);
  const int tid = threadIdx.x, r = tid & 63, kq = tid >> 6;  // 4 threads per row
  const int row = r0 + r;
  const bool rok = row < d.M;
  __shared__ double red[4][64];
  double mx = 0.0;
  if (rok) {
#pragma unroll 8
    for (int k = kq; k < d.K; k += 4) mx = fmax(mx, fabs(A[row + (int64_t)k * lda]));
  }
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
      const int k = k0 + 16 * kq;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const fm_gptr<int8_t> p = sl + fm_oz_chunk<S>(k, s, d.rows_pad, row);
        *(fm_gptr<fm_i32x4>)(p) = *reinterpret_cast<const fm_i32x4*>(&dig[s][0]);
      }
    }
  }
}

// ---- B side: one work group per 64 columns of one operand ------------------
// Column exponents to ws + b_exp, digits to ws + b_slc.  The exponent of a
// column is a reduction along its K contiguous entries (one wave per column,
// coalesced).  Each 64-deep K chunk of the 64 columns is then loaded along
// K (one wave per column, 512-byte runs) into LDS and cut with the 64 lanes
// of a wave on 64 columns, each lane 16 consecutive k of its column, so the
// digit groups go out as consecutive 16-byte slots.
template <int S>
__global__ void __launch_bounds__(256) fm_ozaki_prep_cols(FmOzakiArgs args) {
  const FmOzakiDesc d = args.descs[blockIdx.y];
  const int c0 = blockIdx.x * 64;
  if (c0 >= d.cols_pad) return;
  const fm_gptr<const double> B = fm_global(reinterpret_cast<const double*>(fm_base(args.bases, d.b.role)) + d.b.off);
  const int64_t ldb = d.b.ld;
  const fm_gptr<char> ws = fm_global(reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]));
  const fm_gptr<int32_t> ex = (fm_gptr<int32_t>)(ws + d.b_exp);
  const fm_gptr<int8_t> sl = (fm_gptr<int8_t>)(ws + d.b_slc);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ int es[64];
  for (int q = 0; q < 16; ++q) {
    const int cl = wave * 16 + q, col = c0 + cl;
    double mx = 0.0;
    if (col < d.N) {
      const fm_gptr<const double> bc = B + (int64_t)col * ldb;
#pragma unroll 8
      for (int k = lane; k < d.K; k += 64) mx = fmax(mx, fabs(bc[k]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    const int e = fm_exp_bound(mx);
    if (lane == 0) {
      es[cl] = e;
      if (col < d.N) ex[col] = e;
    }
  }
  __shared__ double tile[64][65];  // [column][k], padded
  __syncthreads();                 // exponents published
  const int cl = tid & 63, kq = tid >> 6, col = c0 + cl;
  const int e = es[cl];
  for (int k0 = 0; k0 < d.Kp; k0 += 64) {
    if (k0 > 0) __syncthreads();  // previous chunk consumed
    for (int q = 0; q < 16; ++q) {
      const int c2 = wave * 16 + q, k = k0 + lane;
      tile[c2][lane] = (c0 + c2 < d.N && k < d.K) ? B[(int64_t)(c0 + c2) * ldb + k] : 0.0;
    }
    __syncthreads();
    const int kb = k0 + 16 * kq;
    int8_t dig[S][16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const double v = ldexp(tile[cl][16 * kq + q], -e);
      int8_t t[S];
      fm_digits<S>(v, t);
#pragma unroll
      for (int s = 0; s < S; ++s) dig[s][q] = t[s];
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const fm_gptr<int8_t> p = sl + fm_oz_chunk<S>(kb, s, d.cols_pad, col);
      *(fm_gptr<fm_i32x4>)(p) = *reinterpret_cast<const fm_i32x4*>(&dig[s][0]);
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
  const fm_gptr<char> ws = fm_global(reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]));
  const fm_gptr<const int8_t> sa = (fm_gptr<const int8_t>)(ws + d.a_slc);
  const fm_gptr<const int8_t> sb = (fm_gptr<const int8_t>)(ws + d.b_slc);
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
      const fm_i32x4 va = *(fm_gptr<const fm_i32x4>)(sa + fm_oz_chunk<S>(k0 + lc, s, d.rows_pad, m0 + lr));
      const fm_i32x4 vb = *(fm_gptr<const fm_i32x4>)(sb + fm_oz_chunk<S>(k0 + lc, s, d.cols_pad, n0 + lr));
      *reinterpret_cast<fm_i32x4*>(&As[s][lr][lc]) = va;
      *reinterpret_cast<fm_i32x4*>(&Bs[s][lr][lc]) = vb;
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

  const fm_gptr<const int32_t> ea = (fm_gptr<const int32_t>)(ws + d.a_exp);
  const fm_gptr<const int32_t> eb = (fm_gptr<const int32_t>)(ws + d.b_exp);
  const fm_gptr<double> C = fm_global(reinterpret_cast<double*>(fm_base(args.bases, d.c.role)) + d.c.off);
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
        const fm_gptr<double> p = C + row + (int64_t)col * ldc;
        if (d.beta != 0.0) r = __dadd_rn(r, __dmul_rn(d.beta, *p));
        *p = r;
      }
    }
  }
}

// ---- grouped digit product, 32 x 32 x 32 int8 MFMA -------------------------------
// Work group: 2 x 2 waves, each a 32 x 32 block computed by
// v_mfma_i32_32x32x32_i8 (64 x 64 output tile), two work groups per CU.
// K goes in stages of KS = 2 slices of 32: A's digit groups of a stage are
// staged through LDS (double-buffered, 2 x 32 KB for S = 8) in their global
// order (plane half j = 2 s + h, then row), so the loads, the LDS stores
// (lane l: row l) and the fragment reads (lane l: row l % 32, half l / 32)
// all move consecutive 16-byte slots.  Each wave loads its B fragments
// straight from global memory into registers (the two waves that share a
// column block both load them; L2 serves the second), which halves the LDS
// traffic per MFMA.  The products of a slice run by descending B digit t,
// and the register of beta_t takes the next slice's beta_t right after its
// last product, so every B load has most of a slice to land and B needs
// one register set; the A half-stage loaded at the start of a slice is
// stored to the free LDS buffer at its end.  One barrier per stage.  The
// tiles of a product go 8 tile rows at a time, so the 64 work groups that
// run together on one XCD share 512 rows of A and 512 columns of B in its
// L2.  The per-scale int32 sums are exact, so the result equals
// fm_ozaki_gemm's bit for bit.
typedef int fm_i32x16 __attribute__((ext_vector_type(16)));

template <int S>
__global__ void __launch_bounds__(256, 2) fm_ozaki_gemm32(FmOzakiArgs args) {
  constexpr int KS = 2;
  constexpr int BM = 64, J = 2 * S, JS = KS * J, GROUP = 8;
  constexpr int NCH = (J * BM + 255) / 256;  // per 32-deep slice
  __shared__ __align__(16) fm_i32x4 As[2][JS][BM];

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
  const int per_group = GROUP * d.tiles_n, grp = local / per_group;
  const int first = grp * GROUP, rows_g = min(GROUP, d.tiles_m - first);
  const int rem = local - grp * per_group;
  const int m0 = (first + rem % rows_g) * BM, n0 = (rem / rows_g) * 64;
  const fm_gptr<char> ws = fm_global(reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]));
  const fm_gptr<const fm_i32x4> sa = (fm_gptr<const fm_i32x4>)(ws + d.a_slc) + m0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int fr = lane & 31, fh = lane >> 5;
  const int64_t pa = d.rows_pad, pb = d.cols_pad;
  // this lane's B column (n0 + wn 32 + fr < cols_pad) and K half
  const fm_gptr<const fm_i32x4> sbl = (fm_gptr<const fm_i32x4>)(ws + d.b_slc) + n0 + wn * 32 + fr + fh * pb;

  fm_i32x16 acc[S];
#pragma unroll
  for (int u = 0; u < S; ++u)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[u][e] = 0;

  // A staging, one 32-deep slice h of stage kc at a time
  fm_i32x4 ra[NCH];
  auto gload = [&](int kc, int h) {
#pragma unroll
    for (int u = 0; u < NCH; ++u) {
      const int o = tid + u * 256, row = o & (BM - 1), j = o / BM;
      if ((J * BM) % 256 == 0 || o < J * BM) ra[u] = sa[(int64_t)(kc * JS + h * J + j) * pa + row];
    }
  };
  auto sstore = [&](int buf, int h) {
#pragma unroll
    for (int u = 0; u < NCH; ++u) {
      const int o = tid + u * 256, row = o & (BM - 1), j = o / BM;
      if ((J * BM) % 256 == 0 || o < J * BM) As[buf][h * J + j][row] = ra[u];
    }
  };

  const int nkc = d.Kp / (32 * KS);  // Kp is a multiple of 64
  fm_i32x4 bf[S];
  if (nkc > 0) {
#pragma unroll
    for (int h = 0; h < KS; ++h) {
      gload(0, h);
      sstore(0, h);
    }
#pragma unroll
    for (int t = 0; t < S; ++t) bf[t] = sbl[(int64_t)(2 * t) * pb];
  }
  __syncthreads();
  for (int kc = 0; kc < nkc; ++kc) {
    const int buf = kc & 1;
    // The last stage reloads and restores itself (in bounds, never read)
    // instead of branching: at a branch join the compiler would wait for
    // every outstanding load, the fresh A loads included, before the first
    // product.  Buffer buf ^ 1 is free: every wave finished reading it
    // before the previous barrier.
    const int kn = min(kc + 1, nkc - 1);
#pragma unroll
    for (int h = 0; h < KS; ++h) {
      gload(kn, h);
      // keep every load where it is written: the scheduler would sink them
      // towards their uses, which leaves them no time to land
      __builtin_amdgcn_sched_barrier(0);
      // B of the next 32-deep slice
      const int64_t nb = (int64_t)((h + 1 < KS ? kc * KS + h + 1 : kn * KS) * J) * pb;
      fm_i32x4 af[S];
#pragma unroll
      for (int s = 0; s < S; ++s) af[s] = As[buf][h * J + 2 * s + fh][wm * 32 + fr];
#pragma unroll
      for (int t = S - 1; t >= 0; --t) {
#pragma unroll
        for (int s = 0; s + t < S; ++s)
          acc[s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bf[t], af[s], acc[s + t], 0, 0, 0);
        bf[t] = sbl[nb + (int64_t)(2 * t) * pb];
        __builtin_amdgcn_sched_barrier(0);
      }
      sstore(buf ^ 1, h);
    }
    __syncthreads();
  }

  const fm_gptr<const int32_t> ea = (fm_gptr<const int32_t>)(ws + d.a_exp);
  const fm_gptr<const int32_t> eb = (fm_gptr<const int32_t>)(ws + d.b_exp);
  const fm_gptr<double> C = fm_global(reinterpret_cast<double*>(fm_base(args.bases, d.c.role)) + d.c.off);
  const int64_t ldc = d.c.ld;
  const int row = m0 + wm * 32 + fr;
  if (row >= d.M) return;
  const int er = ea[row];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int col = n0 + wn * 32 + (e & 3) + 4 * fh + 8 * (e >> 2);
    if (col >= d.N) continue;
    double sum = 0.0;
#pragma unroll
    for (int u = S - 1; u >= 0; --u) sum = __dadd_rn(sum, ldexp((double)acc[u][e], -7 * (u + 2)));
    double r = __dmul_rn(ldexp(sum, er + eb[col]), d.alpha);
    const fm_gptr<double> p = C + row + (int64_t)col * ldc;
    if (d.beta != 0.0) r = __dadd_rn(r, __dmul_rn(d.beta, *p));
    *p = r;
  }
}

// ---- grouped digit product, 64 x 128 tiles -------------------------------------
// fm_ozaki_gemm32 with each wave on a 32 x 64 block: two B fragment columns
// share every A fragment, so a wave issues 72 MFMAs per 32-deep slice for
// 8 A and 16 B fragments (fm_ozaki_gemm32: 36 for 8 and 8).  Global and L2
// traffic per MFMA drops to a little over half; the 2 x 8 accumulator sets
// (256 registers per lane) leave one work group of four waves per CU.  B
// digit planes are padded to 128 columns for it.  Same per-scale int32
// sums and FP64 reconstruction as fm_ozaki_gemm32, so the result is the
// same bit for bit.
template <int S>
__global__ void __launch_bounds__(256, 1) fm_ozaki_gemm_w(FmOzakiArgs args) {
  constexpr int KS = 2;
  constexpr int BM = 64, BN = 128, TN = 2, J = 2 * S, JS = KS * J, GROUP = 8;
  constexpr int NCH = (J * BM + 255) / 256;  // per 32-deep slice
  __shared__ __align__(16) fm_i32x4 As[2][JS][BM];

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
  const int per_group = GROUP * d.tiles_n, grp = local / per_group;
  const int first = grp * GROUP, rows_g = min(GROUP, d.tiles_m - first);
  const int rem = local - grp * per_group;
  const int m0 = (first + rem % rows_g) * BM, n0 = (rem / rows_g) * BN;
  const fm_gptr<char> ws = fm_global(reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]));
  const fm_gptr<const fm_i32x4> sa = (fm_gptr<const fm_i32x4>)(ws + d.a_slc) + m0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int fr = lane & 31, fh = lane >> 5;
  const int64_t pa = d.rows_pad, pb = d.cols_pad;
  // this lane's B columns (n0 + wn 64 + 32 j + fr < cols_pad) and K half
  const fm_gptr<const fm_i32x4> sbl =
      (fm_gptr<const fm_i32x4>)(ws + d.b_slc) + n0 + wn * 64 + fr + fh * pb;

  fm_i32x16 acc[TN][S];
#pragma unroll
  for (int j = 0; j < TN; ++j)
#pragma unroll
    for (int u = 0; u < S; ++u)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[j][u][e] = 0;

  fm_i32x4 ra[NCH];
  auto gload = [&](int kc, int h) {
#pragma unroll
    for (int u = 0; u < NCH; ++u) {
      const int o = tid + u * 256, row = o & (BM - 1), jj = o / BM;
      if ((J * BM) % 256 == 0 || o < J * BM) ra[u] = sa[(int64_t)(kc * JS + h * J + jj) * pa + row];
    }
  };
  auto sstore = [&](int buf, int h) {
#pragma unroll
    for (int u = 0; u < NCH; ++u) {
      const int o = tid + u * 256, row = o & (BM - 1), jj = o / BM;
      if ((J * BM) % 256 == 0 || o < J * BM) As[buf][h * J + jj][row] = ra[u];
    }
  };

  const int nkc = d.Kp / (32 * KS);  // Kp is a multiple of 64
  fm_i32x4 bf[TN][S];
  if (nkc > 0) {
#pragma unroll
    for (int h = 0; h < KS; ++h) {
      gload(0, h);
      sstore(0, h);
    }
#pragma unroll
    for (int t = 0; t < S; ++t)
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j][t] = sbl[(int64_t)(2 * t) * pb + 32 * j];
  }
  __syncthreads();
  for (int kc = 0; kc < nkc; ++kc) {
    const int buf = kc & 1;
    // as in fm_ozaki_gemm32: the last stage reloads itself rather than
    // branching, so no branch join makes the compiler wait for every load
    const int kn = min(kc + 1, nkc - 1);
#pragma unroll
    for (int h = 0; h < KS; ++h) {
      gload(kn, h);
      __builtin_amdgcn_sched_barrier(0);
      const int64_t nb = (int64_t)((h + 1 < KS ? kc * KS + h + 1 : kn * KS) * J) * pb;
      fm_i32x4 af[S];
#pragma unroll
      for (int s = 0; s < S; ++s) af[s] = As[buf][h * J + 2 * s + fh][wm * 32 + fr];
#pragma unroll
      for (int t = S - 1; t >= 0; --t) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
#pragma unroll
          for (int s = 0; s + t < S; ++s)
            acc[j][s + t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(bf[j][t], af[s], acc[j][s + t], 0, 0, 0);
          bf[j][t] = sbl[nb + (int64_t)(2 * t) * pb + 32 * j];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      sstore(buf ^ 1, h);
    }
    __syncthreads();
  }

  const fm_gptr<const int32_t> ea = (fm_gptr<const int32_t>)(ws + d.a_exp);
  const fm_gptr<const int32_t> eb = (fm_gptr<const int32_t>)(ws + d.b_exp);
  const fm_gptr<double> C = fm_global(reinterpret_cast<double*>(fm_base(args.bases, d.c.role)) + d.c.off);
  const int64_t ldc = d.c.ld;
  const int row = m0 + wm * 32 + fr;
  if (row >= d.M) return;
  const int er = ea[row];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int col = n0 + wn * 64 + 32 * j + (e & 3) + 4 * fh + 8 * (e >> 2);
      if (col >= d.N) continue;
      double sum = 0.0;
#pragma unroll
      for (int u = S - 1; u >= 0; --u) sum = __dadd_rn(sum, ldexp((double)acc[j][u][e], -7 * (u + 2)));
      double r = __dmul_rn(ldexp(sum, er + eb[col]), d.alpha);
      const fm_gptr<double> p = C + row + (int64_t)col * ldc;
      if (d.beta != 0.0) r = __dadd_rn(r, __dmul_rn(d.beta, *p));
      *p = r;
    }
  }
}

template <int S>
static hipError_t fm_ozaki_run(const FmOzakiArgs& a, int max_rows, int max_cols, int phases,
                               hipStream_t s) {
  if (phases & FM_PHASE_PREP) {
    hipLaunchKernelGGL((fm_ozaki_prep_rows<S>), dim3((max_rows + 63) / 64, a.ndesc), dim3(256), 0,
                       s, a);
    hipLaunchKernelGGL((fm_ozaki_prep_cols<S>), dim3((max_cols + 63) / 64, a.ndesc), dim3(256), 0,
                       s, a);
  }
  if (!(phases & FM_PHASE_PRODUCT)) return hipGetLastError();
  if (a.variant == 0)
    hipLaunchKernelGGL((fm_ozaki_gemm<S>), dim3(a.total_tiles), dim3(256), 0, s, a);
  else if (a.variant == 1)
    hipLaunchKernelGGL((fm_ozaki_gemm32<S>), dim3(a.total_tiles), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((fm_ozaki_gemm_w<S>), dim3(a.total_tiles), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t fm_ozaki_launch(int slices, const FmOzakiArgs& a, int max_rows, int max_cols,
                           int phases, hipStream_t s) {
  if (a.total_tiles <= 0) return hipSuccess;
  switch (slices) {
    case 6: return fm_ozaki_run<6>(a, max_rows, max_cols, phases, s);
    case 7: return fm_ozaki_run<7>(a, max_rows, max_cols, phases, s);
    case 8: return fm_ozaki_run<8>(a, max_rows, max_cols, phases, s);
    default: return hipErrorInvalidValue;
  }
}
