// gemm_split.hip -- FP32-accurate leaf products from split bfloat16 slices.
//
// Each FP32 operand value x is split on its way into LDS into three bf16
// slices x = x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1);
// every difference is exact in FP32, and the three 8-bit significands carry
// the 24 bits of x).  A B is then the sum of the slice products A_i B_j with
// i + j <= 2 (six v_mfma_f32_16x16x32_bf16 per output fragment and K step,
// smallest terms first), accumulated in FP32: the dropped terms are below
// 2^-24 relative to |A||B|, the accumulation error is that of an FP32 GEMM.
// The leaf contract is the reference's (linalg/linalg.hpp:261-269):
// C = alpha * A B + beta * C, column-major, grouped over many products.
//
// LDS holds each slice plane with K contiguous per row of A and per column
// of B (rows padded to 40 bf16 = 80 B so eight lanes' 16-byte reads hit
// distinct banks).  One LDS buffer, register prefetch of the next K slice,
// two barriers per slice.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gemm_mfma.hpp"
#include "gmem.hpp"
#include "split_bf16.hpp"

typedef __bf16 fm_bf16x8 __attribute__((ext_vector_type(8)));
typedef float fm_f32x4 __attribute__((ext_vector_type(4)));
// native 16-byte vector (register-resident in arrays, unlike the uint4 struct)
typedef uint32_t fm_u32x4 __attribute__((ext_vector_type(4)));

// split 8 values into three slice planes (8 bf16 each, packed in 4 uints);
// fm_split1 (split_bf16.hpp) per value
__device__ __forceinline__ void fm_split3(const float (&x)[8], uint4 (&out)[3]) {
  uint32_t h[8], m[8], l[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) fm_split1(x[e], h[e], m[e], l[e]);
  out[0] = make_uint4(h[0] | h[1] << 16, h[2] | h[3] << 16, h[4] | h[5] << 16, h[6] | h[7] << 16);
  out[1] = make_uint4(m[0] | m[1] << 16, m[2] | m[3] << 16, m[4] | m[5] << 16, m[6] | m[7] << 16);
  out[2] = make_uint4(l[0] | l[1] << 16, l[2] | l[3] << 16, l[4] | l[5] << 16, l[6] | l[7] << 16);
}

__device__ __forceinline__ int fm_xcd_remap2(int wg, int nwg) {
  const int full = nwg & ~7;
  if (wg < full) return (wg & 7) * (full >> 3) + (wg >> 3);
  return wg;
}

template <int BM, int BN, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN) fm_gemm_bf16x3_kernel(FmGemmArgs args) {
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int BK = 32, KP = BK + 8;            // bf16 per padded LDS row
  constexpr int A_OCT = BM * BK / 8 / NT;         // 8-element octets per thread
  constexpr int B_OCT = BN * BK / 8 / NT;
  static_assert(A_OCT * NT * 8 == BM * BK && B_OCT * NT * 8 == BN * BK, "tile/thread mismatch");
  static_assert(TM >= 1 && TN >= 1, "bad tile");

  __shared__ __align__(16) uint16_t As[3][BM][KP];
  __shared__ __align__(16) uint16_t Bs[3][BN][KP];

  // persistent when the grid is smaller than the tile count
  for (int tile = blockIdx.x; tile < args.total_tiles; tile += gridDim.x) {
  const int wg = fm_xcd_remap2(tile, args.total_tiles);
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
  const fm_gptr<const float> __restrict__ A =
      fm_global(reinterpret_cast<const float*>(fm_base(args.bases, d.a.role)) + d.a.off);
  const fm_gptr<const float> __restrict__ B =
      fm_global(reinterpret_cast<const float*>(fm_base(args.bases, d.b.role)) + d.b.off);
  const fm_gptr<float> __restrict__ C = fm_global(reinterpret_cast<float*>(fm_base(args.bases, d.c.role)) + d.c.off);
  const int64_t lda = d.a.ld, ldb = d.b.ld, ldc = d.c.ld;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;

  float ra[A_OCT][8], rb[B_OCT][8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int u = 0; u < A_OCT; ++u) {
      const int o = tid + u * NT;
      const int row = o % BM, kq = o / BM;
      const bool rok = m0 + row < M;
      const fm_gptr<const float> p = A + (m0 + row) + (int64_t)(k0 + 8 * kq) * lda;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        ra[u][e] = (rok && k0 + 8 * kq + e < K) ? p[(int64_t)e * lda] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < B_OCT; ++u) {
      const int o = tid + u * NT;
      const int kq = o % (BK / 8), col = o / (BK / 8);
      const bool cok = n0 + col < N;
      const fm_gptr<const float> p = B + (k0 + 8 * kq) + (int64_t)(n0 + col) * ldb;
#pragma unroll
      for (int e = 0; e < 8; ++e) rb[u][e] = (cok && k0 + 8 * kq + e < K) ? p[e] : 0.f;
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int u = 0; u < A_OCT; ++u) {
      const int o = tid + u * NT;
      const int row = o % BM, kq = o / BM;
      uint4 sl[3];
      fm_split3(ra[u], sl);
#pragma unroll
      for (int p = 0; p < 3; ++p) *reinterpret_cast<uint4*>(&As[p][row][8 * kq]) = sl[p];
    }
#pragma unroll
    for (int u = 0; u < B_OCT; ++u) {
      const int o = tid + u * NT;
      const int kq = o % (BK / 8), col = o / (BK / 8);
      uint4 sl[3];
      fm_split3(rb[u], sl);
#pragma unroll
      for (int p = 0; p < 3; ++p) *reinterpret_cast<uint4*>(&Bs[p][col][8 * kq]) = sl[p];
    }
  };

  fm_f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = fm_f32x4{0.f, 0.f, 0.f, 0.f};

  const int nk = (K + BK - 1) / BK;
  const int kofs = 8 * (lane >> 4);
  gload(0);
  for (int kt = 0; kt < nk; ++kt) {
    sstore();
    __syncthreads();
    if (kt + 1 < nk) gload((kt + 1) * BK);
    fm_bf16x8 af[3][TM], bf[3][TN];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
        af[p][i] = *reinterpret_cast<const fm_bf16x8*>(&As[p][wm * WTM + i * 16 + (lane & 15)][kofs]);
#pragma unroll
      for (int j = 0; j < TN; ++j)
        bf[p][j] = *reinterpret_cast<const fm_bf16x8*>(&Bs[p][wn * WTN + j * 16 + (lane & 15)][kofs]);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        // slice products A_a B_b with a + b <= 2, smallest first; the MFMA's
        // first operand is the B slice (rows = C columns), its second the A slice
        fm_f32x4 c = acc[i][j];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[2][j], af[0][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[1][j], af[1][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[2][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[1][j], af[0][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[1][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[0][i], c, 0, 0, 0);
        acc[i][j] = c;
      }
    __syncthreads();
  }

  const float alpha = (float)d.alpha, beta = (float)d.beta;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int row = m0 + wm * WTM + i * 16 + (lane & 15);
    if (row >= M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = n0 + wn * WTN + j * 16 + 4 * (lane >> 4) + v;
        if (col < N) {
          const fm_gptr<float> p = C + row + (int64_t)col * ldc;
          float r = alpha * acc[i][j][v];
          if (beta != 0.f) r += beta * *p;
          *p = r;
        }
      }
    }
  }
  }  // tile loop
}

hipError_t fm_gemm_launch_bf16x3(int cfg, const FmGemmArgs& a, hipStream_t s) {
  if (a.total_tiles <= 0) return hipSuccess;
  const int grid = a.grid > 0 && a.grid < a.total_tiles ? a.grid : a.total_tiles;
  switch (cfg) {
    case FM_TILE_64x64:
      hipLaunchKernelGGL((fm_gemm_bf16x3_kernel<64, 64, 2, 2>), dim3(grid), dim3(256), 0,
                         s, a);
      break;
    case FM_TILE_128x128_W8:
      hipLaunchKernelGGL((fm_gemm_bf16x3_kernel<128, 128, 2, 4>), dim3(grid),
                         dim3(512), 0, s, a);
      break;
    case FM_TILE_128x64:
      hipLaunchKernelGGL((fm_gemm_bf16x3_kernel<128, 64, 2, 2>), dim3(grid), dim3(256),
                         0, s, a);
      break;
    case FM_TILE_64x128:
      hipLaunchKernelGGL((fm_gemm_bf16x3_kernel<64, 128, 2, 2>), dim3(grid), dim3(256),
                         0, s, a);
      break;
    default:
      hipLaunchKernelGGL((fm_gemm_bf16x3_kernel<128, 128, 2, 2>), dim3(grid),
                         dim3(256), 0, s, a);
      break;
  }
  return hipGetLastError();
}

// ---- slice planes prepared once per operand -------------------------------------
// One work group cuts 64 rows x FM_PREP_KCH consecutive 32-deep K slices of
// one operand: per slice each thread 8 consecutive k of one row, the 4
// threads of a row in adjacent lanes, so a wave writes 16 consecutive
// 192-byte row chunks.  The launch is flat: work group b belongs to the
// operand whose wg_begin is the largest one <= b (binary search), so no
// work group of an uneven batch is empty.  Sources with K contiguous (B)
// load 128-byte runs along each row; sources with rows contiguous (A) go
// through LDS: each wave loads 256-byte runs down 64 rows of a column, and
// every thread then reads its row's 8 k across the staged slice.
__global__ void __launch_bounds__(256) fm_split_prep(FmSplitArgs args) {
  int lo = 0, hi = args.nprep - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (args.preps[mid].wg_begin <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const FmSplitPrep d = args.preps[lo];
  const int nkc = d.Kp >> 5, nkb = (nkc + FM_PREP_KCH - 1) / FM_PREP_KCH;
  const int local = b - d.wg_begin, rb = local / nkb;
  const int r0 = rb * 64, kc0 = (local - rb * nkb) * FM_PREP_KCH;
  const fm_gptr<const float> __restrict__ src =
      fm_global(reinterpret_cast<const float*>(fm_base(args.bases, d.src.role)) + d.src.off);
  const fm_gptr<char> ws = fm_global(reinterpret_cast<char*>(args.bases.p[FM_ROLE_WS]));
  const int tid = threadIdx.x;
  const int r = tid >> 2, q = tid & 3;
  const int row = r0 + r;
  const bool rok = row < d.rows;
  const int64_t ld = d.src.ld;
  const int kc1 = min(kc0 + FM_PREP_KCH, nkc);
  // [k][row] of one slice (A sources); rows padded to 66 words so that the
  // reads below (16 rows of 4 k groups per wave, 8 k apart) hit 64 banks
  __shared__ float stage[32][64 + 2];
  __shared__ fm_u32x4 ostage[64 * 12];  // the slice's planes, 64 rows x 192 bytes
  // A sources: thread t loads row t % 64 of columns t / 64 + 4 i
  const int lr = tid & 63, lk = tid >> 6;
  const bool lrok = r0 + lr < d.rows;
  for (int kc = kc0; kc < kc1; ++kc) {
    const int k = kc * 32 + 8 * q;
    float x[8];
    if (d.kcontig) {
      const fm_gptr<const float> p = src + (int64_t)row * ld + k;
#pragma unroll
      for (int e = 0; e < 8; ++e) x[e] = (rok && k + e < d.K) ? p[e] : 0.f;
    } else {
      const fm_gptr<const float> p = src + (r0 + lr) + (int64_t)(kc * 32 + lk) * ld;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int kk = lk + 4 * i;
        stage[kk][lr] = (lrok && kc * 32 + kk < d.K) ? p[(int64_t)(4 * i) * ld] : 0.f;
      }
      __syncthreads();
#pragma unroll
      for (int e = 0; e < 8; ++e) x[e] = stage[8 * q + e][r];
      __syncthreads();  // the next slice overwrites the stage
    }
    uint4 sl[3];
    fm_split3(x, sl);
    // the slice's 64 rows x 192 bytes are one contiguous run of the
    // planes: staged in LDS in that order, then written 16 bytes a lane
    // over consecutive addresses (each store instruction a 1 KB run)
#pragma unroll
    for (int p = 0; p < 3; ++p) ostage[r * 12 + p * 4 + q] = *reinterpret_cast<const fm_u32x4*>(&sl[p]);
    __syncthreads();
    const fm_gptr<fm_u32x4> dst = (fm_gptr<fm_u32x4>)(ws + d.dst + ((int64_t)kc * d.rows_pad + r0) * 192);
#pragma unroll
    for (int u = 0; u < 3; ++u) dst[tid + 256 * u] = ostage[tid + 256 * u];
    __syncthreads();  // the next slice overwrites the stage
  }
}

// ---- grouped products on slice planes -------------------------------------------
// K in 32-deep slices; each slice of a tile is BM (A) + BN (B) consecutive
// rows of 192 bytes, copied from the planes into LDS with 16-byte loads (no
// bounds checks: the planes are padded to whole tiles and K slices, their
// padded row counts tiles_m BM and tiles_n BN).  LDS is
// double-buffered: the loads of slice kt + 1 are in flight while slice kt
// is multiplied, then stored to the other buffer; one barrier per slice.
// With DB false the LDS holds one slice (two work groups fit on a CU): the
// next slice waits in registers and is stored after a second barrier.
// In LDS each plane row is 64 bytes (4 chunks of 8 bf16); chunk c of row r
// sits at position c ^ ((r >> 1) & 3).  Measured on MI355X (SQ_LDS_BANK_CONFLICT
// over the fragment read pattern: rows r..r+15 from a multiple of 16, one
// chunk per 16 lanes) this is the swizzle without bank conflicts: the plain
// layout, 80-byte padded rows and swizzles that also mix in row bit 3 all
// conflict and read at about half the rate.  The stores (4 chunks of 16
// consecutive rows per wave) are conflict-free with it too.
// The slice products and their order per output fragment are those of
// fm_gemm_bf16x3_kernel, so both kernels give the same bits.
__device__ __forceinline__ int fm_sp_swz(int r) { return (r >> 1) & 3; }

template <int BM, int BN, int WM, int WN, bool DB, bool IL = false>
__global__ void __launch_bounds__(64 * WM * WN) fm_split_gemm(FmSplitArgs args) {
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int A_CH = BM * 12 / NT, B_CH = BN * 12 / NT;  // 16-byte chunks per thread and slice
  static_assert(A_CH * NT == BM * 12 && B_CH * NT == BN * 12, "tile/thread mismatch");
  static_assert(TM >= 1 && TN >= 1, "bad tile");

  constexpr int NBUF = DB ? 2 : 1;
  __shared__ __align__(16) fm_u32x4 As[NBUF][3][BM][4];
  __shared__ __align__(16) fm_u32x4 Bs[NBUF][3][BN][4];

  for (int tile = blockIdx.x; tile < args.total_tiles; tile += gridDim.x) {
  const int wg = fm_xcd_remap2(tile, args.total_tiles);
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
  const FmSplitDesc d = args.descs[g];
  // tiles of a product in groups of SP_GROUP tile rows: consecutive tiles go
  // down the rows of a group, then across its columns, so the 32 to 64 work
  // groups that run together on one XCD cover a square of about 1024 x 1024
  // outputs and share its A and B panels in the XCD's L2
  constexpr int SP_GROUP = BM == 256 ? 4 : 8;
  const int local = wg - d.tile_begin;
  const int per_group = SP_GROUP * d.tiles_n, grp = local / per_group;
  const int first = grp * SP_GROUP, rows_g = min(SP_GROUP, d.tiles_m - first);
  const int rem = local - grp * per_group;
  const int m0 = (first + rem % rows_g) * BM, n0 = (rem / rows_g) * BN;
  const int nkc = d.Kp >> 5;
  const fm_gptr<const char> ws = fm_global(reinterpret_cast<const char*>(args.bases.p[FM_ROLE_WS]));
  // chunk o of a slice: plane o / (4 * rows), row (o / 4) % rows, chunk o % 4;
  // consecutive K slices of a plane set are apa (bpa) rows apart
  const int64_t apa = (int64_t)d.tiles_m * BM, bpa = (int64_t)d.tiles_n * BN;
  const fm_gptr<const fm_u32x4> Ap = (fm_gptr<const fm_u32x4>)(ws + d.a_pl) + (int64_t)m0 * 12;
  const fm_gptr<const fm_u32x4> Bp = (fm_gptr<const fm_u32x4>)(ws + d.b_pl) + (int64_t)n0 * 12;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;

  fm_u32x4 ra[A_CH], rb[B_CH];
  auto gload = [&](int kt) {
#pragma unroll
    for (int u = 0; u < A_CH; ++u) {
      const int o = tid + u * NT, c = o & 3, row = (o >> 2) % BM, p = (o >> 2) / BM;
      ra[u] = Ap[((int64_t)kt * apa + row) * 12 + p * 4 + c];
    }
#pragma unroll
    for (int u = 0; u < B_CH; ++u) {
      const int o = tid + u * NT, c = o & 3, col = (o >> 2) % BN, p = (o >> 2) / BN;
      rb[u] = Bp[((int64_t)kt * bpa + col) * 12 + p * 4 + c];
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int u = 0; u < A_CH; ++u) {
      const int o = tid + u * NT, c = o & 3, row = (o >> 2) % BM, p = (o >> 2) / BM;
      As[buf][p][row][c ^ fm_sp_swz(row)] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < B_CH; ++u) {
      const int o = tid + u * NT, c = o & 3, col = (o >> 2) % BN, p = (o >> 2) / BN;
      Bs[buf][p][col][c ^ fm_sp_swz(col)] = rb[u];
    }
  };

  fm_f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = fm_f32x4{0.f, 0.f, 0.f, 0.f};

  const int fr = lane & 15, fc = lane >> 4;
  const int fsw = fc ^ fm_sp_swz(fr);  // rows of a fragment start at multiples of 16
  if (nkc > 0) {
    gload(0);
    sstore(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nkc; ++kt) {
    const int buf = DB ? kt & 1 : 0;
    if (kt + 1 < nkc) gload(kt + 1);
    // B fragments and the first A row first; each further A row is read
    // while the previous row's MFMAs run (pinned with scheduling-group
    // barriers when IL: 15 reads, then per row 3 x (8 MFMAs, 1 read))
    fm_bf16x8 af[3][TM], bf[3][TN];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int j = 0; j < TN; ++j)
        bf[p][j] = *reinterpret_cast<const fm_bf16x8*>(&Bs[buf][p][wn * WTN + j * 16 + fr][fsw]);
#pragma unroll
    for (int p = 0; p < 3; ++p)
      af[p][0] = *reinterpret_cast<const fm_bf16x8*>(&As[buf][p][wm * WTM + fr][fsw]);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      if (i + 1 < TM) {
#pragma unroll
        for (int p = 0; p < 3; ++p)
          af[p][i + 1] =
              *reinterpret_cast<const fm_bf16x8*>(&As[buf][p][wm * WTM + (i + 1) * 16 + fr][fsw]);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        fm_f32x4 c = acc[i][j];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[2][j], af[0][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[1][j], af[1][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[2][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[1][j], af[0][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[1][i], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[0][i], c, 0, 0, 0);
        acc[i][j] = c;
      }
    }
    if constexpr (IL) {
      __builtin_amdgcn_sched_group_barrier(0x100, 3 * TN + 3, 0);  // B fragments, A row 0
#pragma unroll
      for (int i = 0; i + 1 < TM; ++i)
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          __builtin_amdgcn_sched_group_barrier(0x008, 2 * TN, 0);  // a third of row i's MFMAs
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);       // one read of row i + 1
        }
      __builtin_amdgcn_sched_group_barrier(0x008, 6 * TN, 0);      // the last row
      __builtin_amdgcn_sched_barrier(0);
    }
    if (DB) {
      if (kt + 1 < nkc) sstore(buf ^ 1);
      __syncthreads();
    } else if (kt + 1 < nkc) {
      __syncthreads();
      sstore(0);
      __syncthreads();
    }
  }

  const float alpha = (float)d.alpha, beta = (float)d.beta;
  const fm_gptr<float> __restrict__ C = fm_global(reinterpret_cast<float*>(fm_base(args.bases, d.c.role)) + d.c.off);
  const int64_t ldc = d.c.ld;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int row = m0 + wm * WTM + i * 16 + fr;
    if (row >= d.M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = n0 + wn * WTN + j * 16 + 4 * fc + v;
        if (col < d.N) {
          const fm_gptr<float> p = C + row + (int64_t)col * ldc;
          float r = alpha * acc[i][j][v];
          if (beta != 0.f) r += beta * *p;
          *p = r;
        }
      }
    }
  }
  }  // tile loop
}

// fm_split_gemm with the B fragments loaded by each wave straight from the
// planes into registers (the WM waves that share a column block each load
// it; L2 serves the repeats) and only A staged through LDS (double-
// buffered): less LDS traffic per MFMA.  The products of a slice run by
// column block j, and the registers of block j take the next slice's B
// fragments right after its last product, so every B load has most of a
// slice to land with one register set.  The slice products of each output
// fragment and their order are fm_split_gemm's, so the results are
// bit-identical.  KS 32-deep slices share one LDS stage and one barrier
// (KS = 2 only when every product has an even number of slices): the A
// slice h of the next stage loads at the start of slice h and is stored to
// the free buffer at its end.  One tile per work group (launched with one
// work group per tile).
template <int BM, int BN, int WM, int WN, int KS>
__global__ void __launch_bounds__(64 * WM * WN) fm_split_gemm_bd(FmSplitArgs args) {
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int A_CH = BM * 12 / NT;  // 16-byte chunks per thread and slice
  static_assert(A_CH * NT == BM * 12, "tile/thread mismatch");
  static_assert(TM >= 1 && TN >= 1, "bad tile");
  __shared__ __align__(16) fm_u32x4 As[2][KS][3][BM][4];

  const int wg = fm_xcd_remap2(blockIdx.x, args.total_tiles);
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
  const FmSplitDesc d = args.descs[g];
  constexpr int SP_GROUP = BM == 256 ? 4 : 8;
  const int local = wg - d.tile_begin;
  const int per_group = SP_GROUP * d.tiles_n, grp = local / per_group;
  const int first = grp * SP_GROUP, rows_g = min(SP_GROUP, d.tiles_m - first);
  const int rem = local - grp * per_group;
  const int m0 = (first + rem % rows_g) * BM, n0 = (rem / rows_g) * BN;
  const int nkc = d.Kp >> 5;
  const fm_gptr<const char> ws = fm_global(reinterpret_cast<const char*>(args.bases.p[FM_ROLE_WS]));
  const int64_t apa = (int64_t)d.tiles_m * BM, bpa = (int64_t)d.tiles_n * BN;
  const fm_gptr<const fm_u32x4> Ap = (fm_gptr<const fm_u32x4>)(ws + d.a_pl) + (int64_t)m0 * 12;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int fr = lane & 15, fc = lane >> 4;
  const int fsw = fc ^ fm_sp_swz(fr);
  // this lane's B fragment bytes: column n0 + wn WTN + 16 j + fr, chunk fc
  // of each plane (bf16 k = 8 fc .. 8 fc + 7 of the slice)
  const fm_gptr<const char> bl = ws + d.b_pl + (int64_t)(n0 + wn * WTN + fr) * 192 + fc * 16;

  fm_u32x4 ra[A_CH];
  auto gload = [&](int ks) {  // A of 32-deep slice ks
#pragma unroll
    for (int u = 0; u < A_CH; ++u) {
      const int o = tid + u * NT, c = o & 3, row = (o >> 2) % BM, p = (o >> 2) / BM;
      ra[u] = Ap[((int64_t)ks * apa + row) * 12 + p * 4 + c];
    }
  };
  auto sstore = [&](int buf, int h) {
#pragma unroll
    for (int u = 0; u < A_CH; ++u) {
      const int o = tid + u * NT, c = o & 3, row = (o >> 2) % BM, p = (o >> 2) / BM;
      As[buf][h][p][row][c ^ fm_sp_swz(row)] = ra[u];
    }
  };
  fm_bf16x8 bf[3][TN];
  auto bload = [&](int kt, int j) {
#pragma unroll
    for (int p = 0; p < 3; ++p)
      bf[p][j] = *(fm_gptr<const fm_bf16x8>)(bl + ((int64_t)kt * bpa + 16 * j) * 192 + p * 64);
  };

  fm_f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = fm_f32x4{0.f, 0.f, 0.f, 0.f};

  const int nst = nkc / KS;  // stages (nkc is a multiple of KS)
  if (nst > 0) {
#pragma unroll
    for (int h = 0; h < KS; ++h) {
      gload(h);
      sstore(0, h);
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) bload(0, j);
  }
  __syncthreads();
  for (int kt = 0; kt < nst; ++kt) {
    const int buf = kt & 1;
    // The last stage reloads and restores itself (in bounds, never read)
    // instead of branching: at a branch join the compiler would wait for
    // every outstanding load before the first product.  Buffer buf ^ 1 is
    // free: every wave finished reading it before the previous barrier.
    const int kn = min(kt + 1, nst - 1);
#pragma unroll
    for (int h = 0; h < KS; ++h) {
      gload(kn * KS + h);
      // keep every load where it is written: the scheduler would sink them
      // towards their uses, which leaves them no time to land
      __builtin_amdgcn_sched_barrier(0);
      // B of the next 32-deep slice
      const int nb = h + 1 < KS ? kt * KS + h + 1 : kn * KS;
      fm_bf16x8 af[3][TM];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int p = 0; p < 3; ++p)
          af[p][i] = *reinterpret_cast<const fm_bf16x8*>(&As[buf][h][p][wm * WTM + i * 16 + fr][fsw]);
#pragma unroll
      for (int j = 0; j < TN; ++j) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          fm_f32x4 c = acc[i][j];
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[2][j], af[0][i], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[1][j], af[1][i], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[2][i], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[1][j], af[0][i], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[1][i], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[0][j], af[0][i], c, 0, 0, 0);
          acc[i][j] = c;
        }
        bload(nb, j);
        __builtin_amdgcn_sched_barrier(0);
      }
      sstore(buf ^ 1, h);
    }
    __syncthreads();
  }

  const float alpha = (float)d.alpha, beta = (float)d.beta;
  const fm_gptr<float> __restrict__ C = fm_global(reinterpret_cast<float*>(fm_base(args.bases, d.c.role)) + d.c.off);
  const int64_t ldc = d.c.ld;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int row = m0 + wm * WTM + i * 16 + fr;
    if (row >= d.M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = n0 + wn * WTN + j * 16 + 4 * fc + v;
        if (col < d.N) {
          const fm_gptr<float> p = C + row + (int64_t)col * ldc;
          float r = alpha * acc[i][j][v];
          if (beta != 0.f) r += beta * *p;
          *p = r;
        }
      }
    }
  }
}

hipError_t fm_split_launch(int cfg, const FmSplitArgs& a, int phases, hipStream_t s) {
  if ((phases & FM_PHASE_PREP) && a.nprep > 0 && a.prep_wgs > 0) {
    hipLaunchKernelGGL(fm_split_prep, dim3(a.prep_wgs), dim3(256), 0, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (!(phases & FM_PHASE_PRODUCT) || a.total_tiles <= 0) return hipSuccess;
  const int grid = a.grid > 0 && a.grid < a.total_tiles ? a.grid : a.total_tiles;
  switch (cfg) {
    case FM_TILE_SP_128x256_BD:  // one work group per tile
      if (a.even_k)
        hipLaunchKernelGGL((fm_split_gemm_bd<128, 256, 2, 4, 2>), dim3(a.total_tiles), dim3(512), 0, s, a);
      else
        hipLaunchKernelGGL((fm_split_gemm_bd<128, 256, 2, 4, 1>), dim3(a.total_tiles), dim3(512), 0, s, a);
      break;
    case FM_TILE_SP_256x128:
      hipLaunchKernelGGL((fm_split_gemm<256, 128, 4, 2, true>), dim3(grid), dim3(512), 0, s, a);
      break;
    case FM_TILE_SP_128x256:
      hipLaunchKernelGGL((fm_split_gemm<128, 256, 2, 4, true>), dim3(grid), dim3(512), 0, s, a);
      break;
    case FM_TILE_SP_256x128_IL:
      hipLaunchKernelGGL((fm_split_gemm<256, 128, 4, 2, true, true>), dim3(grid), dim3(512), 0, s, a);
      break;
    case FM_TILE_SP_128x256_IL:
      hipLaunchKernelGGL((fm_split_gemm<128, 256, 2, 4, true, true>), dim3(grid), dim3(512), 0, s, a);
      break;
    case FM_TILE_SP_128x128_DB:
      hipLaunchKernelGGL((fm_split_gemm<128, 128, 2, 2, true>), dim3(grid), dim3(256), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((fm_split_gemm<128, 128, 2, 2, false>), dim3(grid), dim3(256), 0, s, a);
      break;
  }
  return hipGetLastError();
}
