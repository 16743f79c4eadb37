// gemm_mfma.hip -- grouped leaf GEMM on CDNA4 matrix cores (FP64 and FP32).
//
// One launch runs every product of a recursion level (the BFS form of the
// reference's leaf dgemm calls, linalg/linalg.hpp:261-269 reached from
// algorithms/strassen.hpp:385-388) plus any peeling products
// (linalg/linalg.hpp:213-243).  Work-group w takes output tile w of the
// concatenated tile list; products are found by division (uniform shapes) or
// by binary search over tile_begin.
//
// Tile: BM x BN per work group of WM x WN waves, each wave a
// (BM/WM) x (BN/WN) block of 16 x 16 MFMA tiles, K staged through LDS in
// BK-deep slices, double-buffered with a register prefetch of the next slice
// (16-byte loads on interior tiles, predicated element loads on edges).
// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 are fed with the B tile as
// the instruction's first operand and the A tile as the second, so the
// instruction's output row index runs along C's columns and its lane index
// along C's rows: each store instruction writes 16 consecutive rows of 4
// columns, i.e. contiguous 128-byte runs of column-major C.
#include <hip/hip_runtime.h>

#include "gemm_mfma.hpp"
#include "gmem.hpp"

template <typename T>
struct FmMfma;

// out_row(g, v): output row held in accumulator item v by lane group g
// (g = lane / 16; the output column is lane % 16).
template <>
struct FmMfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ __forceinline__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int out_row(int g, int v) { return g + 4 * v; }
};

template <>
struct FmMfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ __forceinline__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ __forceinline__ static int out_row(int g, int v) { return 4 * g + v; }
};

// Map a work-group id to a tile id so that each of the 8 XCDs (which receive
// work groups round-robin) walks a contiguous range of tiles: neighbouring
// tiles share A/B panels in that XCD's L2.
constexpr int FM_TILE_GROUP = 8;

__device__ __forceinline__ int fm_xcd_remap(int wg, int nwg) {
  const int full = nwg & ~7;
  if (wg < full) return (wg & 7) * (full >> 3) + (wg >> 3);
  return wg;
}

// MINW: minimum waves per SIMD the register allocation must allow (2 with
// four-wave work groups lets two work groups share a CU, so one group's
// barrier wait is covered by the other's MFMAs).  UNR: unroll of the
// 4-deep K steps inside one LDS slice (less unrolling, fewer live operand
// registers).
// PIPE: software-pipelined main loop (see the loop below); PIPE == 2 also
// interleaves the next step's LDS reads with the current step's MFMAs
// (one read, then two MFMAs) through scheduling-group barriers; PIPE == 3
// is PIPE == 2 with the next slice's LDS stores and the slice after's
// global loads spread over the K steps of the current slice (FULL and EVEN
// launches; others run as PIPE == 2).  FULL: every product is a whole
// number of tiles and K slices with 16-byte aligned columns (the launch
// checks), so only the unpredicated 16-byte load path is compiled.  EVEN:
// every product has M and K multiples of the vector width and aligned
// columns, so every tile, edge tiles and a partial last K slice included,
// takes 16-byte loads.
template <typename T, int BM, int BN, int BK, int WM, int WN, int MINW = 1, int UNR = BK / 4,
          int PIPE = 0, bool FULL = false, bool EVEN = false>
__global__ void __launch_bounds__(64 * WM * WN, MINW) fm_gemm_kernel(FmGemmArgs args) {
  using Op = FmMfma<T>;
  using acc_t = typename Op::acc_t;
  constexpr int NT = 64 * WM * WN;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int TM = WTM / 16, TN = WTN / 16;
  constexpr int BKP = BK + 1;
  // scalar (edge) load mapping
  constexpr int A_PER = BM * BK / NT, B_PER = BK * BN / NT;
  constexpr int A_KSTEP = NT / BM;  // k columns covered by one pass of the A tile
  constexpr int B_NSTEP = NT / BK;  // n columns covered by one pass of the B tile
  // 16-byte (interior) load mapping: VL consecutive elements per load
  constexpr int VL = 16 / sizeof(T);
  constexpr bool CAN_VEC = A_PER % VL == 0 && B_PER % VL == 0 && BM % VL == 0 &&
                           BK % VL == 0 && (NT * VL) % BM == 0 && (NT * VL) % BK == 0;
  constexpr int AV_PER = CAN_VEC ? A_PER / VL : 1, BV_PER = CAN_VEC ? B_PER / VL : 1;
  constexpr int AV_KSTEP = CAN_VEC ? NT * VL / BM : 1;
  constexpr int BV_NSTEP = CAN_VEC ? NT * VL / BK : 1;
  static_assert(NT % BM == 0 && NT % BK == 0, "tile/thread mismatch");
  static_assert(A_PER * NT == BM * BK && B_PER * NT == BK * BN, "tile/thread mismatch");
  static_assert(TM >= 1 && TN >= 1 && BK % 4 == 0, "bad tile");
  typedef T vec_t __attribute__((ext_vector_type(VL)));
  // the spread loop (PIPE == 3 on FULL and EVEN launches); other PIPE == 3
  // launches run the PIPE == 2 loop
  constexpr bool SPREAD = PIPE == 3 && (FULL || EVEN);
  constexpr int PIPE_EFF = PIPE == 3 ? 2 : PIPE;
  // Operand pointers: global address space (global loads, which the LDS
  // wait counter does not track) for FP32 and for the spread loop, which
  // waits on each part's load separately; generic for the other FP64
  // loops.  Measured on MI355X at 8192^3: FP32 tiles gain 1-4% with global
  // loads; the FP64 PIPE == 2 128 x 128 two-groups-per-CU tile loses 4%
  // (70.8 -> 67.7 TFLOP/s), other FP64 tiles move within +-2%.
  constexpr bool GM = sizeof(T) == 4 || SPREAD;

  __shared__ __align__(16) T As[2][BK][BM];
  __shared__ __align__(16) T Bs[2][BN][BKP];

  // Persistent when the grid is smaller than the tile count: work group b
  // takes tiles b, b + grid, b + 2 grid, ...
  for (int tile = blockIdx.x; tile < args.total_tiles; tile += gridDim.x) {
  const int wg = fm_xcd_remap(tile, args.total_tiles);
  int g;
  if (args.uniform) {
    const int per = args.descs[0].tiles_m * args.descs[0].tiles_n;
    g = wg / per;
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
  // Tiles go in groups of FM_TILE_GROUP tile rows, all tile columns of a
  // group, rows fastest: the work groups that run together on one XCD share
  // a few A row panels and walk B's column panels once per group, so a tall
  // product reads A from HBM about once (row-major order would read it once
  // per tile column; column-major, once per tile column too when the
  // product is taller than the tiles in flight).
  const int per_group = FM_TILE_GROUP * d.tiles_n, grp = local / per_group;
  const int first = grp * FM_TILE_GROUP, rows_g = min(FM_TILE_GROUP, d.tiles_m - first);
  const int rem = local - grp * per_group;
  const int tmi = first + rem % rows_g, tni = rem / rows_g;
  const int m0 = tmi * BM, n0 = tni * BN;
  const int M = d.M, N = d.N, K = d.K;
  const fm_ptr<GM, const T> __restrict__ A =
      (fm_ptr<GM, const T>)(reinterpret_cast<const T*>(fm_base(args.bases, d.a.role)) + d.a.off);
  const fm_ptr<GM, const T> __restrict__ B =
      (fm_ptr<GM, const T>)(reinterpret_cast<const T*>(fm_base(args.bases, d.b.role)) + d.b.off);
  const fm_ptr<GM, T> __restrict__ C = (fm_ptr<GM, T>)(reinterpret_cast<T*>(fm_base(args.bases, d.c.role)) + d.c.off);
  const int64_t lda = d.a.ld, ldb = d.b.ld, ldc = d.c.ld;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;

  // Tiles of products whose A and B columns start 16-byte aligned (checked
  // per product: one misaligned product of a grouped launch does not slow
  // the others; args.vec covers the base pointers) take 16-byte loads for
  // every full K slice: unpredicated inside the product, and on its right
  // and bottom edges predicated per 16-byte vector (a B vector is VL
  // consecutive k of one column, an A vector VL consecutive rows, which lie
  // wholly inside or wholly outside the product when VL divides M).  Tiles
  // of other products, and partial K slices, take the predicated element
  // loads.  Out-of-range elements read as zero either way, so the MFMA
  // sequence and the results are the same.
  static_assert(!FULL || CAN_VEC, "full-tile variant needs the vector load mapping");
  const bool d_aligned = d.a.off % VL == 0 && d.a.ld % VL == 0 && d.b.off % VL == 0 && d.b.ld % VL == 0;
  const bool edge_free = FULL || ((m0 + BM <= M) && (n0 + BN <= N));
  const bool interior =
      FULL || EVEN || (CAN_VEC && args.vec && d_aligned && (m0 + BM <= M || M % VL == 0));

  // element-load coordinates
  const int a_row = tid % BM, a_k = tid / BM;
  const int b_k = tid % BK, b_col = tid / BK;
  const bool a_row_ok = (m0 + a_row) < M;
  const fm_ptr<GM, const T> a_ptr = A + (m0 + a_row) + (int64_t)a_k * lda;
  const fm_ptr<GM, const T> b_ptr = B + b_k + (int64_t)(n0 + b_col) * ldb;
  // vector-load coordinates
  const int av_row = (tid % (BM / VL)) * VL, av_k = tid / (BM / VL);
  const int bv_k = (tid % (BK / VL)) * VL, bv_col = tid / (BK / VL);
  const fm_ptr<GM, const T> av_ptr = A + (m0 + av_row) + (int64_t)av_k * lda;
  const fm_ptr<GM, const T> bv_ptr = B + bv_k + (int64_t)(n0 + bv_col) * ldb;

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  T ra[A_PER], rb[B_PER];

  const bool av_in = m0 + av_row < M;  // this lane's A vector rows lie inside the product
  auto gload_vec = [&](int k0) {
    if constexpr (!CAN_VEC) return;
    if (edge_free) {
#pragma unroll
      for (int i = 0; i < AV_PER; ++i) {
        const vec_t v = *(const fm_ptr<GM, const vec_t>)(av_ptr + (int64_t)(k0 + i * AV_KSTEP) * lda);
#pragma unroll
        for (int e = 0; e < VL; ++e) ra[i * VL + e] = v[e];
      }
#pragma unroll
      for (int i = 0; i < BV_PER; ++i) {
        const vec_t v = *(const fm_ptr<GM, const vec_t>)(bv_ptr + k0 + (int64_t)(i * BV_NSTEP) * ldb);
#pragma unroll
        for (int e = 0; e < VL; ++e) rb[i * VL + e] = v[e];
      }
    } else {
#pragma unroll
      for (int i = 0; i < AV_PER; ++i) {
        vec_t v = vec_t{};
        if (av_in) v = *(const fm_ptr<GM, const vec_t>)(av_ptr + (int64_t)(k0 + i * AV_KSTEP) * lda);
#pragma unroll
        for (int e = 0; e < VL; ++e) ra[i * VL + e] = v[e];
      }
#pragma unroll
      for (int i = 0; i < BV_PER; ++i) {
        vec_t v = vec_t{};
        if (n0 + bv_col + i * BV_NSTEP < N)
          v = *(const fm_ptr<GM, const vec_t>)(bv_ptr + k0 + (int64_t)(i * BV_NSTEP) * ldb);
#pragma unroll
        for (int e = 0; e < VL; ++e) rb[i * VL + e] = v[e];
      }
    }
  };
  auto sstore_vec = [&](int buf) {
    if constexpr (!CAN_VEC) return;
#pragma unroll
    for (int i = 0; i < AV_PER; ++i) {
      vec_t v;
#pragma unroll
      for (int e = 0; e < VL; ++e) v[e] = ra[i * VL + e];
      *reinterpret_cast<vec_t*>(&As[buf][av_k + i * AV_KSTEP][av_row]) = v;
    }
#pragma unroll
    for (int i = 0; i < BV_PER; ++i)
#pragma unroll
      for (int e = 0; e < VL; ++e) Bs[buf][bv_col + i * BV_NSTEP][bv_k + e] = rb[i * VL + e];
  };
  auto gload = [&](int k0) {
    if constexpr (FULL) return;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) {
      const int gk = k0 + a_k + i * A_KSTEP;
      ra[i] = (a_row_ok && gk < K) ? a_ptr[(int64_t)(k0 + i * A_KSTEP) * lda] : T(0);
    }
#pragma unroll
    for (int i = 0; i < B_PER; ++i) {
      const int gk = k0 + b_k;
      const int gn = n0 + b_col + i * B_NSTEP;
      rb[i] = (gk < K && gn < N) ? b_ptr[k0 + (int64_t)(i * B_NSTEP) * ldb] : T(0);
    }
  };
  auto sstore = [&](int buf) {
    if constexpr (FULL) return;
#pragma unroll
    for (int i = 0; i < A_PER; ++i) As[buf][a_k + i * A_KSTEP][a_row] = ra[i];
#pragma unroll
    for (int i = 0; i < B_PER; ++i) Bs[buf][b_col + i * B_NSTEP][b_k] = rb[i];
  };

  const int nk = (K + BK - 1) / BK;
  // (FULL: K is a positive multiple of BK; the constant keeps a tile's
  // epilogue stores out of the next tile's main loop, where the compiler
  // would otherwise have to wait for every load and store in flight)
  bool vec_stage = FULL || EVEN || (interior && BK <= K);
  if constexpr (!SPREAD) {  // the spread loop stages its first slice by parts
    if (vec_stage) {
      gload_vec(0);
      sstore_vec(0);
    } else {
      gload(0);
      sstore(0);
    }
    __syncthreads();
  }
  if constexpr (SPREAD) {
    // The PIPE == 2 loop, with the LDS stores of slice kt + 1 and the global
    // loads of slice kt + 2 cut in NS4 parts, one per K step of slice kt
    // (part q: the 16-byte vectors i with i % NS4 == q).  Buffer cur ^ 1 is
    // free during slice kt: its last fragment reads came before the
    // previous barrier.  Each part's load has NS4 steps to land before its
    // store, and the barrier then waits for one part's stores only.
    constexpr int NS4 = BK / 4;
    T af[2][TM], bf[2][TN];
    auto frag = [&](int buf, int k4, int set) {
      const int kidx = k4 * 4 + (lane >> 4);
#pragma unroll
      for (int i = 0; i < TM; ++i) af[set][i] = As[buf][kidx][wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[set][j] = Bs[buf][wn * WTN + j * 16 + (lane & 15)][kidx];
    };
    auto mma_step = [&](int set) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = Op::mma(bf[set][j], af[set][i], acc[i][j]);
    };
    // Edge tiles and a partial last K slice (EVEN launches: K a multiple
    // of the vector width, so a B vector's VL k lie wholly inside or
    // outside): a vector outside the product loads from the product's first
    // row, column or k instead, in bounds, so every lane loads
    // unconditionally (no branch for the compiler's wait counts to merge
    // over), and is zeroed when stored to LDS (zeroing it at the load would
    // make the compiler wait for the load right there).
    const bool a_ok = edge_free || av_in;
    const fm_ptr<GM, const T> av_src = a_ok ? av_ptr : A + (int64_t)av_k * lda;
    auto b_ok = [&](int i) { return edge_free || n0 + bv_col + i * BV_NSTEP < N; };
    auto ak_ok = [&](int k0, int i) { return FULL || k0 + av_k + i * AV_KSTEP < K; };
    auto bk_ok = [&](int k0) { return FULL || k0 + bv_k < K; };
    // store part q of the slice at k0 (loaded by load_part(k0, q))
    auto store_part = [&](int buf, int q, int k0) {
#pragma unroll
      for (int i = 0; i < AV_PER; ++i)
        if (i % NS4 == q) {
          vec_t v;
#pragma unroll
          for (int e = 0; e < VL; ++e) v[e] = ra[i * VL + e];
          if constexpr (!FULL) v = a_ok && ak_ok(k0, i) ? v : vec_t{};
          *reinterpret_cast<vec_t*>(&As[buf][av_k + i * AV_KSTEP][av_row]) = v;
        }
#pragma unroll
      for (int i = 0; i < BV_PER; ++i)
        if (i % NS4 == q) {
          const bool ok = FULL || (b_ok(i) && bk_ok(k0));
#pragma unroll
          for (int e = 0; e < VL; ++e)
            Bs[buf][bv_col + i * BV_NSTEP][bv_k + e] = ok ? rb[i * VL + e] : T(0);
        }
    };
    auto load_part = [&](int k0, int q) {
#pragma unroll
      for (int i = 0; i < AV_PER; ++i)
        if (i % NS4 == q) {
          const int64_t kc = ak_ok(k0, i) ? (int64_t)(k0 + i * AV_KSTEP) : -(int64_t)av_k;
          const vec_t v = *(const fm_ptr<GM, const vec_t>)(av_src + kc * lda);
#pragma unroll
          for (int e = 0; e < VL; ++e) ra[i * VL + e] = v[e];
        }
#pragma unroll
      for (int i = 0; i < BV_PER; ++i)
        if (i % NS4 == q) {
          const int64_t col = FULL || b_ok(i) ? (int64_t)(i * BV_NSTEP) : -(int64_t)(n0 + bv_col);
          const int kr = bk_ok(k0) ? k0 : -bv_k;
          const vec_t v = *(const fm_ptr<GM, const vec_t>)(bv_ptr + kr + col * ldb);
#pragma unroll
          for (int e = 0; e < VL; ++e) rb[i * VL + e] = v[e];
        }
    };
    // first slice
#pragma unroll
    for (int q = 0; q < NS4; ++q) {
      load_part(0, q);
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int q = 0; q < NS4; ++q) store_part(0, q, 0);
    __syncthreads();
    // The stores and loads run unconditionally (the last slices store into
    // a buffer no one reads again and reload slice nk - 1, in bounds):
    // under a branch the compiler would wait for every load in flight
    // before each store instead of the one that store needs.
    // (in the loop's part order: the compiler's wait counts at the loop
    // head hold for both ways in only when the loads in flight are ordered
    // alike)
#pragma unroll
    for (int q = 0; q < NS4; ++q) {
      load_part(min(1, nk - 1) * BK, q);
      __builtin_amdgcn_sched_barrier(0);
    }
    frag(0, 0, 0);
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      const int kl = min(kt + 2, nk - 1) * BK;
#pragma unroll
      for (int q = 0; q < NS4; ++q) {
        store_part(cur ^ 1, q, (kt + 1) * BK);
        load_part(kl, q);
        __builtin_amdgcn_sched_barrier(0);
        if (q < NS4 - 1) {
          frag(cur, q + 1, (q + 1) & 1);
          mma_step(q & 1);
#pragma unroll
          for (int r = 0; r < TM + TN; ++r) {
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
            __builtin_amdgcn_sched_group_barrier(0x008, (TM * TN) / (TM + TN), 0);  // MFMAs
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      __syncthreads();
      frag(cur ^ 1, 0, NS4 & 1);
      __builtin_amdgcn_sched_barrier(0);
      mma_step((NS4 - 1) & 1);
    }
  } else if constexpr (PIPE_EFF) {
    // Fragments of K step s + 1 are read from LDS before the MFMAs of step s
    // (two register sets), and the MFMAs of a slice's last step are issued
    // after the barrier that publishes the next slice, behind that slice's
    // first fragment reads: no MFMA waits on an LDS read or a barrier unless
    // the reads take longer than a step's MFMAs.
    constexpr int NS4 = BK / 4;
    T af[2][TM], bf[2][TN];
    auto frag = [&](int buf, int k4, int set) {
      const int kidx = k4 * 4 + (lane >> 4);
#pragma unroll
      for (int i = 0; i < TM; ++i) af[set][i] = As[buf][kidx][wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[set][j] = Bs[buf][wn * WTN + j * 16 + (lane & 15)][kidx];
    };
    auto mma_step = [&](int set) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = Op::mma(bf[set][j], af[set][i], acc[i][j]);
    };
    if (nk > 1) {
      if (interior && 2 * BK <= K)
        gload_vec(BK);
      else
        gload(BK);
    }
    frag(0, 0, 0);
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
#pragma unroll
      for (int k4 = 0; k4 < NS4 - 1; ++k4) {
        frag(cur, k4 + 1, (k4 + 1) & 1);
        mma_step(k4 & 1);
        if constexpr (PIPE_EFF == 2) {
#pragma unroll
          for (int r = 0; r < TM + TN; ++r) {
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
            __builtin_amdgcn_sched_group_barrier(0x008, (TM * TN) / (TM + TN), 0);  // MFMAs
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      // the last step's fragments are in registers: publish the next slice
      if (kt + 1 < nk) {
        if (interior && (kt + 2) * BK <= K)
          sstore_vec(cur ^ 1);
        else
          sstore(cur ^ 1);
      }
      __syncthreads();
      if (kt + 2 < nk) {
        if (interior && (kt + 3) * BK <= K)
          gload_vec((kt + 2) * BK);
        else
          gload((kt + 2) * BK);
      }
      if (kt + 1 < nk) frag(cur ^ 1, 0, NS4 & 1);
      __builtin_amdgcn_sched_barrier(0);
      mma_step((NS4 - 1) & 1);
    }
  } else
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    vec_stage = interior && (kt + 2) * BK <= K;
    if (more) {
      if (vec_stage)
        gload_vec((kt + 1) * BK);
      else
        gload((kt + 1) * BK);
    }
#pragma unroll UNR
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      const int kidx = k4 * 4 + (lane >> 4);
      T af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[cur][kidx][wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[cur][wn * WTN + j * 16 + (lane & 15)][kidx];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = Op::mma(bf[j], af[i], acc[i][j]);
    }
    if (more) {
      if (vec_stage)
        sstore_vec(cur ^ 1);
      else
        sstore(cur ^ 1);
    }
    __syncthreads();
  }

  const T alpha = (T)d.alpha, beta = (T)d.beta;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int row = m0 + wm * WTM + i * 16 + (lane & 15);
    if (row >= M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = n0 + wn * WTN + j * 16 + Op::out_row(lane >> 4, v);
        if (col < N) {
          const fm_ptr<GM, T> p = C + row + (int64_t)col * ldc;
          T r = alpha * acc[i][j][v];
          if (beta != T(0)) r += beta * *p;
          *p = r;
        }
      }
    }
  }
  }  // tile loop
}

template <typename T, int BM, int BN, int BK, int WM, int WN, int MINW = 1, int UNR = BK / 4,
          int PIPE = 0>
static void fm_gemm_go(const FmGemmArgs& a, hipStream_t s) {
  const int grid = a.grid > 0 && a.grid < a.total_tiles ? a.grid : a.total_tiles;
  constexpr int NT = 64 * WM * WN, VL = 16 / sizeof(T);
  constexpr bool can_full = (BM * BK / NT) % VL == 0 && (BK * BN / NT) % VL == 0 &&
                            BM % VL == 0 && BK % VL == 0 && (NT * VL) % BM == 0 &&
                            (NT * VL) % BK == 0;
  if constexpr (can_full) {
    if (a.full && a.vec) {
      hipLaunchKernelGGL((fm_gemm_kernel<T, BM, BN, BK, WM, WN, MINW, UNR, PIPE, true>),
                         dim3(grid), dim3(NT), 0, s, a);
      return;
    }
    if constexpr (PIPE == 3) {
      if (a.even && a.vec) {
        hipLaunchKernelGGL((fm_gemm_kernel<T, BM, BN, BK, WM, WN, MINW, UNR, PIPE, false, true>),
                           dim3(grid), dim3(NT), 0, s, a);
        return;
      }
    }
  }
  hipLaunchKernelGGL((fm_gemm_kernel<T, BM, BN, BK, WM, WN, MINW, UNR, PIPE, false>), dim3(grid),
                     dim3(NT), 0, s, a);
}

template <typename T>
static hipError_t fm_gemm_launch(int cfg, const FmGemmArgs& a, hipStream_t s) {
  if (a.total_tiles <= 0) return hipSuccess;
  switch (cfg) {
    case FM_TILE_128x128: fm_gemm_go<T, 128, 128, 16, 2, 2>(a, s); break;
    case FM_TILE_64x64: fm_gemm_go<T, 64, 64, 16, 2, 2>(a, s); break;
    case FM_TILE_128x64: fm_gemm_go<T, 128, 64, 16, 2, 2>(a, s); break;
    case FM_TILE_64x128: fm_gemm_go<T, 64, 128, 16, 2, 2>(a, s); break;
    case FM_TILE_128x128_K32: fm_gemm_go<T, 128, 128, 32, 2, 2>(a, s); break;
    case FM_TILE_128x128_W8: fm_gemm_go<T, 128, 128, 16, 2, 4>(a, s); break;
    case FM_TILE_128x128_K32_W8: fm_gemm_go<T, 128, 128, 32, 2, 4>(a, s); break;
    case FM_TILE_256x128_W8: fm_gemm_go<T, 256, 128, 16, 4, 2>(a, s); break;
    case FM_TILE_128x256_W8: fm_gemm_go<T, 128, 256, 16, 2, 4>(a, s); break;
    case FM_TILE_256x256_W16: fm_gemm_go<T, 256, 256, 16, 4, 4>(a, s); break;
    case FM_TILE_64x64_K32: fm_gemm_go<T, 64, 64, 32, 2, 2>(a, s); break;
    case FM_TILE_128x128_W16: fm_gemm_go<T, 128, 128, 16, 4, 4>(a, s); break;
    case FM_TILE_128x64_W8: fm_gemm_go<T, 128, 64, 16, 2, 4>(a, s); break;
    case FM_TILE_64x128_W8: fm_gemm_go<T, 64, 128, 16, 4, 2>(a, s); break;
    case FM_TILE_128x128_O2: fm_gemm_go<T, 128, 128, 16, 2, 2, 2>(a, s); break;
    case FM_TILE_256x128_W8_U1: fm_gemm_go<T, 256, 128, 16, 4, 2, 1, 1>(a, s); break;
    case FM_TILE_128x128_O2_U1: fm_gemm_go<T, 128, 128, 16, 2, 2, 2, 1>(a, s); break;
    case FM_TILE_128x256_W8_U1: fm_gemm_go<T, 128, 256, 16, 2, 4, 1, 1>(a, s); break;
    case FM_TILE_PP_256x128_W8: fm_gemm_go<T, 256, 128, 16, 4, 2, 1, 4, 1>(a, s); break;
    case FM_TILE_PP_128x256_W8: fm_gemm_go<T, 128, 256, 16, 2, 4, 1, 4, 1>(a, s); break;
    case FM_TILE_PP_128x128_W8: fm_gemm_go<T, 128, 128, 16, 2, 4, 1, 4, 1>(a, s); break;
    case FM_TILE_PP_128x128_O2: fm_gemm_go<T, 128, 128, 16, 2, 2, 2, 4, 1>(a, s); break;
    case FM_TILE_PI_128x128_O2: fm_gemm_go<T, 128, 128, 16, 2, 2, 2, 4, 2>(a, s); break;
    case FM_TILE_PI_128x256_W8: fm_gemm_go<T, 128, 256, 16, 2, 4, 1, 4, 2>(a, s); break;
    case FM_TILE_PI_128x96_O2: fm_gemm_go<T, 128, 96, 16, 2, 2, 2, 4, 2>(a, s); break;
    case FM_TILE_PS_128x128_O2: fm_gemm_go<T, 128, 128, 16, 2, 2, 2, 4, 3>(a, s); break;
    case FM_TILE_PS_128x256_W8: fm_gemm_go<T, 128, 256, 16, 2, 4, 1, 4, 3>(a, s); break;
    case FM_TILE_PS_128x96_O2: fm_gemm_go<T, 128, 96, 16, 2, 2, 2, 4, 3>(a, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t fm_gemm_launch_f64(int cfg, const FmGemmArgs& a, hipStream_t s) {
  return fm_gemm_launch<double>(cfg, a, s);
}

hipError_t fm_gemm_launch_f32(int cfg, const FmGemmArgs& a, hipStream_t s) {
  return fm_gemm_launch<float>(cfg, a, s);
}
