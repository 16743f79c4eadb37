// stage_kernels.hpp -- HBM-bound linear-combination stages of one recursion
// level, shared by every generated algorithm (csrc/gen/alg_*.hip).
//
// One launch covers all nodes of a level (grid.y = node): the BFS form of the
// reference's per-product S_Add/T_Add/M_Add calls (algorithms/strassen.hpp:9-376).
// Each lane owns VEC consecutive rows of one column of a block position,
// loads that position from every block of the node's operand once, and the
// generated combine function produces every output from registers, so each
// source element is read from HBM once per level and each output written once.
//
// Arithmetic goes through fm_add/fm_sub/fm_mul/fm_neg, which never contract
// into FMA: the combination of terms is rounded exactly like the reference's
// separate multiply/add expressions (codegen/gen.py:216-233).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "alg_desc.hpp"
#include "gmem.hpp"
#include "split_bf16.hpp"

__device__ __forceinline__ double fm_add(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double fm_sub(double a, double b) { return __dadd_rn(a, -b); }
__device__ __forceinline__ double fm_mul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double fm_neg(double a) { return -a; }
__device__ __forceinline__ float fm_add(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float fm_sub(float a, float b) { return __fadd_rn(a, -b); }
__device__ __forceinline__ float fm_mul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float fm_neg(float a) { return -a; }

template <typename T, int VEC>
struct FmVec {
  T v[VEC];
};

template <typename T, int VEC>
__device__ __forceinline__ FmVec<T, VEC> fm_load(fm_gptr<const T> p) {
  FmVec<T, VEC> r;
  if constexpr (VEC == 1) {
    r.v[0] = *p;
  } else {
    typedef T vt __attribute__((ext_vector_type(VEC)));
    vt x = *(fm_gptr<const vt>)p;
#pragma unroll
    for (int i = 0; i < VEC; ++i) r.v[i] = x[i];
  }
  return r;
}

template <typename T, int VEC>
__device__ __forceinline__ void fm_store(fm_gptr<T> p, const FmVec<T, VEC>& r) {
  if constexpr (VEC == 1) {
    *p = r.v[0];
  } else {
    typedef T vt __attribute__((ext_vector_type(VEC)));
    vt x;
#pragma unroll
    for (int i = 0; i < VEC; ++i) x[i] = r.v[i];
    *(fm_gptr<vt>)p = x;
  }
}

constexpr int FM_STAGE_THREADS = 256;

// SIDE 0: A operand (m x k blocks, U/SX, outputs S); SIDE 1: B (k x n, V/TX, T).
template <class Cmb, typename T, int VEC, int SIDE>
__global__ void __launch_bounds__(FM_STAGE_THREADS) fm_in_stage(FmStageArgs a) {
  constexpr int BR = SIDE == 0 ? Cmb::m : Cmb::k;
  constexpr int BC = SIDE == 0 ? Cmb::k : Cmb::n;
  constexpr int NB = BR * BC;
  constexpr int NOUT = SIDE == 0 ? Cmb::NS : Cmb::NT;
  constexpr int NOUT_ARR = NOUT > 0 ? NOUT : 1;
  const int q = blockIdx.y;
  const FmNodeView nv = fm_node_view(a, q, NOUT, 0);
  const FmOp op = nv.op;
  const fm_gptr<const T> src = fm_global(reinterpret_cast<const T*>(fm_base(a.bases, op.role)) + op.off);
  const fm_gptr<T> dst = fm_global(reinterpret_cast<T*>(a.bases.p[FM_ROLE_WS]) + nv.out);
  const uint32_t rv = (uint32_t)nv.rows / VEC;
  const uint32_t total = rv * (uint32_t)nv.cols;
  const int64_t ld = op.ld, rows = nv.rows, cols = nv.cols, out_ld = nv.out_ld, slab = nv.out_slab;
  const double x = nv.x;
  for (uint32_t e = blockIdx.x * FM_STAGE_THREADS + threadIdx.x; e < total;
       e += gridDim.x * FM_STAGE_THREADS) {
    const uint32_t j = e / rv;
    const uint32_t i = (e - j * rv) * VEC;
    FmVec<T, VEC> v[NB];
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int bi = p / BC, bj = p % BC;
      v[p] = fm_load<T, VEC>(src + (int64_t)bi * rows + i + ((int64_t)bj * cols + j) * ld);
    }
    FmVec<T, VEC> o[NOUT_ARR];
#pragma unroll
    for (int l = 0; l < VEC; ++l) {
      T in[NB];
#pragma unroll
      for (int p = 0; p < NB; ++p) in[p] = v[p].v[l];
      T out[NOUT_ARR];
      if constexpr (SIDE == 0)
        Cmb::s_combine(in, out, x);
      else
        Cmb::t_combine(in, out, x);
#pragma unroll
      for (int t = 0; t < NOUT; ++t) o[t].v[l] = out[t];
    }
#pragma unroll
    for (int t = 0; t < NOUT; ++t)
      fm_store<T, VEC>(dst + (int64_t)t * slab + i + (int64_t)j * out_ld, o[t]);
  }
}

// C_ij (+)= sum_q W[ij, q] M_q for every node; M_q of node q at
// ws + out_off + (q * R + r) * out_slab.
template <class Cmb, typename T, int VEC, bool BETA>
__global__ void __launch_bounds__(FM_STAGE_THREADS) fm_w_stage(FmStageArgs a) {
  constexpr int R = Cmb::R;
  constexpr int MN = Cmb::m * Cmb::n;
  const int q = blockIdx.y;
  const FmNodeView nv = fm_node_view(a, q, R, 0);
  const FmOp op = nv.op;
  const fm_gptr<T> c = fm_global(reinterpret_cast<T*>(fm_base(a.bases, op.role)) + op.off);
  const fm_gptr<const T> mb = fm_global(reinterpret_cast<const T*>(a.bases.p[FM_ROLE_WS]) + nv.out);
  const uint32_t rv = (uint32_t)nv.rows / VEC;
  const uint32_t total = rv * (uint32_t)nv.cols;
  const int64_t ld = op.ld, rows = nv.rows, cols = nv.cols, out_ld = nv.out_ld, slab = nv.out_slab;
  const T beta = (T)nv.beta;
  const double x = nv.x;
  for (uint32_t e = blockIdx.x * FM_STAGE_THREADS + threadIdx.x; e < total;
       e += gridDim.x * FM_STAGE_THREADS) {
    const uint32_t j = e / rv;
    const uint32_t i = (e - j * rv) * VEC;
    FmVec<T, VEC> mv[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      mv[r] = fm_load<T, VEC>(mb + (int64_t)r * slab + i + (int64_t)j * out_ld);
    FmVec<T, VEC> o[MN];
#pragma unroll
    for (int l = 0; l < VEC; ++l) {
      T in[R];
#pragma unroll
      for (int r = 0; r < R; ++r) in[r] = mv[r].v[l];
      T out[MN];
      Cmb::w_combine(in, out, x);
#pragma unroll
      for (int t = 0; t < MN; ++t) o[t].v[l] = out[t];
    }
#pragma unroll
    for (int t = 0; t < MN; ++t) {
      const int bi = t / Cmb::n, bj = t % Cmb::n;
      const fm_gptr<T> p = c + (int64_t)bi * rows + i + ((int64_t)bj * cols + j) * ld;
      if constexpr (BETA) {
        FmVec<T, VEC> old = fm_load<T, VEC>(p);
#pragma unroll
        for (int l = 0; l < VEC; ++l) o[t].v[l] = fm_add(o[t].v[l], fm_mul(beta, old.v[l]));
      }
      fm_store<T, VEC>(p, o[t]);
    }
  }
}

// Compile-time loop: f(std::integral_constant<int, I>) for I = B .. E-1.
template <int B, int E, class F>
__device__ __forceinline__ void fm_static_for(F& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>());
    fm_static_for<B + 1, E>(f);
  }
}

// Two recursion levels of S (SIDE 0) or T (SIDE 1) in one pass.  A lane owns
// one element position of the level-(l+2) blocks and loads it from every
// sub-block of every block of its node's operand (NB * NB values, each read
// from HBM once).  For each product r it forms the level-(l+1) operand's
// sub-block values exactly as the one-level stage would (the same generated
// expression of output a_index[r], or the block itself when the product
// uses it unchanged), then applies the level-(l+1) combination to them and
// writes the outputs, plus the view sub-blocks a materialised operand's
// children read directly.  Every value is computed by the same operations in
// the same order as two one-level stages, so results are bit-identical; only
// the level-(l+1) outputs are never stored.
template <class Cmb, typename T, int SIDE>
__global__ void __launch_bounds__(FM_STAGE_THREADS) fm_in_stage2(FmStageArgs a) {
  constexpr int BC = SIDE == 0 ? Cmb::k : Cmb::n;
  constexpr int NB = (SIDE == 0 ? Cmb::m : Cmb::k) * BC;
  constexpr int R = Cmb::R;
  constexpr int NOUT = SIDE == 0 ? Cmb::NS : Cmb::NT;
  constexpr int NOUT_ARR = NOUT > 0 ? NOUT : 1;
  constexpr int NV = SIDE == 0 ? Cmb::a_nviews : Cmb::b_nviews;
  const int q = blockIdx.y;
  const FmNodeView nv = fm_node_view(a, q, R * NOUT, R * NV);
  const FmOp op = nv.op;
  const fm_gptr<const T> src = fm_global(reinterpret_cast<const T*>(fm_base(a.bases, op.role)) + op.off);
  const fm_gptr<T> ws = fm_global(reinterpret_cast<T*>(a.bases.p[FM_ROLE_WS]));
  const fm_gptr<T> dst = ws + nv.out;
  const fm_gptr<T> vdst = ws + nv.view;
  const uint32_t rows = (uint32_t)nv.rows2;
  const uint32_t total = rows * (uint32_t)nv.cols2;
  const int64_t ld = op.ld, brows = nv.rows, bcols = nv.cols, brows2 = nv.rows2,
                bcols2 = nv.cols2, slab = nv.out_slab;
  const double x = nv.x;
  for (uint32_t e = blockIdx.x * FM_STAGE_THREADS + threadIdx.x; e < total;
       e += gridDim.x * FM_STAGE_THREADS) {
    const uint32_t j = e / rows;
    const uint32_t i = e - j * rows;
    T v[NB][NB];  // v[block][sub-block]
    // one lane address; each load adds its (lane-uniform) block and
    // sub-block offsets, a block's loads at a time, so that the 64-bit
    // addresses of all NB * NB loads are not live together
    const fm_gptr<const T> lsrc = src + i + (int64_t)j * ld;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const fm_gptr<const T> bsrc = lsrc + (int64_t)(b / BC) * brows + (int64_t)(b % BC) * bcols * ld;
#pragma unroll
      for (int sb = 0; sb < NB; ++sb)
        v[b][sb] = bsrc[(int64_t)(sb / BC) * brows2 + (int64_t)(sb % BC) * bcols2 * ld];
      __builtin_amdgcn_sched_barrier(0);
    }
    const int64_t opos = i + (int64_t)j * nv.out_ld;
    auto product = [&](auto rc) {
      constexpr int r = decltype(rc)::value;
      constexpr int kind = SIDE == 0 ? Cmb::a_kind_c[r] : Cmb::b_kind_c[r];
      constexpr int idx = SIDE == 0 ? Cmb::a_index_c[r] : Cmb::b_index_c[r];
      T par[NB];
#pragma unroll
      for (int sb = 0; sb < NB; ++sb) {
        if constexpr (kind == 0) {
          par[sb] = v[idx][sb];
        } else {
          T in[NB];
#pragma unroll
          for (int b = 0; b < NB; ++b) in[b] = v[b][sb];
          if constexpr (SIDE == 0)
            par[sb] = Cmb::template s_one<idx, T>(in, x);
          else
            par[sb] = Cmb::template t_one<idx, T>(in, x);
        }
      }
      T out[NOUT_ARR];
      if constexpr (SIDE == 0)
        Cmb::s_combine(par, out, x);
      else
        Cmb::t_combine(par, out, x);
      // one running address per product, advanced a slab per output: the
      // outputs' slab offsets are not all kept live at once
      fm_gptr<T> op = dst + (int64_t)r * NOUT * slab + opos;
#pragma unroll
      for (int t = 0; t < NOUT; ++t) {
        *op = out[t];
        op += slab;
      }
      if constexpr (kind == 1) {
#pragma unroll
        for (int sb = 0; sb < NB; ++sb) {
          const int slot = SIDE == 0 ? Cmb::a_view_slot[sb] : Cmb::b_view_slot[sb];
          if (slot >= 0) vdst[((int64_t)r * NV + slot) * slab + opos] = par[sb];
        }
      }
    };
    fm_static_for<0, R>(product);
  }
}

// Two-level S (SIDE 0) or T (SIDE 1) of FP32 data written straight as the
// leaves' split-bf16 slice planes (layout in split_bf16.hpp), for leaf
// launches on slice planes: the values of fm_in_stage2 (same expressions,
// same order), each split by fm_split1 as the preparation pass splits a
// stored FP32 slab, so the planes are bit-identical to preparing
// fm_in_stage2's output.  Here out_off, view_off and out_slab are bytes (one
// slab's planes: rows2 * cols2 * 6 bytes, no padding: the plan takes this
// path only when the leaf operands' rows are whole tiles and their K whole
// 32-deep slices).  A lane owns VEC consecutive k of one plane row (A side:
// row i, columns k; B side: column j, rows k), so 32 / VEC consecutive lanes
// write one 64-byte run of each plane.
template <class Cmb, int SIDE>
constexpr int fm_planes_vec() {
  constexpr int NB = SIDE == 0 ? Cmb::m * Cmb::k : Cmb::k * Cmb::n;
  return NB * NB * 4 <= 96 ? 4 : NB * NB * 2 <= 96 ? 2 : 1;
}

template <int VEC>
__device__ __forceinline__ void fm_put_planes(fm_gptr<char> p, const float (&val)[VEC]) {
  uint32_t h[VEC], m[VEC], l[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) fm_split1(val[e], h[e], m[e], l[e]);
  if constexpr (VEC == 4) {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    *(fm_gptr<u2>)p = u2{h[0] | h[1] << 16, h[2] | h[3] << 16};
    *(fm_gptr<u2>)(p + 64) = u2{m[0] | m[1] << 16, m[2] | m[3] << 16};
    *(fm_gptr<u2>)(p + 128) = u2{l[0] | l[1] << 16, l[2] | l[3] << 16};
  } else if constexpr (VEC == 2) {
    *(fm_gptr<uint32_t>)p = h[0] | h[1] << 16;
    *(fm_gptr<uint32_t>)(p + 64) = m[0] | m[1] << 16;
    *(fm_gptr<uint32_t>)(p + 128) = l[0] | l[1] << 16;
  } else {
    *(fm_gptr<uint16_t>)p = (uint16_t)h[0];
    *(fm_gptr<uint16_t>)(p + 64) = (uint16_t)m[0];
    *(fm_gptr<uint16_t>)(p + 128) = (uint16_t)l[0];
  }
}

template <class Cmb, int SIDE>
__global__ void __launch_bounds__(FM_STAGE_THREADS) fm_in_stage2_planes(FmStageArgs a) {
  using T = float;
  constexpr int BC = SIDE == 0 ? Cmb::k : Cmb::n;
  constexpr int NB = (SIDE == 0 ? Cmb::m : Cmb::k) * BC;
  constexpr int R = Cmb::R;
  constexpr int NOUT = SIDE == 0 ? Cmb::NS : Cmb::NT;
  constexpr int NOUT_ARR = NOUT > 0 ? NOUT : 1;
  constexpr int NV = SIDE == 0 ? Cmb::a_nviews : Cmb::b_nviews;
  constexpr int VEC = fm_planes_vec<Cmb, SIDE>();
  constexpr int KQ = 32 / VEC;  // lanes per 32-deep run of a plane row
  constexpr int RT = FM_STAGE_THREADS / KQ;  // plane rows of one work group's lanes
  // A side: the sources go through LDS (below) where their tile leaves room
  // for two work groups per CU (measured, S stage: grey424, 8 blocks, width
  // 1, 74 KB: 3.21 -> 2.52 ms at 16384 x 8192; Strassen, width 4, 68 KB:
  // 1.32 -> 1.19 ms at 16384^3; smirnov333, 9 blocks, 93 KB, one group per
  // CU: 1.20 -> 1.67 ms, so not staged)
  constexpr bool STAGED = SIDE == 0 && NB * NB * 32 * (RT + 1) * (int)sizeof(T) <= 80 * 1024;
  __shared__ T stg[STAGED ? NB * NB * 32 : 1][RT + 1];  // [source x column][row]
  const int q = blockIdx.y;
  const FmNodeView nv = fm_node_view(a, q, R * NOUT, R * NV);
  const FmOp op = nv.op;
  const fm_gptr<const T> src = fm_global(reinterpret_cast<const T*>(fm_base(a.bases, op.role)) + op.off);
  const fm_gptr<char> ws = fm_global(reinterpret_cast<char*>(a.bases.p[FM_ROLE_WS]));
  const fm_gptr<char> dst = ws + nv.out;
  const fm_gptr<char> vdst = ws + nv.view;
  const uint32_t prow = SIDE == 0 ? (uint32_t)nv.rows2 : (uint32_t)nv.cols2;  // plane rows
  const uint32_t nkc = (SIDE == 0 ? (uint32_t)nv.cols2 : (uint32_t)nv.rows2) / 32;
  const uint32_t total = prow * nkc * KQ;
  const int64_t ld = op.ld, brows = nv.rows, bcols = nv.cols, brows2 = nv.rows2,
                bcols2 = nv.cols2, slab = nv.out_slab;
  // the step between a lane's VEC elements: columns (A side) or rows (B side)
  const int64_t estep = SIDE == 0 ? ld : 1;
  const double x = nv.x;
  for (uint32_t e = blockIdx.x * FM_STAGE_THREADS + threadIdx.x; e < total;
       e += gridDim.x * FM_STAGE_THREADS) {
    // k fastest: adjacent lanes write adjacent bytes of a plane run, which
    // the store path merges (with rows fastest instead, the A side's loads
    // coalesce better but its stores do not: measured slower, S of
    // grey424 16384 x 8192 3.2 -> 4.3 ms)
    const uint32_t kq = e % KQ;
    const uint32_t rest = e / KQ;
    const uint32_t r = rest % prow;
    const uint32_t kc = rest / prow;
    const uint32_t k0 = kc * 32 + kq * VEC;
    const uint32_t i = SIDE == 0 ? r : k0, j = SIDE == 0 ? k0 : r;
    T v[NB][NB][VEC];  // v[block][sub-block][element]
    if constexpr (STAGED) {
      // The work group's lanes cover RT consecutive plane rows (A rows
      // r0 ..) of one 32-deep run (columns kc * 32 ..): every lane takes the
      // same trip count (the total is a multiple of 256, plane rows a
      // multiple of RT).  Each source's RT x 32 tile is loaded rows-first
      // (RT consecutive rows of a column per load) into LDS, then each lane
      // reads its k-first elements.  The k-first loads straight from HBM
      // fetched 32 columns' lines for 2 rows per wave instruction.
      const uint32_t r0 = ((e - threadIdx.x) / KQ) % prow;
      const uint32_t sr = threadIdx.x % RT, sc = threadIdx.x / RT;
      const fm_gptr<const T> ssrc = src + (r0 + sr) + (int64_t)(kc * 32 + sc) * ld;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const fm_gptr<const T> bsrc = ssrc + (int64_t)(b / BC) * brows + (int64_t)(b % BC) * bcols * ld;
#pragma unroll
        for (int sb = 0; sb < NB; ++sb) {
          const fm_gptr<const T> p = bsrc + (int64_t)(sb / BC) * brows2 + (int64_t)(sb % BC) * bcols2 * ld;
#pragma unroll
          for (int pp = 0; pp < VEC; ++pp) stg[(b * NB + sb) * 32 + sc + pp * KQ][sr] = p[(int64_t)pp * KQ * ld];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      __syncthreads();
      const uint32_t rl = threadIdx.x / KQ;  // r - r0
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int sb = 0; sb < NB; ++sb)
#pragma unroll
          for (int l = 0; l < VEC; ++l) v[b][sb][l] = stg[(b * NB + sb) * 32 + kq * VEC + l][rl];
      __syncthreads();  // the next trip overwrites the tile
    } else {
      const fm_gptr<const T> lsrc = src + i + (int64_t)j * ld;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const fm_gptr<const T> bsrc = lsrc + (int64_t)(b / BC) * brows + (int64_t)(b % BC) * bcols * ld;
#pragma unroll
        for (int sb = 0; sb < NB; ++sb) {
          const fm_gptr<const T> p = bsrc + (int64_t)(sb / BC) * brows2 + (int64_t)(sb % BC) * bcols2 * ld;
#pragma unroll
          for (int l = 0; l < VEC; ++l) v[b][sb][l] = p[l * estep];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // the lane's run within one slab's planes
    const int64_t pos = ((int64_t)kc * prow + r) * 192 + kq * (VEC * 2);
    auto product = [&](auto rc) {
      constexpr int rr = decltype(rc)::value;
      constexpr int kind = SIDE == 0 ? Cmb::a_kind_c[rr] : Cmb::b_kind_c[rr];
      constexpr int idx = SIDE == 0 ? Cmb::a_index_c[rr] : Cmb::b_index_c[rr];
      T par[NB][VEC];
      T out[NOUT_ARR][VEC];
#pragma unroll
      for (int l = 0; l < VEC; ++l) {
        T pl[NB];
#pragma unroll
        for (int sb = 0; sb < NB; ++sb) {
          if constexpr (kind == 0) {
            pl[sb] = v[idx][sb][l];
          } else {
            T in[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) in[b] = v[b][sb][l];
            if constexpr (SIDE == 0)
              pl[sb] = Cmb::template s_one<idx, T>(in, x);
            else
              pl[sb] = Cmb::template t_one<idx, T>(in, x);
          }
          par[sb][l] = pl[sb];
        }
        T o[NOUT_ARR];
        if constexpr (SIDE == 0)
          Cmb::s_combine(pl, o, x);
        else
          Cmb::t_combine(pl, o, x);
#pragma unroll
        for (int t = 0; t < NOUT; ++t) out[t][l] = o[t];
      }
#pragma unroll
      for (int t = 0; t < NOUT; ++t) fm_put_planes<VEC>(dst + ((int64_t)rr * NOUT + t) * slab + pos, out[t]);
      if constexpr (kind == 1) {
#pragma unroll
        for (int sb = 0; sb < NB; ++sb) {
          const int slot = SIDE == 0 ? Cmb::a_view_slot[sb] : Cmb::b_view_slot[sb];
          if (slot >= 0) fm_put_planes<VEC>(vdst + ((int64_t)rr * NV + slot) * slab + pos, par[sb]);
        }
      }
    };
    fm_static_for<0, R>(product);
  }
}

// Operands with at most this many blocks take the two-level stage
// (NB * NB values per lane in registers).
constexpr int FM_STAGE2_MAX_BLOCKS = 9;

// Two recursion levels of W in one pass (levels l+1 then l), for nodes q
// of level l: the level-(l+1) products M1[q,r] (r < R, blocks rows x cols)
// are never stored.  A work group takes P consecutive rows of one column of
// the level-(l+2) block position and one node:
//   1. for every r, the level-(l+1) W of node q*R + r at those positions:
//      w_combine over M2[(q*R + r)*R + r2] (r2 < R) gives the m*n sub-blocks
//      t of M1[q,r], kept in LDS;
//   2. for every sub-block t, the level-l W: w_combine over M1[q,r][t]
//      (r < R) gives C_q's blocks ij at sub-block t (+ beta * C).
// Same operations in the same order as the two one-level W stages.
template <class Cmb, typename T>
constexpr int fm_w2_rows() {
  // rows per work group: LDS holds R * m*n * P values, at most 32 KB for
  // 64 or 128 rows (small tables: longer runs of each product read), 64 KB
  // for 32 rows and fewer
  constexpr int per = Cmb::R * Cmb::m * Cmb::n * (int)sizeof(T);
  return per * 128 <= 32768 ? 128 : per * 64 <= 32768 ? 64 : per * 32 <= 65536 ? 32
       : per * 16 <= 65536 ? 16 : per * 8 <= 65536 ? 8 : 0;
}

template <class Cmb, typename T, bool BETA>
__global__ void __launch_bounds__(FM_STAGE_THREADS) fm_w_stage2(FmStageArgs a) {
  constexpr int R = Cmb::R, MN = Cmb::m * Cmb::n, P = fm_w2_rows<Cmb, T>();
  static_assert(P > 0, "level-(l+1) products do not fit in LDS");
  __shared__ T m1[R][MN][P];
  const int q = blockIdx.y;
  const FmNodeView nv = fm_node_view(a, q, R * R, 0);
  const int rows2 = nv.rows2;
  const int chunks = (rows2 + P - 1) / P;
  if ((int)blockIdx.x >= chunks * nv.cols2) return;  // grouped launch: a smaller node
  const int j = blockIdx.x / chunks;           // column within the sub-block
  const int i0 = (blockIdx.x - j * chunks) * P;  // first row
  const fm_gptr<const T> mb = fm_global(reinterpret_cast<const T*>(a.bases.p[FM_ROLE_WS]) + nv.out);
  const int64_t col_off = (int64_t)j * nv.out_ld;
  const int64_t slab = nv.out_slab;
  const double x = nv.x;
  // 1. level-(l+1) W, one (row, r) item per thread and pass
  for (int item = threadIdx.x; item < P * R; item += FM_STAGE_THREADS) {
    const int p = item % P, r = item / P;
    const int i = i0 + p;
    if (i >= rows2) continue;
    T mv[R];
    const fm_gptr<const T> src = mb + (int64_t)r * R * slab + i + col_off;
#pragma unroll
    for (int r2 = 0; r2 < R; ++r2) mv[r2] = src[(int64_t)r2 * slab];
    T out[MN];
    Cmb::w_combine(mv, out, x);
#pragma unroll
    for (int t = 0; t < MN; ++t) m1[r][t][p] = out[t];
  }
  __syncthreads();
  // 2. level-l W, one (row, sub-block t) item per thread and pass
  const FmOp op = nv.op;
  const fm_gptr<T> c = fm_global(reinterpret_cast<T*>(fm_base(a.bases, op.role)) + op.off);
  const int64_t ld = op.ld;
  const T beta = (T)nv.beta;
  for (int item = threadIdx.x; item < P * MN; item += FM_STAGE_THREADS) {
    const int p = item % P, t = item / P;
    const int i = i0 + p;
    if (i >= rows2) continue;
    T mv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) mv[r] = m1[r][t][p];
    T out[MN];
    Cmb::w_combine(mv, out, x);
    const int ti = t / Cmb::n, tj = t % Cmb::n;
    const int64_t r0 = (int64_t)ti * rows2 + i, c0 = (int64_t)tj * nv.cols2 + j;
#pragma unroll
    for (int b = 0; b < MN; ++b) {
      const int bi = b / Cmb::n, bj = b % Cmb::n;
      const fm_gptr<T> dst = c + (int64_t)bi * nv.rows + r0 + ((int64_t)bj * nv.cols + c0) * ld;
      T v = out[b];
      if constexpr (BETA) v = fm_add(v, fm_mul(beta, *dst));
      *dst = v;
    }
  }
}

inline dim3 fm_stage_grid(const FmStageArgs& a, int vec) {
  const uint64_t total = (uint64_t)(a.rows / vec) * (uint64_t)a.cols;
  uint64_t gx = (total + FM_STAGE_THREADS - 1) / FM_STAGE_THREADS;
  // enough work groups to fill 256 CUs several times over, across all nodes
  uint64_t cap = 8192 / (uint64_t)(a.nodes > 0 ? a.nodes : 1);
  if (cap < 4) cap = 4;
  if (gx > cap) gx = cap;
  if (gx < 1) gx = 1;
  return dim3((unsigned)gx, (unsigned)a.nodes, 1);
}

template <class Cmb, typename T, int SIDE>
hipError_t fm_launch_in_stage(const FmStageArgs& a, hipStream_t s) {
  constexpr int VMAX = sizeof(T) == 8 ? 2 : 4;
  if (a.vec == VMAX) {
    hipLaunchKernelGGL((fm_in_stage<Cmb, T, VMAX, SIDE>), fm_stage_grid(a, VMAX),
                       dim3(FM_STAGE_THREADS), 0, s, a);
  } else {
    hipLaunchKernelGGL((fm_in_stage<Cmb, T, 1, SIDE>), fm_stage_grid(a, 1),
                       dim3(FM_STAGE_THREADS), 0, s, a);
  }
  return hipGetLastError();
}

template <class Cmb, typename T>
hipError_t fm_launch_w_stage(const FmStageArgs& a, hipStream_t s) {
  constexpr int VMAX = sizeof(T) == 8 ? 2 : 4;
  const bool beta = a.beta != 0.0;
  if (a.vec == VMAX) {
    if (beta)
      hipLaunchKernelGGL((fm_w_stage<Cmb, T, VMAX, true>), fm_stage_grid(a, VMAX),
                         dim3(FM_STAGE_THREADS), 0, s, a);
    else
      hipLaunchKernelGGL((fm_w_stage<Cmb, T, VMAX, false>), fm_stage_grid(a, VMAX),
                         dim3(FM_STAGE_THREADS), 0, s, a);
  } else {
    if (beta)
      hipLaunchKernelGGL((fm_w_stage<Cmb, T, 1, true>), fm_stage_grid(a, 1),
                         dim3(FM_STAGE_THREADS), 0, s, a);
    else
      hipLaunchKernelGGL((fm_w_stage<Cmb, T, 1, false>), fm_stage_grid(a, 1),
                         dim3(FM_STAGE_THREADS), 0, s, a);
  }
  return hipGetLastError();
}

inline dim3 fm_stage2_grid(const FmStageArgs& a) {
  const uint64_t total = (uint64_t)a.rows2 * (uint64_t)a.cols2;
  uint64_t gx = (total + FM_STAGE_THREADS - 1) / FM_STAGE_THREADS;
  uint64_t cap = 8192 / (uint64_t)(a.nodes > 0 ? a.nodes : 1);
  if (cap < 4) cap = 4;
  if (gx > cap) gx = cap;
  if (gx < 1) gx = 1;
  return dim3((unsigned)gx, (unsigned)a.nodes, 1);
}

template <class Cmb, typename T, int SIDE>
hipError_t fm_launch_in_stage2(const FmStageArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((fm_in_stage2<Cmb, T, SIDE>), fm_stage2_grid(a), dim3(FM_STAGE_THREADS), 0, s,
                     a);
  return hipGetLastError();
}

template <class Cmb, typename T>
hipError_t fm_launch_w_stage2(const FmStageArgs& a, hipStream_t s) {
  constexpr int P = fm_w2_rows<Cmb, T>();
  const unsigned chunks = (unsigned)((a.rows2 + P - 1) / P);
  const dim3 grid(chunks * (unsigned)a.cols2, (unsigned)a.nodes, 1);
  if (a.beta != 0.0)
    hipLaunchKernelGGL((fm_w_stage2<Cmb, T, true>), grid, dim3(FM_STAGE_THREADS), 0, s, a);
  else
    hipLaunchKernelGGL((fm_w_stage2<Cmb, T, false>), grid, dim3(FM_STAGE_THREADS), 0, s, a);
  return hipGetLastError();
}

template <class Cmb, typename T>
constexpr FmStageLaunch fm_w2_fn() {
  if constexpr (fm_w2_rows<Cmb, T>() > 0)
    return &fm_launch_w_stage2<Cmb, T>;
  else
    return nullptr;
}

template <class Cmb, int SIDE>
constexpr bool fm_stage2_ok() {
  return (SIDE == 0 ? Cmb::m * Cmb::k : Cmb::k * Cmb::n) <= FM_STAGE2_MAX_BLOCKS &&
         (SIDE == 0 ? Cmb::NS : Cmb::NT) > 0;
}

template <class Cmb, typename T, int SIDE>
constexpr FmStageLaunch fm_stage2_fn() {
  if constexpr (fm_stage2_ok<Cmb, SIDE>())
    return &fm_launch_in_stage2<Cmb, T, SIDE>;
  else
    return nullptr;
}

template <class Cmb, int SIDE>
hipError_t fm_launch_in_stage2_planes(const FmStageArgs& a, hipStream_t s) {
  constexpr int VEC = fm_planes_vec<Cmb, SIDE>();
  const uint64_t total = (uint64_t)a.rows2 * (uint64_t)a.cols2 / VEC;
  uint64_t gx = (total + FM_STAGE_THREADS - 1) / FM_STAGE_THREADS;
  uint64_t cap = 8192 / (uint64_t)(a.nodes > 0 ? a.nodes : 1);
  if (cap < 4) cap = 4;
  if (gx > cap) gx = cap;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL((fm_in_stage2_planes<Cmb, SIDE>), dim3((unsigned)gx, (unsigned)a.nodes, 1),
                     dim3(FM_STAGE_THREADS), 0, s, a);
  return hipGetLastError();
}

// Plane-writing stages for tables up to NB^2 R = 2048 (every table up to
// <3,3,3;23> and <4,2,4;26>): beyond that (the R = 40 and R = 29 families,
// classical333) the fully unrolled kernel triples the file's compile time
// (<6,3,3;40>: 57 -> 182 s) and those tables keep the separate preparation.
template <class Cmb, int SIDE>
constexpr int fm_planes_weight() {
  constexpr int NB = SIDE == 0 ? Cmb::m * Cmb::k : Cmb::k * Cmb::n;
  return NB * NB * Cmb::R;
}

template <class Cmb, int SIDE>
constexpr FmStageLaunch fm_stage2_planes_fn() {
  if constexpr (fm_stage2_ok<Cmb, SIDE>() && fm_planes_weight<Cmb, SIDE>() <= 2048)
    return &fm_launch_in_stage2_planes<Cmb, SIDE>;
  else
    return nullptr;
}

#define FM_DEFINE_ALGORITHM(NAME, GNS, APA)                                                   \
  static hipError_t fm_##NAME##_s_d(const FmStageArgs& a, hipStream_t s) {                   \
    return fm_launch_in_stage<GNS::Combine, double, 0>(a, s);                                 \
  }                                                                                          \
  static hipError_t fm_##NAME##_t_d(const FmStageArgs& a, hipStream_t s) {                   \
    return fm_launch_in_stage<GNS::Combine, double, 1>(a, s);                                 \
  }                                                                                          \
  static hipError_t fm_##NAME##_w_d(const FmStageArgs& a, hipStream_t s) {                   \
    return fm_launch_w_stage<GNS::Combine, double>(a, s);                                     \
  }                                                                                          \
  static hipError_t fm_##NAME##_s_f(const FmStageArgs& a, hipStream_t s) {                   \
    return fm_launch_in_stage<GNS::Combine, float, 0>(a, s);                                  \
  }                                                                                          \
  static hipError_t fm_##NAME##_t_f(const FmStageArgs& a, hipStream_t s) {                   \
    return fm_launch_in_stage<GNS::Combine, float, 1>(a, s);                                  \
  }                                                                                          \
  static hipError_t fm_##NAME##_w_f(const FmStageArgs& a, hipStream_t s) {                   \
    return fm_launch_w_stage<GNS::Combine, float>(a, s);                                      \
  }                                                                                          \
  const FmAlgDesc* fm_alg_##NAME() {                                                        \
    static const FmAlgDesc d = {                                                             \
        #NAME, GNS::Combine::m, GNS::Combine::k, GNS::Combine::n, GNS::Combine::R,           \
        GNS::Combine::NS, GNS::Combine::NT, APA, GNS::a_kind, GNS::a_index, GNS::a_coef,     \
        GNS::b_kind, GNS::b_index, GNS::b_coef, fm_##NAME##_s_d, fm_##NAME##_t_d,           \
        fm_##NAME##_w_d, fm_##NAME##_s_f, fm_##NAME##_t_f, fm_##NAME##_w_f,                   \
        fm_stage2_fn<GNS::Combine, double, 0>(), fm_stage2_fn<GNS::Combine, double, 1>(),     \
        fm_stage2_fn<GNS::Combine, float, 0>(), fm_stage2_fn<GNS::Combine, float, 1>(),       \
        GNS::Combine::a_nviews, GNS::Combine::b_nviews, GNS::Combine::a_view_slot,            \
        GNS::Combine::b_view_slot, fm_w2_fn<GNS::Combine, double>(),                          \
        fm_w2_fn<GNS::Combine, float>(), fm_stage2_planes_fn<GNS::Combine, 0>(),             \
        fm_stage2_planes_fn<GNS::Combine, 1>()};                                             \
    return &d;                                                                               \
  }
