// alg_desc.hpp -- host-side description of one generated fast algorithm and
// the argument blocks of the stage kernels.
//
// Every device address in a plan is written as (role, offset): role selects
// one of four base pointers that are only known at call time (the caller's
// A, B, C and the workspace arena), so a plan built once for a shape can be
// replayed for any buffers (and captured in a graph) without re-uploading
// descriptors.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Phases of a leaf launch on prepared planes (split-bf16 slices, int8
// digits): the preparation kernels that cut the operands into planes, the
// products on them, or both.
enum FmLeafPhase : int { FM_PHASE_PREP = 1, FM_PHASE_PRODUCT = 2, FM_PHASE_ALL = 3 };

enum FmRole : int32_t { FM_ROLE_A = 0, FM_ROLE_B = 1, FM_ROLE_C = 2, FM_ROLE_WS = 3 };
// plan-time marker of a leaf operand whose split-bf16 slice planes a stage
// writes (FmOp::off: their byte offset in the workspace); never a kernel's role
constexpr int32_t FM_ROLE_PLANES = -2;

// Roles 0-3 come from p[]; a batched plan numbers its problems' operands
// from 4 on and resolves them through a device table (ext[role - 4]).
struct FmBases {
  void* p[4];
  void* const* ext;
};

__host__ __device__ __forceinline__ void* fm_base(const FmBases& b, int32_t role) {
  return role < 4 ? b.p[role] : b.ext[role - 4];
}

// A column-major operand: element (i, j) at base[role] + off + i + j * ld.
struct FmOp {
  int32_t role;
  int32_t pad;
  int64_t off;  // in elements
  int64_t ld;   // in elements
};

// S/T stage: for every node q, read the node's operand (ops[q]) split in
// bm x bn blocks of rows x cols, write NOUT outputs of rows x cols at
// ws + out_off + (q * NOUT + o) * out_slab with leading dimension out_ld.
// W stage: read R products of node q from the same slab layout and write the
// m x n blocks of ops[q] (C_ij = sum W M (+ beta C_ij)).
//
// Two-level S/T stage (one launch for levels l and l+1): ops[q] are the
// level-l operands, split in bm x bn blocks of rows x cols, each block split
// again in bm x bn sub-blocks of rows2 x cols2.  The level-(l+1) outputs are
// never written; the level-(l+2) outputs of child node q * R + r go to
// ws + out_off + ((q * R + r) * NOUT + o) * out_slab, and the sub-blocks of
// a materialised child that its own children use unchanged (the "views")
// go to ws + view_off + ((q * R + r) * NVIEW + slot) * out_slab.
// Two-level W stage: ops[q] are the level-l C operands (blocks rows x cols,
// sub-blocks rows2 x cols2); the level-(l+2) products of node q are at
// ws + out_off + ((q * R + r) * R + r2) * out_slab.
// Two-level S/T stage writing split-bf16 slice planes (FP32 data): the same
// slab numbering, with out_off, view_off and out_slab in bytes (one slab:
// the planes of a rows2 x cols2 operand) and out_ld unused.
//
// Grouped stage launch (the nodes of several problems, whose blocks differ
// in shape, in one launch): `node` holds each node's operand, shape, output
// placement and scalars, and the launch-wide fields hold the largest shape
// (which sizes the grid).  A node's outputs start at ws + out_off with the
// same slab layout as above for its own index 0.
struct FmStageNode {
  FmOp op;
  int32_t rows, cols;    // blocks (one-level) / blocks of the level-l operand (two-level)
  int32_t rows2, cols2;  // two-level stages: sub-blocks
  int64_t out_off;       // the node's first output slab (workspace elements)
  int64_t out_ld, out_slab;
  int64_t view_off;      // two-level S/T: the node's first view slab
  double x, beta;
};

struct FmStageArgs {
  FmBases bases;
  const FmOp* ops;
  int32_t rows, cols, nodes, vec;
  int64_t out_off, out_ld, out_slab;
  double x;
  double beta;
  int32_t rows2, cols2;
  int64_t view_off;
  const FmStageNode* node;  // grouped launch: per-node descriptors (nullptr otherwise)
};

// The placement of node q of a stage launch, uniform or grouped.  nout: output
// slabs per node of the one-level layout (S/T: NS/NT; W: R; two-level S/T:
// R * NOUT; two-level W: R * R); nview: view slabs per node (R * NV).
struct FmNodeView {
  FmOp op;
  int32_t rows, cols, rows2, cols2;
  int64_t out, out_ld, out_slab, view;
  double x, beta;
};

__host__ __device__ __forceinline__ FmNodeView fm_node_view(const FmStageArgs& a, int q, int nout,
                                                           int nview) {
  FmNodeView v;
  if (a.node) {
    const FmStageNode& n = a.node[q];
    v.op = n.op;
    v.rows = n.rows;
    v.cols = n.cols;
    v.rows2 = n.rows2;
    v.cols2 = n.cols2;
    v.out = n.out_off;
    v.out_ld = n.out_ld;
    v.out_slab = n.out_slab;
    v.view = n.view_off;
    v.x = n.x;
    v.beta = n.beta;
  } else {
    v.op = a.ops[q];
    v.rows = a.rows;
    v.cols = a.cols;
    v.rows2 = a.rows2;
    v.cols2 = a.cols2;
    v.out = a.out_off + (int64_t)q * nout * a.out_slab;
    v.out_ld = a.out_ld;
    v.out_slab = a.out_slab;
    v.view = a.view_off + (int64_t)q * nview * a.out_slab;
    v.x = a.x;
    v.beta = a.beta;
  }
  return v;
}

typedef hipError_t (*FmStageLaunch)(const FmStageArgs& a, hipStream_t s);

struct FmAlgDesc {
  const char* name;
  int m, k, n, R;
  int ns, nt;  // materialised S / T outputs per node
  bool apa;
  // per product r: kind 1 = materialised output index, 0 = block index
  const int* a_kind;
  const int* a_index;
  double (*a_coef)(int r, double x);
  const int* b_kind;
  const int* b_index;
  double (*b_coef)(int r, double x);
  FmStageLaunch s_d, t_d, w_d, s_f, t_f, w_f;
  // two-level S/T stages (nullptr where the operand has too many blocks to
  // hold one element of every sub-block in registers)
  FmStageLaunch s2_d, t2_d, s2_f, t2_f;
  // per block of the A (B) side: slot of its view buffer in a two-level
  // stage, or -1 when no product uses that block unchanged
  int a_nviews, b_nviews;
  const int* a_view_slot;
  const int* b_view_slot;
  // two-level W stage (nullptr where a level's products do not fit in LDS)
  FmStageLaunch w2_d, w2_f;
  // two-level S/T stages of FP32 data writing the leaves' split-bf16 slice
  // planes (nullptr where s2_f / t2_f are)
  FmStageLaunch s2p_f, t2p_f;
};

// generated registry (csrc/gen/registry.cpp): one accessor per algorithm
typedef const FmAlgDesc* (*FmAlgAccessor)();
extern const FmAlgAccessor fm_all_algorithms[];
extern const int fm_num_all_algorithms;
