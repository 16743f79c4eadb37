"""Non-default schedules and tiles against the oracle itself (not only
against the default GPU plan): depth-first levels, the automatic cut-off
under a workspace limit, the remainder split and carve-out off, one-level
stages, forced leaf tiles, and the split-bf16 slice-plane kernels.

Same bar as the golden parity tests: the oracle first reproduces the
reference's digest for the case, then
max|C_gpu - C_oracle| / (max|A| max|B| K) <= 2 e_ref + floor, e_ref being the
reference's own normwise error against classical GEMM on that case
(golden.json), floor 1e-15 (FP64) / 1e-6 (FP32) times max(1,|alpha|)(1+|beta|).
"""
import numpy as np
import pytest
import torch

from cases import golden_cases, make_inputs, sha, x_for

pytestmark = pytest.mark.gpu

DEV = "cuda:0"

SUBSET = ["strassen_2_366x500x200", "grey424_26_206_2_366x500x200",
          "smirnov333_23_139_2_300x500x100", "grey423_20_144_2_257x500x55",
          "grey522_18_99_2_300x500x100", "ab_strassen_1_2.5_2_1007x400x679",
          "ab_smirnov333_23_128_7_-1.3_2_1001x200x309", "f32_strassen_2", "f32_grey424_26_206_2",
          "u11_grey424_26_206_2_512"]

OPTS = {
    "dfs1": dict(dfs_levels=1),
    "dfs2": dict(dfs_levels=2),
    "ws_limit": None,  # automatic cut-off under the fully depth-first plan's workspace
    "no_split_remainder": dict(split_remainder=0),
    "one_level_stages": dict(stage_fusion=0),
    "tile64": dict(gemm_tile=1),
    "tile128x64_w8": dict(gemm_tile=12),
    "tile128x96": dict(gemm_tile=30),
    # the spread kernels: whole-tile, whole-slice edge and predicated variants
    "tile_spread128": dict(gemm_tile=32),
    "tile_spread128x256": dict(gemm_tile=33),
    "tile_spread128x96": dict(gemm_tile=34),
}


def _dev(a, fm):
    if a is None:
        return None
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


def _normwise(C, D, A, B):
    return float(np.abs(C.astype(np.float64) - D.astype(np.float64)).max()) / (
        float(np.abs(A).max()) * float(np.abs(B).max()) * A.shape[1])


def _case(cid):
    return {c[0]: c for c in golden_cases()}[cid]


def _oracle_and_inputs(oracle_mod, golden, cid):
    (cid, alg, m, k, n, steps, alpha, beta, dist, dt, seed) = _case(cid)
    rec = {r["id"]: r for r in golden["cases"]}[cid]
    dtype = np.float64 if dt == "f64" else np.float32
    A, B, C0 = make_inputs(m, k, n, seed, dist, dtype, with_c=beta != 0.0)
    x = x_for(alg, steps)
    Co = oracle_mod.fastmatmul(alg, A, B, C0, steps, alpha, beta, x)
    if oracle_mod.reproducible():
        assert sha(Co) == rec["sha256"], "oracle no longer reproduces the reference"
    floor = (1e-15 if dtype == np.float64 else 1e-6) * max(1.0, abs(alpha)) * (1 + abs(beta))
    return alg, steps, alpha, beta, x, A, B, C0, Co, 2 * rec["normwise"] + floor


@pytest.mark.parametrize("opt", sorted(OPTS))
@pytest.mark.parametrize("cid", SUBSET)
def test_schedule_vs_oracle(fm, oracle_mod, golden, cid, opt):
    alg, steps, alpha, beta, x, A, B, C0, Co, tol = _oracle_and_inputs(oracle_mod, golden, cid)
    opts = OPTS[opt]
    if opts is None:
        m, k = A.shape
        n = B.shape[1]
        tdt = torch.float64 if A.dtype == np.float64 else torch.float32
        full_dfs = fm.Plan(alg, m, k, n, steps, dtype=tdt, x=x, alpha=alpha, beta=beta,
                           opts=dict(dfs_levels=steps)).workspace_bytes()
        opts = dict(dfs_levels=-1, workspace_limit=max(full_dfs, 1))
    C = fm.fastmatmul(alg, _dev(A, fm), _dev(B, fm), _dev(C0, fm), steps, x, alpha, beta,
                      opts=opts)
    assert _normwise(C.cpu().numpy(), Co, A, B) <= tol


# split-bf16 leaves on every slice-plane kernel, and on the in-loop kernel
SPLIT_TILES = [-1, 0, 24, 25, 26, 27, 28, 29, 31]


@pytest.mark.parametrize("tile", SPLIT_TILES)
@pytest.mark.parametrize("cid", ["f32_strassen_2", "f32_grey424_26_206_2", "f32_smirnov333_23_139_1",
                                 "f32_grey423_20_144_2"])
def test_split_bf16_tiles_vs_oracle(fm, oracle_mod, golden, cid, tile):
    """FP32 data with split-bf16 leaves (FP32-accurate products) against the
    oracle's FP32 run of the same algorithm (MKL sgemm leaves): the same bar
    as plain FP32 leaves."""
    alg, steps, alpha, beta, x, A, B, C0, Co, tol = _oracle_and_inputs(oracle_mod, golden, cid)
    C = fm.fastmatmul(alg, _dev(A, fm), _dev(B, fm), _dev(C0, fm), steps, x, alpha, beta,
                      mode="split_bf16", opts=dict(gemm_tile=tile))
    assert _normwise(C.cpu().numpy(), Co, A, B) <= tol


# Split-bf16 leaves whose operands' slice planes the last two-level S/T
# stage writes itself (the default where the leaves' rows are multiples of
# 256 and their K of 32; the golden cases are too small for that path):
# against the oracle's FP32 run on the same seeded inputs, with e_ref the
# oracle's own normwise error against the FP64 product.
@pytest.mark.parametrize("alg,steps,shape", [("strassen", 2, (1024, 1024, 1024)),
                                             ("grey424_26_206", 2, (4096, 256, 4096)),
                                             ("smirnov333_23_139", 2, (2304, 288, 2304))])
def test_split_planes_from_stages_vs_oracle(fm, oracle_mod, alg, steps, shape):
    m, k, n = shape
    A, B, _ = make_inputs(m, k, n, 51, "u11", np.float32)
    x = x_for(alg, steps)
    Co = oracle_mod.fastmatmul(alg, A, B, None, steps, 1.0, 0.0, x)
    D = A.astype(np.float64) @ B.astype(np.float64)
    e_ref = _normwise(Co, D, A, B)
    try:
        fm.set_stage_fusion(1)
        fused = fm.plan_info(alg, m, k, n, steps, dtype=torch.float32, mode="split_bf16")
        fm.set_stage_fusion(2)
        apart = fm.plan_info(alg, m, k, n, steps, dtype=torch.float32, mode="split_bf16")
    finally:
        fm.set_stage_fusion(1)
    assert fused["prep_bytes"] < apart["prep_bytes"]  # the planes come from the stage
    C = fm.fastmatmul(alg, _dev(A, fm), _dev(B, fm), None, steps, x, mode="split_bf16")
    assert _normwise(C.cpu().numpy(), Co, A, B) <= 2 * e_ref + 1e-6
