"""Plan structure on the host (no GPU): two-level stage pairing, depth-first
schedules and the leaf remainder split, through fm_plan_query.

The expected counts follow from the recursion (algorithms/strassen.hpp:
378-659) and the schedule rules: R^l nodes per level, levels paired from
the deepest up when the middle level has no peeling, one child sub-plan per
depth-first node, and the last part-empty round of leaf tiles moved to a
second launch."""
import pytest
import torch

import fastmm


@pytest.fixture(autouse=True)
def _reset():
    yield
    fastmm.set_schedule(0, 0, True)
    fastmm.set_stage_fusion(True)


def test_bench_workload_plan():
    # grey424_26_206, 2 levels, 8192^3: one two-level S and T launch, W per
    # level, 676 leaves of 512 x 2048 x 512 = 10816 tiles of 128 x 128
    # (FmTileCfg 32, two work groups per CU) on 256 CUs: 21 full rounds
    # (672 products) + 4 products on 64 x 64 tiles (FmTileCfg 1)
    p = fastmm.plan_info("grey424_26_206", 8192, 8192, 8192, 2)
    assert p["launches"] == {"S": 1, "T": 1, "W": 1, "LEAF": 2, "PEEL": 0}
    assert p["two_level_stages"] == 3
    assert p["leaf"] == [{"cfg": 32, "products": 672, "tiles": 10752},
                         {"cfg": 1, "products": 4, "tiles": 256}]
    fastmm.set_schedule(0, 0, False)
    p = fastmm.plan_info("grey424_26_206", 8192, 8192, 8192, 2)
    assert p["leaf"] == [{"cfg": 32, "products": 676, "tiles": 10816}]


def test_split_workload_plan():
    # Strassen, 2 levels, 16384^3 FP32 with split-bf16 leaves: 49 leaves of
    # 4096^3 on the kernel that loads B straight into registers (FmTileCfg
    # 31, 128 x 256 tiles): 49 x 32 x 16 = 25088 tiles in one launch.
    # With the planes cut by a separate pass (stage_fusion 2): the slice
    # planes of 98 distinct operands (rows x K x 6 bytes each) on top of the
    # plain FP32 plan's workspace.  With the last S/T stage writing them
    # (default): its 2 x 49 outputs and views take 6 bytes an element instead
    # of 4, and only the 8 leaf operands that are sub-blocks of A or B get
    # planes from the preparation pass
    q = fastmm.plan_info("strassen", 16384, 16384, 16384, 2, dtype=torch.float32)
    p = fastmm.plan_info("strassen", 16384, 16384, 16384, 2, dtype=torch.float32, mode="split_bf16")
    assert p["leaf"] == [{"cfg": 31, "products": 49, "tiles": 25088}]
    assert p["workspace_bytes"] - q["workspace_bytes"] == 4096 * 4096 * (2 * 49 * 2 + 8 * 6)
    assert p["prep_bytes"] == 8 * 4096 * 4096 * (4 + 6)
    try:
        fastmm.set_stage_fusion(2)
        p = fastmm.plan_info("strassen", 16384, 16384, 16384, 2, dtype=torch.float32, mode="split_bf16")
    finally:
        fastmm.set_stage_fusion(1)
    assert p["leaf"] == [{"cfg": 31, "products": 49, "tiles": 25088}]
    assert p["workspace_bytes"] - q["workspace_bytes"] == 98 * 4096 * 4096 * 6
    assert p["prep_bytes"] == 98 * 4096 * 4096 * (4 + 6)


def test_one_level_stages_when_middle_level_peels():
    p = fastmm.plan_info("strassen", 130, 130, 130, 2)
    assert p["two_level_stages"] == 0
    assert p["launches"]["S"] == 2 and p["launches"]["T"] == 2 and p["launches"]["W"] == 2
    assert p["launches"]["PEEL"] == 1   # 130 splits evenly, 65 does not: level 1 peels
    fastmm.set_stage_fusion(False)
    p = fastmm.plan_info("grey424_26_206", 8192, 8192, 8192, 2)
    assert p["two_level_stages"] == 0 and p["launches"]["S"] == 2


def test_depth_first_top_level():
    fastmm.set_schedule(1)
    p = fastmm.plan_info("grey424_26_206", 8192, 8192, 8192, 2)
    # level-0 S/T/W plus, per child, its own S/T/W and one leaf launch
    assert p["launches"] == {"S": 27, "T": 27, "W": 27, "LEAF": 26, "PEEL": 0}
    assert p["leaf"][0] == {"cfg": 32, "products": 26, "tiles": 416}
    # level-0 S/T/M blocks plus one child's S/T/M blocks, 8-byte elements
    lvl0 = 22 * 2048 * 4096 + 20 * 4096 * 2048 + 26 * 2048 * 2048
    child = 22 * 512 * 2048 + 20 * 2048 * 512 + 26 * 512 * 512
    assert p["workspace_bytes"] == 8 * (lvl0 + child)


def test_workspace_limit_selects_depth_first():
    bfs = fastmm.plan_info("strassen", 256, 256, 256, 2)["workspace_bytes"]
    fastmm.set_schedule(1)
    dfs = fastmm.plan_info("strassen", 256, 256, 256, 2)["workspace_bytes"]
    # level 0: 5 S + 5 T + 7 M blocks of 128^2; one child: 5 + 5 + 7 of 64^2
    assert dfs == 8 * (17 * 128 * 128 + 17 * 64 * 64) < bfs
    fastmm.set_schedule(0, bfs)
    assert fastmm.plan_info("strassen", 256, 256, 256, 2)["workspace_bytes"] == bfs
    fastmm.set_schedule(0, bfs - 1)
    info = fastmm.plan_info("strassen", 256, 256, 256, 2)
    assert info["workspace_bytes"] == dfs and info["dfs_levels"] == 1
    # the automatic cut-off under the same limit chooses the same plan
    fastmm.set_schedule(-1, bfs - 1)
    info = fastmm.plan_info("strassen", 256, 256, 256, 2)
    assert info["workspace_bytes"] == dfs and info["dfs_levels"] == 1
    fastmm.set_schedule(0, 1000)
    with pytest.raises(fastmm.FastMMError):
        fastmm.plan_info("strassen", 256, 256, 256, 2)
    with pytest.raises(ValueError):
        fastmm.set_schedule(-2)


def test_remainder_split_modes():
    p = fastmm.plan_info("strassen", 4096, 4096, 4096, 2, dtype=torch.float32, mode="split_bf16")
    assert [l["products"] for l in p["leaf"]] == [48, 1]
    # int8-digit leaves keep one launch
    p = fastmm.plan_info("strassen", 4096, 4096, 4096, 2, mode="ozaki_int8")
    assert p["launches"]["LEAF"] == 1 and p["leaf"][0]["cfg"] == -1


def test_split_leaves_on_slice_planes():
    """split_bf16 leaves run on bf16 slice planes prepared in the workspace:
    per distinct operand rows_pad x Kp x 3 planes x 2 bytes (rows padded to
    the tile, K to 32), on top of the plain FP32 plan's workspace."""
    p = fastmm.plan_info("strassen", 4096, 4096, 4096, 1, dtype=torch.float32, mode="split_bf16")
    q = fastmm.plan_info("strassen", 4096, 4096, 4096, 1, dtype=torch.float32)
    assert p["leaf"][0]["cfg"] in (24, 25, 26, 27, 28, 29, 31)
    # Strassen level 1: 7 products of 2048^3, each with its own A-side
    # operand (5 sums, the blocks A11 and A22) and B-side operand: 14 planes
    planes = p["workspace_bytes"] - q["workspace_bytes"]
    assert planes == 14 * 2048 * 2048 * 6
    # ragged shapes: rows padded to the tile, K to 32
    fastmm.set_gemm_tile(24)
    try:
        p = fastmm.plan_info("strassen", 300, 100, 200, 0, dtype=torch.float32, mode="split_bf16")
    finally:
        fastmm.set_gemm_tile(-1)
    assert p["leaf"] == [{"cfg": 24, "products": 1, "tiles": 3 * 2}]
    assert p["workspace_bytes"] == 384 * 128 * 6 + 256 * 128 * 6


def test_split_leaves_beyond_the_prep_grid_split_on_the_fly():
    # the preparation is one flat launch: a very deep K still takes the
    # slice planes; more than 2^31 - 1 preparation work groups (2^20 x 2^24
    # x 2^20: 32767 row blocks x 131072 K blocks) take the in-loop split
    # kernel (FmTileCfg 0)
    p = fastmm.plan_info("strassen", 64, 2200000, 64, 0, dtype=torch.float32, mode="split_bf16")
    assert p["leaf"][0]["cfg"] >= 24
    p = fastmm.plan_info("strassen", 1 << 20, 1 << 24, 1 << 20, 0, dtype=torch.float32,
                         mode="split_bf16")
    assert p["leaf"][0]["cfg"] == 0
    # beyond 2^31 - 1 output tiles of 64 x 64 no launch can number its tiles
    with pytest.raises(ValueError):
        fastmm.plan_info("strassen", 1 << 24, 64, 1 << 24, 0, dtype=torch.float32)


def test_stage_blocks_beyond_32_bit_counts_are_refused():
    # Strassen L=1 on 131072^3: 65536^2 = 2^32-element blocks; L=2 is fine
    with pytest.raises(ValueError):
        fastmm.plan_info("strassen", 131072, 131072, 131072, 1, dtype=torch.float32)
    p = fastmm.plan_info("strassen", 131072, 131072, 131072, 2, dtype=torch.float32)
    assert p["launches"]["LEAF"] >= 1


def test_plane_tiles_need_split_mode():
    fastmm.set_gemm_tile(25)
    try:
        with pytest.raises(ValueError):
            fastmm.plan_info("strassen", 1024, 1024, 1024, 1, dtype=torch.float32)
    finally:
        fastmm.set_gemm_tile(-1)


def test_stage_bytes_match_closed_forms():
    """Algorithmic bytes per stage, grey424_26_206 L=2 FP64 8192^3: one-level
    S = 8 (blocks of level 1 + 26 x blocks of level 2) x (8 reads + 22 writes),
    the per-level figures 8.56 / 7.99 / 3.70 GB; two-level S
    reads A once and writes 26 x (22 outputs + 4 views of 22 products)."""
    fastmm.set_stage_fusion(False)
    b = fastmm.plan_info("grey424_26_206", 8192, 8192, 8192, 2)["stage_bytes"]
    assert b["S"] == 8 * 30 * (2048 * 4096 + 26 * 512 * 2048)
    assert b["T"] == 8 * 28 * (4096 * 2048 + 26 * 2048 * 512)
    assert b["W"] == 8 * (26 + 16) * (2048 * 2048 + 26 * 512 * 512)
    assert b["LEAF"] == 8 * 676 * (512 * 2048 + 2048 * 512 + 512 * 512)
    fastmm.set_stage_fusion(True)
    b = fastmm.plan_info("grey424_26_206", 8192, 8192, 8192, 2)["stage_bytes"]
    assert b["S"] == 8 * 512 * 2048 * (64 + 26 * 22 + 22 * 4)
    assert b["T"] == 8 * 2048 * 512 * (64 + 26 * 20 + 20 * 6)
    # two-level W: read the 676 level-2 products once, write C's 16 x 16
    # sub-blocks once
    assert b["W"] == 8 * 512 * 512 * (676 + 256)


def test_benchmark_output_format_reads_reference_data():
    """fastmm.benchmark's parser reads the reference's own data files
    (data/README.md:15-27) and its eff_gflops matches the README's formula on
    a published point: Strassen L=1 at 2000^3 took 667.291 ms = 23.97 GFLOP/s."""
    import os
    from fastmm import benchmark as bm
    path = os.path.join(os.path.dirname(__file__), "golden", "matlab_square_sample.txt")
    data = bm.parse(open(path).read())
    assert set(data) == {"MKL_0", "STRASSEN_1", "STRASSEN_2"}
    pt = [p for p in data["STRASSEN_1"] if p[0] == 2000][0]
    assert pt[:4] == (2000, 2000, 2000, 1) and pt[4] == 667.291
    assert abs(bm.eff_gflops(pt) - 23.97) < 0.01
    # round trip of our own output line
    line = "%s_%d = [ 8192 8192 8192 2 14.2 ;  4096 4096 4096 2 2.1 ; ];" % (
        bm.enum_name("grey424_26_206"), 2)
    assert line.startswith("FAST424_26_206_2 = [")
    assert bm.parse(line) == {"FAST424_26_206_2": [(8192, 8192, 8192, 2, 14.2),
                                                    (4096, 4096, 4096, 2, 2.1)]}
    assert bm.enum_name("strassen") == "STRASSEN" and bm.enum_name("MKL") == "MKL"


def test_split_planes_from_stages_only_in_breadth_first_plans():
    # Strassen, 3 levels, 8192^3 FP32 split-bf16: breadth-first, the
    # two-level S/T stage of levels 1-3 writes the leaves' planes; the leaf
    # operands that are sub-blocks of a level-1 operand (4 per side for each
    # of the 7 level-1 nodes) take the preparation pass: 56 x 1024^2 x
    # (4 + 6) bytes.  Depth-first children cut all their planes in the
    # preparation pass.
    args = ("strassen", 8192, 8192, 8192, 3)
    p = fastmm.plan_info(*args, dtype=torch.float32, mode="split_bf16")
    assert p["prep_bytes"] == 56 * 1024 * 1024 * (4 + 6)
    try:
        fastmm.set_stage_fusion(2)
        q = fastmm.plan_info(*args, dtype=torch.float32, mode="split_bf16")
        fastmm.set_stage_fusion(1)
        fastmm.set_schedule(1, 0, True)
        d1 = fastmm.plan_info(*args, dtype=torch.float32, mode="split_bf16")
        fastmm.set_stage_fusion(2)
        d2 = fastmm.plan_info(*args, dtype=torch.float32, mode="split_bf16")
    finally:
        fastmm.set_stage_fusion(1)
        fastmm.set_schedule(0, 0, True)
    assert p["prep_bytes"] < q["prep_bytes"]
    assert d1["dfs_levels"] == 1 and d1["prep_bytes"] == d2["prep_bytes"]
