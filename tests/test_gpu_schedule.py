"""Recursion schedules on the GPU: depth-first levels, the workspace limit and
the leaf remainder split against the default breadth-first plan.

Every schedule runs the same element-wise operations (the stage expressions
and whole leaf products with the same instructions per output element);
only launch order, launch grouping and workspace placement change.  The
breadth-first result is the reference here and the tolerance is zero.
"""
import numpy as np
import pytest
import torch

from cases import make_inputs

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _dev(a, fm):
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


@pytest.fixture(autouse=True)
def _reset(fm):
    yield
    fm.set_schedule(0, 0, True)


def _run(fm, alg, A, B, C0, steps, sched, alpha=1.0, beta=0.0, mode="plain", workspace=None):
    fm.set_schedule(*sched)
    C = None if C0 is None else _dev(C0, fm)
    fm.profile_enable(True)
    fm.profile_read()
    try:
        C = fm.fastmatmul(alg, A, B, C, num_steps=steps, alpha=alpha, beta=beta, mode=mode,
                          workspace=workspace)
        torch.cuda.synchronize()
        prof = fm.profile_read()
    finally:
        fm.profile_enable(False)
    return C, prof


CASES = [
    ("strassen", (256, 256, 256), 2, np.float64),
    ("strassen", (257, 255, 263), 3, np.float64),      # peeling at every level
    ("strassen", (512, 384, 256), 3, np.float32),
    ("grey424_26_206", (512, 256, 512), 2, np.float64),
    ("smirnov333_23_139", (243, 243, 243), 2, np.float64),
    ("grey234_20_144", (130, 290, 250), 2, np.float32),
]


@pytest.mark.parametrize("dfs", [1, 2])
@pytest.mark.parametrize("alg,shape,steps,dtype", CASES)
def test_depth_first_bit_identical(fm, alg, shape, steps, dtype, dfs):
    m, k, n = shape
    A, B, C0 = make_inputs(m, k, n, 21, "u11", dtype, with_c=True)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    Cb, pb = _run(fm, alg, Ad, Bd, C0, steps, (0, 0, True), alpha=1.5, beta=-0.5)
    Cd, pd = _run(fm, alg, Ad, Bd, C0, steps, (dfs, 0, True), alpha=1.5, beta=-0.5)
    assert torch.equal(Cb, Cd), float((Cb - Cd).abs().max())
    # at least one leaf launch per sub-plan below the depth-first levels
    R = {"strassen": 7, "grey424_26_206": 26, "smirnov333_23_139": 23, "grey234_20_144": 20}[alg]
    assert pd["LEAF"][1] >= R ** min(dfs, steps)
    assert pb["LEAF"][1] <= 2


@pytest.mark.parametrize("mode,dtype", [("split_bf16", np.float32), ("ozaki_int8", np.float64)])
def test_depth_first_leaf_modes(fm, mode, dtype):
    A, B, _ = make_inputs(256, 256, 256, 22, "u11", dtype)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    Cb, _ = _run(fm, "strassen", Ad, Bd, None, 2, (0, 0, True), mode=mode)
    Cd, pd = _run(fm, "strassen", Ad, Bd, None, 2, (1, 0, True), mode=mode)
    assert torch.equal(Cb, Cd)
    assert pd["LEAF"][1] == 7


def test_workspace_limit_with_caller_workspace(fm):
    """A workspace limit below the breadth-first need selects depth-first
    levels; a caller-owned buffer of exactly the reported size suffices."""
    A, B, _ = make_inputs(512, 256, 512, 23, "u11", np.float64)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    bfs = fm.workspace_bytes("grey424_26_206", 512, 256, 512, 2)
    Cb, _ = _run(fm, "grey424_26_206", Ad, Bd, None, 2, (0, 0, True))
    fm.set_schedule(0, bfs - 1)
    need = fm.workspace_bytes("grey424_26_206", 512, 256, 512, 2)
    assert need < bfs
    ws = torch.empty(need, dtype=torch.uint8, device=DEV)
    Cd, pd = _run(fm, "grey424_26_206", Ad, Bd, None, 2, (0, bfs - 1, True), workspace=ws)
    assert torch.equal(Cb, Cd)
    assert pd["LEAF"][1] == 26


def test_remainder_split_bit_identical(fm):
    # 676 leaves of 256 x 1024 x 256: 2704 tiles of 128 x 128 (one work group
    # per CU), i.e. 10 full rounds on 256 CUs (640 products) and 36 products
    # in a second launch
    info = fm.plan_info("grey424_26_206", 4096, 4096, 4096, 2)
    assert [l["products"] for l in info["leaf"]] == [640, 36]
    A, B, _ = make_inputs(4096, 4096, 4096, 24, "u11", np.float64)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    Cs, ps = _run(fm, "grey424_26_206", Ad, Bd, None, 2, (0, 0, True))
    Cu, pu = _run(fm, "grey424_26_206", Ad, Bd, None, 2, (0, 0, False))
    assert (ps["LEAF"][1], pu["LEAF"][1]) == (2, 1)
    assert torch.equal(Cs, Cu)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_whole_tile_carve_out_bit_identical(fm, dtype):
    # classical 8191 x 512 x 4100 (FP32: 12286 x 512 x 4100) with an aligned
    # leading dimension: M is not a multiple of the 16-byte vector width, so
    # the spread kernel's edge variant cannot take it and the product is cut
    # into a whole-tile core and row / column strips, against one
    # edge-kernel launch
    vl = 2 if dtype == np.float64 else 4
    m = 8191 if dtype == np.float64 else 12286
    k, n = 512, 4100
    A, B, _ = make_inputs(m, k, n, 25, "u11", dtype)
    pad = np.zeros((m + vl - m % vl, k), dtype=dtype)
    pad[:m] = A
    Ad = _dev(pad, fm)[:m]
    assert Ad.stride(1) % vl == 0
    Bd = _dev(B, fm)
    Cs, ps = _run(fm, "strassen", Ad, Bd, None, 0, (0, 0, True))
    Cu, pu = _run(fm, "strassen", Ad, Bd, None, 0, (0, 0, False))
    assert (ps["LEAF"][1], pu["LEAF"][1]) == (2, 1)
    assert torch.equal(Cs, Cu)


@pytest.mark.parametrize("tiles", [(34, 30), (32, 16), (33, 19)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spread_edge_launch_bit_identical(fm, dtype, tiles):
    # strassen L=1 on 4200 x 512 x 4200: 7 leaves of 2100 x 256 x 2100 with
    # whole K slices and aligned columns but partial tiles, on a spread
    # tile forced (its edge variant: 128 x 96, 128 x 128, 128 x 256) against
    # a predicated edge kernel of the same tile forced: the same bits
    spread, edge = tiles
    m, k, n = 4200, 512, 4200
    A, B, _ = make_inputs(m, k, n, 25, "u11", dtype)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    out = []
    try:
        for cfg in (spread, edge):
            fm.set_gemm_tile(cfg)
            C, prof = _run(fm, "strassen", Ad, Bd, None, 1, (0, 0, True))
            assert prof["LEAF"][1] == 1
            out.append(C)
    finally:
        fm.set_gemm_tile(-1)
    assert torch.equal(out[0], out[1])
