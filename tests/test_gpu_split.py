"""FP32-accurate leaves from split bfloat16 slices (FM_MODE_SPLIT_BF16).

Pins (tolerances written at each check):
* exact: integers of 9-11 bits need two bf16 slices (8 significand bits
  each) with no remainder, so every kept slice product is exact and, with K
  small enough that every partial sum stays below 2^24, so is the FP32
  accumulation: the result must equal the exact product bit for bit (this
  also pins the instruction's output layout and the K-slot pairing of the
  two operands);
* random uniform(-1,1) and row/column-scaled 10^u inputs: normwise error
  against the FP64 product <= 8 * 2^-24 * sqrt(K)-free bound below, i.e. the
  level of a plain FP32 GEMM, for the classical split product and for
  Strassen with 1 and 2 levels.
"""
import numpy as np
import pytest
import torch

from cases import make_inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a, fm):
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


def _nw(C, D, A, B):
    return float(np.abs(C.astype(np.float64) - D).max()) / (
        float(np.abs(A).max()) * float(np.abs(B).max()) * A.shape[1])


@pytest.mark.parametrize("shape,bits", [((64, 32, 64), 9), ((200, 8, 130), 10),
                                        ((129, 1, 257), 11), ((1, 32, 1), 9),
                                        ((300, 500, 100), 5), ((512, 64, 256), 9)])
def test_split_gemm_exact_on_integers(fm, shape, bits):
    m, k, n = shape
    rng = np.random.default_rng(bits)
    lim = 2 ** bits - 1
    A = rng.integers(-lim, lim + 1, (m, k)).astype(np.float32)
    B = rng.integers(-lim, lim + 1, (k, n)).astype(np.float32)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    C = fm.gemm(_dev(A, fm), _dev(B, fm), mode="split_bf16").cpu().numpy()
    # exact when every partial sum of K terms |a||b| stays below 2^24
    if k * lim * lim < 2 ** 24:
        assert np.array_equal(C.astype(np.float64), exact)
    else:
        assert np.abs(C - exact).max() <= 2.0 ** -23 * lim * lim * k


@pytest.mark.parametrize("alg,steps", [("strassen", 1), ("strassen", 2), ("grey424_26_206", 1)])
def test_split_fast_exact_on_small_integers(fm, alg, steps):
    A, B, C0 = make_inputs(402, 128, 562, 5, "int", np.float32, with_c=True)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    C = fm.fastmatmul(alg, _dev(A, fm), _dev(B, fm), _dev(C0, fm), steps, alpha=2.0, beta=-1.0,
                      mode="split_bf16").cpu().numpy()
    assert np.array_equal(C.astype(np.float64), 2 * exact - C0.astype(np.float64))


def _scaled(m, k, n, seed):
    rng = np.random.default_rng(seed)
    A = rng.random((m, k)) * 2 - 1
    B = rng.random((k, n)) * 2 - 1
    A *= 10.0 ** rng.uniform(-3, 3, (m, 1))
    B *= 10.0 ** rng.uniform(-3, 3, (1, n))
    return np.asfortranarray(A.astype(np.float32)), np.asfortranarray(B.astype(np.float32))


@pytest.mark.parametrize("dist", ["u11", "scaled"])
@pytest.mark.parametrize("alg,steps", [(None, 0), ("strassen", 1), ("strassen", 2)])
def test_split_accuracy_fp32_level(fm, oracle_mod, dist, alg, steps):
    m = k = n = 1024
    if dist == "u11":
        A, B, _ = make_inputs(m, k, n, 31, "u11", np.float32)
    else:
        A, B = _scaled(m, k, n, 32)
    # quad-precision values of a sample of 4096 entries
    rng = np.random.default_rng(7)
    rows, cols = rng.integers(0, m, 4096), rng.integers(0, n, 4096)
    D, absab = oracle_mod.quad_entries(A.astype(np.float64), B.astype(np.float64), rows, cols)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    if alg is None:
        C = fm.gemm(Ad, Bd, mode="split_bf16").cpu().numpy()
        Cp = fm.gemm(Ad, Bd).cpu().numpy()
    else:
        C = fm.fastmatmul(alg, Ad, Bd, None, steps, mode="split_bf16").cpu().numpy()
        Cp = fm.fastmatmul(alg, Ad, Bd, None, steps).cpu().numpy()
    err = np.abs(C[rows, cols].astype(np.float64) - D) / absab
    errp = np.abs(Cp[rows, cols].astype(np.float64) - D) / absab
    if dist == "u11" or steps == 0:
        # componentwise: |C - AB| <= c 2^-24 (|A||B|), c = 128 (1 + levels)
        assert err.max() <= 128 * (1 + steps) * 2.0 ** -24, (err.max(), errp.max())
    # fast algorithms are stable only normwise: with rows/columns scaled by
    # 10^+-3 the small rows lose their digits in S = A11 + A22 whatever the
    # leaf arithmetic; the split leaves must not add to that loss
    assert err.max() <= 4 * errp.max() + 2.0 ** -24, (err.max(), errp.max())
    nw = np.abs(C[rows, cols].astype(np.float64) - D).max() / (np.abs(A).max() * np.abs(B).max() * k)
    nwp = np.abs(Cp[rows, cols].astype(np.float64) - D).max() / (np.abs(A).max() * np.abs(B).max() * k)
    assert nw <= 4 * nwp + 2.0 ** -24


# FmTileCfg: 0 = in-loop split kernel (128 x 128), 24-29 = kernels on
# prepared slice planes (28, 29: fragment reads interleaved with the MFMAs)
INLOOP, PLANE_CFGS = 0, (24, 25, 26, 27, 28, 29, 31)


@pytest.fixture
def _tile_reset(fm):
    yield
    fm.set_gemm_tile(-1)


@pytest.mark.parametrize("cfg", PLANE_CFGS)
def test_plane_kernels_match_inloop_kernel_bits(fm, _tile_reset, cfg):
    """Both kernels form the same six slice products per 32-deep K slice in
    the same order, so they must agree bit for bit, over ragged shapes
    (partial tiles, K not a multiple of 32) and with alpha / beta."""
    for (m, k, n) in [(1, 1, 1), (17, 33, 65), (300, 500, 100), (515, 1031, 263),
                      (256, 64, 384), (129, 2048, 255)]:
        A, B, C0 = make_inputs(m, k, n, 41, "u11", np.float32, with_c=True)
        Ad, Bd = _dev(A, fm), _dev(B, fm)
        fm.set_gemm_tile(INLOOP)
        ref = fm.gemm(Ad, Bd, _dev(C0, fm), 1.5, -0.5, mode="split_bf16").cpu().numpy()
        fm.set_gemm_tile(cfg)
        got = fm.gemm(Ad, Bd, _dev(C0, fm), 1.5, -0.5, mode="split_bf16").cpu().numpy()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (m, k, n)


@pytest.mark.parametrize("alg,steps,shape", [("strassen", 2, (1024, 1024, 1024)),
                                             ("strassen", 2, (1000, 515, 777)),
                                             ("grey424_26_206", 1, (512, 256, 512)),
                                             ("smirnov333_23_139", 1, (384, 384, 384))])
def test_fast_split_planes_match_inloop_bits(fm, _tile_reset, alg, steps, shape):
    """A whole fast multiply with plane leaves equals the one with in-loop
    leaves bit for bit (stages, peeling and W are shared)."""
    m, k, n = shape
    A, B, _ = make_inputs(m, k, n, 43, "u11", np.float32)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    fm.set_gemm_tile(INLOOP)
    ref = fm.fastmatmul(alg, Ad, Bd, None, steps, mode="split_bf16").cpu().numpy()
    fm.set_gemm_tile(-1)
    info = fm.plan_info(alg, m, k, n, steps, dtype=torch.float32, mode="split_bf16")
    assert all(l["cfg"] in PLANE_CFGS for l in info["leaf"])
    got = fm.fastmatmul(alg, Ad, Bd, None, steps, mode="split_bf16").cpu().numpy()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_split_planes_shared_operand(fm, _tile_reset):
    """classical222_8_24 uses each block of A and of B in two products: each
    is prepared once (4 + 4 plane sets, not 16) and both products read it."""
    m = k = n = 512
    info = fm.plan_info("classical222_8_24", m, k, n, 1, dtype=torch.float32, mode="split_bf16")
    plain = fm.plan_info("classical222_8_24", m, k, n, 1, dtype=torch.float32)
    assert info["workspace_bytes"] - plain["workspace_bytes"] == 8 * 256 * 256 * 6
    A, B, _ = make_inputs(m, k, n, 47, "u11", np.float32)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    fm.set_gemm_tile(INLOOP)
    ref = fm.fastmatmul("classical222_8_24", Ad, Bd, None, 1, mode="split_bf16").cpu().numpy()
    fm.set_gemm_tile(-1)
    got = fm.fastmatmul("classical222_8_24", Ad, Bd, None, 1, mode="split_bf16").cpu().numpy()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("tile", [-1, 0])
def test_split_bf16_huge_and_nonfinite_entries(fm, tile):
    """FP32 entries whose rounding to bf16 would overflow (up to FLT_MAX)
    are split with a truncated leading slice, so the split-bf16 product stays
    finite and FP32-accurate where the plain FP32 product is; Inf and NaN
    entries give non-finite results exactly where the plain product has them
    (slice planes: tile -1; in-loop splitting: tile 0)."""
    rng = np.random.default_rng(91)
    m, k, n = 96, 128, 80
    A = (rng.random((m, k)) * 2 - 1).astype(np.float32)
    B = ((rng.random((k, n)) * 2 - 1) * 1e-3).astype(np.float32)
    A[3, 5] = np.float32(3.4028235e38)
    A[7, 9] = -np.float32(3.395e38)
    A[11, 0] = np.float32(3.39e38)
    Ad, Bd = fm.col_major(torch.from_numpy(A).to(DEV)), fm.col_major(torch.from_numpy(B).to(DEV))

    def run(mode, A_, B_):
        P = fm.Plan(None, m, k, n, 0, dtype=torch.float32, mode=mode,
                    opts=dict(gemm_tile=tile) if mode == "split_bf16" else None)
        return P.run(A_, B_, fm.empty(m, n, torch.float32, DEV)).cpu().numpy()

    got, plain = run("split_bf16", Ad, Bd), run("plain", Ad, Bd)
    want = A.astype(np.float64) @ B.astype(np.float64)
    assert np.all(np.isfinite(got)) and np.all(np.isfinite(plain))
    scale = np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
    assert np.all(np.abs(got - want) <= 4 * k * 2.0 ** -24 * scale)
    # Inf and NaN entries
    A2 = A.copy()
    A2[20, 4] = np.inf
    A2[30, 6] = -np.inf
    A2[40, 8] = np.nan
    Ad2 = fm.col_major(torch.from_numpy(A2).to(DEV))
    got2, plain2 = run("split_bf16", Ad2, Bd), run("plain", Ad2, Bd)
    assert np.array_equal(np.isfinite(got2), np.isfinite(plain2))
    assert np.isnan(got2[40]).all()


# Two-level S/T stages writing the leaves' slice planes themselves
# (stage_fusion 1) against the same stages followed by the separate
# preparation pass (stage_fusion 2): bit-identical.  The cases cover a lane
# width of 4 (Strassen, 4 blocks a side), 2 (grey223's B side, 6 blocks) and
# 1 (grey424, 8 blocks; smirnov333, 9 blocks), three levels, and alpha/beta.
@pytest.mark.parametrize("alg,steps,shape,alpha,beta", [
    ("strassen", 2, (1024, 1024, 1024), 1.0, 0.0),
    ("strassen", 2, (2048, 512, 1024), 2.0, -1.0),
    ("strassen", 3, (2048, 256, 2048), 1.0, 0.0),
    ("grey223_11_50", 2, (1024, 512, 2304), 1.0, 0.0),
    ("grey424_26_206", 2, (4096, 256, 4096), 1.0, 0.0),
    ("smirnov333_23_139", 2, (2304, 288, 2304), 1.0, 0.0),
])
def test_split_planes_from_stages_bit_identical(fm, alg, steps, shape, alpha, beta):
    m, k, n = shape
    g = torch.Generator(device=DEV)
    g.manual_seed(m + 7 * k + 13 * n + steps)
    A = fm.col_major(torch.rand((m, k), generator=g, device=DEV) * 2 - 1)
    B = fm.col_major(torch.rand((k, n), generator=g, device=DEV) * 2 - 1)
    C0 = fm.col_major(torch.rand((m, n), generator=g, device=DEV) * 2 - 1)
    # the fused plan prepares only the leaf operands that are sub-blocks of A or B
    try:
        fm.set_stage_fusion(1)
        fused = fm.plan_info(alg, m, k, n, steps, dtype=torch.float32, mode="split_bf16")
        fm.set_stage_fusion(2)
        apart = fm.plan_info(alg, m, k, n, steps, dtype=torch.float32, mode="split_bf16")
    finally:
        fm.set_stage_fusion(1)
    assert fused["prep_bytes"] < apart["prep_bytes"]
    out = []
    for fu in (1, 2):
        C = C0.clone()
        fm.fastmatmul(alg, A, B, C, steps, alpha=alpha, beta=beta, mode="split_bf16",
                      opts=fm.options(stage_fusion=fu))
        out.append(C)
    torch.cuda.synchronize()
    assert torch.equal(out[0], out[1])
    # and accurate: FP32-level normwise error against the FP64 product
    D = alpha * (A.double() @ B.double()) + beta * C0.double()
    nw = float((out[0].double() - D).abs().max()) / k
    assert nw <= 64 * 2.0 ** -24


def test_split_planes_plan_graph_replays(fm):
    """A split-bf16 plan whose last S/T stage writes the slice planes, run
    with a caller workspace and captured into a hipGraph: every replay (with
    eager work between) equals the eager run, which equals the call with the
    planes cut by the separate pass."""
    m = k = n = 1024
    g = torch.Generator(device=DEV)
    g.manual_seed(41)
    A = fm.col_major(torch.rand((m, k), generator=g, device=DEV) * 2 - 1)
    B = fm.col_major(torch.rand((k, n), generator=g, device=DEV) * 2 - 1)
    P = fm.Plan("strassen", m, k, n, 2, dtype=torch.float32, mode="split_bf16")
    ws = torch.empty(max(1, P.workspace_bytes()), dtype=torch.uint8, device=DEV)
    C = fm.empty(m, n, torch.float32, DEV)
    eager = P.run(A, B, C, workspace=ws).clone()
    apart = fm.fastmatmul("strassen", A, B, None, 2, mode="split_bf16",
                          opts=fm.options(stage_fusion=2))
    assert torch.equal(eager, apart)
    s = torch.cuda.Stream(device=DEV)
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=s):
        P.run(A, B, C, workspace=ws)
    for i in range(3):
        C.zero_()
        junk = torch.randn(1000, device=DEV).sum()  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, eager), i
        del junk


def test_split_planes_scaled_bit_identical(fm):
    """The diagonally scaled entry point takes the stage-written planes too:
    bit-identical to the separate preparation pass."""
    m = k = n = 1024
    rng = np.random.default_rng(43)
    A = (rng.uniform(-1, 1, (m, k)) * 10.0 ** rng.uniform(-3, 3, (m, 1))).astype(np.float32)
    B = (rng.uniform(-1, 1, (k, n)) * 10.0 ** rng.uniform(-3, 3, (1, n))).astype(np.float32)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    out = []
    try:
        for fu in (1, 2):
            fm.set_stage_fusion(fu)
            out.append(fm.fastmatmul_scaled("strassen", Ad, Bd, None, 2, mode="split_bf16"))
    finally:
        fm.set_stage_fusion(1)
    torch.cuda.synchronize()
    assert torch.equal(out[0], out[1])
