"""GPU parity: the HIP path (through the C ABI) against the oracle.

Tolerances (FP64 / FP32) are written next to each check:
* integer inputs: bit-exact against the exact product (every intermediate of
  the recursion is a small dyadic rational);
* seeded real inputs (the golden cases): the oracle first reproduces the
  reference's digest (so it IS the reference on this host), then
  max|C_gpu - C_oracle| / (max|A| max|B| K) <= 2 e_ref + 1e-15 (FP64) /
  2 e_ref + 1e-6 (FP32), e_ref being the reference's own normwise error
  against classical GEMM on that case (golden.json, 9e-16 median, 1.2e-14 max
  for FP64), and the GPU's MaxRelativeDiff against classical GEMM is within a
  factor 8 of the reference's own;
* full-size bench workload: errors against the quad oracle on sampled entries
  and a checksum identity C x = A (B x).
"""
import numpy as np
import pytest
import torch

from cases import golden_cases, make_inputs, sha, x_for

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _to_dev(a, fm):
    if a is None:
        return None
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


def _host(t):
    return np.asfortranarray(t.cpu().numpy())


def _normwise(C, D, A, B):
    K = A.shape[1]
    return float(np.abs(C.astype(np.float64) - D.astype(np.float64)).max()) / (
        float(np.abs(A).max()) * float(np.abs(B).max()) * K)


# --------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(1, 1, 1), (17, 33, 65), (64, 64, 64), (128, 128, 128),
                                   (200, 300, 150), (1025, 513, 257), (3, 1000, 2)])
def test_classical_gemm_exact_on_integers(fm, oracle_mod, shape, dtype):
    m, k, n = shape
    A, B, C0 = make_inputs(m, k, n, 3, "int", dtype, with_c=True)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    C = _host(fm.gemm(_to_dev(A, fm), _to_dev(B, fm)))
    assert np.array_equal(C.astype(np.float64), exact)
    C = _host(fm.gemm(_to_dev(A, fm), _to_dev(B, fm), _to_dev(C0, fm), 2.0, -3.0))
    assert np.array_equal(C.astype(np.float64), 2 * exact - 3 * C0.astype(np.float64))


@pytest.mark.parametrize("dtype,tol", [(np.float64, 2e-15), (np.float32, 2e-6)])
def test_classical_gemm_random(fm, oracle_mod, dtype, tol):
    A, B, _ = make_inputs(777, 555, 333, 4, "u11", dtype)
    C = _host(fm.gemm(_to_dev(A, fm), _to_dev(B, fm)))
    D = oracle_mod.quad_matmul(A.astype(np.float64), B.astype(np.float64))
    assert _normwise(C, D, A, B) <= tol


# --------------------------------------------------------------------------
INT_CASES = [
    ("strassen", 1, (257, 500, 55)), ("strassen", 2, (300, 500, 100)),
    ("strassen", 3, (1001, 513, 257)), ("grey424_26_206", 1, (402, 500, 562)),
    ("grey424_26_206", 2, (402, 500, 562)), ("grey424_26_257", 2, (366, 500, 200)),
    ("smirnov333_23_139", 2, (366, 500, 200)), ("smirnov333_23_128", 2, (634, 700, 1000)),
    ("grey423_20_144", 2, (101, 500, 340)), ("grey333_23_152", 2, (600, 500, 671)),
    ("grey333_23_125", 1, (257, 500, 55)), ("grey323_15_89", 2, (600, 500, 671)),
    ("hk323_15_84", 2, (402, 500, 562)), ("grey522_18_99", 2, (1007, 400, 679)),
    ("smirnov336_40_960", 1, (402, 500, 562)), ("classical333_27_81", 1, (100, 100, 100)),
    ("grey252_18_99", 1, (64, 64, 64)), ("strassen", 0, (129, 65, 33)),
    # permuted tables (codegen/convert.py), derived at build time
    ("grey232_11_50", 2, (203, 305, 190)), ("grey224_14_84", 2, (130, 77, 321)),
    ("grey244_26_257", 2, (130, 310, 290)), ("grey334_29_234", 1, (300, 301, 405)),
    ("tichavsky_kovac633_40_960", 1, (610, 300, 310)),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("alg,steps,shape", INT_CASES)
def test_fast_exact_on_integers(fm, alg, steps, shape, dtype):
    m, k, n = shape
    A, B, C0 = make_inputs(m, k, n, 21, "int", dtype, with_c=True)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    if dtype == np.float32 and np.abs(exact).max() * 6 > 2 ** 24:
        pytest.skip("not exact in FP32 at this size")
    Ad, Bd = _to_dev(A, fm), _to_dev(B, fm)
    C = _host(fm.fastmatmul(alg, Ad, Bd, None, steps))
    assert np.array_equal(C.astype(np.float64), exact)
    # alpha/beta, with the reference's steps == 0 rule (beta ignored)
    C = _host(fm.fastmatmul(alg, Ad, Bd, _to_dev(C0, fm), steps, alpha=3.0, beta=-2.0))
    want = 3 * exact - (2 * C0.astype(np.float64) if steps > 0 else 0)
    assert np.array_equal(C.astype(np.float64), want)


DERIVED = ["grey232_11_50", "grey223_11_50", "grey242_14_84", "grey224_14_84", "grey225_18_99",
           "grey244_26_257", "grey334_29_234", "grey233_15_103", "hk233_15_94",
           "tichavsky_kovac633_40_960"]


@pytest.mark.parametrize("alg", DERIVED)
def test_permuted_algorithms_vs_oracle(fm, oracle_mod, alg):
    """Permuted tables against the oracle running the same table: normwise
    within twice the oracle's own error against classical GEMM + 1e-15."""
    A, B, C0 = make_inputs(301, 257, 333, 41, "u11", np.float64, with_c=True)
    want = oracle_mod.fastmatmul(alg, A, B, C0, 2, 0.5, 2.0)
    classical = 0.5 * (A @ B) + 2.0 * C0
    got = _host(fm.fastmatmul(alg, _to_dev(A, fm), _to_dev(B, fm), _to_dev(C0, fm), 2,
                              alpha=0.5, beta=2.0))
    e_ref = _normwise(want, classical, A, B)
    assert _normwise(got, want, A, B) <= 2 * e_ref + 1e-15


def _golden_params():
    return [pytest.param(c, id=c[0]) for c in golden_cases()]


@pytest.mark.parametrize("case", _golden_params())
def test_fast_matches_oracle(fm, oracle_mod, golden, case):
    (cid, alg, m, k, n, steps, alpha, beta, dist, dt, seed) = case
    rec = {r["id"]: r for r in golden["cases"]}[cid]
    dtype = np.float64 if dt == "f64" else np.float32
    A, B, C0 = make_inputs(m, k, n, seed, dist, dtype, with_c=beta != 0.0)
    x = x_for(alg, steps)
    Co = oracle_mod.fastmatmul(alg, A, B, C0, steps, alpha, beta, x)
    if oracle_mod.reproducible():
        assert sha(Co) == rec["sha256"], "oracle no longer reproduces the reference"
    C = _host(fm.fastmatmul(alg, _to_dev(A, fm), _to_dev(B, fm), _to_dev(C0, fm), steps, x,
                            alpha, beta))
    # Two evaluations of the same algorithm differ by at most the sum of their
    # errors: bound the GPU-oracle gap by twice the reference's own normwise
    # error against classical GEMM (golden.json) plus a floor.
    floor = 1e-15 if dtype == np.float64 else 1e-6
    scale = max(1.0, abs(alpha)) * (1 + abs(beta))
    tol = 2 * rec["normwise"] + floor * scale
    if "approx" in alg:
        tol = 1e-6 if steps == 1 else 1e-3
    assert _normwise(C, Co, A, B) <= tol
    D = oracle_mod.gemm(A.astype(np.float64), B.astype(np.float64),
                        None if C0 is None else C0.astype(np.float64), alpha, beta)
    err = oracle_mod.max_rel_diff(C.astype(np.float64), D)
    assert err <= 8 * rec["max_rel_diff"] + (1e-15 if dtype == np.float64 else 1e-6)


# --------------------------------------------------------------------------
def test_matrix_api_mirror(fm, oracle_mod):
    A, B, C0 = make_inputs(257, 500, 55, 12, "int", with_c=True)
    Am, Bm = fm.Matrix.from_numpy(A), fm.Matrix.from_numpy(B)
    Cm = fm.Matrix(257, 55)
    ms = fm.strassen.FastMatmul(Am, Bm, Cm, 1)
    assert ms > 0
    assert np.array_equal(Cm.numpy(), A @ B)
    ms = fm.RunAlgorithm("FAST424_26_206", Am, Bm, Cm, 2)
    assert np.array_equal(Cm.numpy(), A @ B)
    # MKL entry: classical through the same wrapper with 0 steps
    fm.RunAlgorithm("MKL", Am, Bm, Cm, 2)
    assert np.array_equal(Cm.numpy(), A @ B)
    # views: multiply block (2,1) of a 2x2 split (linalg/linalg.hpp:97-102)
    sub = Am.Subblock(2, 2, 2, 1)
    assert sub.m() == 128 and sub.n() == 250 and sub.stride() == 257
    Cs = fm.Matrix(128, 55)
    fm.strassen.FastMatmul(sub, Bm.Submatrix(0, 0, 250, 55), Cs, 1)
    assert np.array_equal(Cs.numpy(), A[128:256, :250] @ B[:250, :])
    # no caller state changes (unlike algorithms/strassen.hpp:669)
    assert Am.multiplier() == 1 and Cm.multiplier() == 1
    Cm.set_multiplier(2.0)
    fm.MatMul(Am, Bm, Cm)
    assert np.array_equal(Cm.numpy(), 2 * (A @ B))


def test_explicit_workspace_and_stream(fm):
    A, B, _ = make_inputs(512, 512, 512, 13, "int")
    Ad, Bd = _to_dev(A, fm), _to_dev(B, fm)
    need = fm.workspace_bytes("grey424_26_206", 512, 512, 512, 2)
    ws = torch.empty(need, dtype=torch.uint8, device=DEV)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        C = fm.fastmatmul("grey424_26_206", Ad, Bd, None, 2, workspace=ws)
    s.synchronize()
    assert np.array_equal(_host(C), A @ B)
    small = torch.empty(need // 2, dtype=torch.uint8, device=DEV)
    with pytest.raises(Exception, match="workspace"):
        fm.fastmatmul("grey424_26_206", Ad, Bd, None, 2, workspace=small)


def test_views_with_large_leading_dimension(fm):
    """Operands that are views into bigger column-major buffers (lda > m)."""
    A, B, _ = make_inputs(600, 700, 500, 14, "int")
    big = torch.zeros((800, 1000), dtype=torch.float64, device=DEV).t()  # 1000 x 800, ld 1000
    big[:600, :700] = torch.from_numpy(A).to(DEV)
    Av = big[:600, :700]
    assert Av.stride() == (1, 1000)
    C = _host(fm.fastmatmul("smirnov333_23_139", Av, _to_dev(B, fm), None, 2))
    assert np.array_equal(C, A @ B)


def test_bench_workload_full_size(fm, oracle_mod):
    """grey424_26_206, 2 steps, FP64, 8192^3, uniform(-1,1): checked at full
    size through sampled quad-precision entries and C x = A (B x)."""
    n = 8192
    g = torch.Generator(device=DEV)
    g.manual_seed(0xFA57)
    A = fm.col_major(torch.rand((n, n), generator=g, device=DEV, dtype=torch.float64) * 2 - 1)
    B = fm.col_major(torch.rand((n, n), generator=g, device=DEV, dtype=torch.float64) * 2 - 1)
    C = fm.fastmatmul("grey424_26_206", A, B, None, 2)
    x = torch.rand((n, 1), generator=g, device=DEV, dtype=torch.float64)
    lhs = C @ x
    rhs = A @ (B @ x)
    assert float((lhs - rhs).abs().max()) <= 1e-10 * n
    rng = np.random.default_rng(1)
    rows = rng.integers(0, n, 64)
    cols = rng.integers(0, n, 64)
    Ah, Bh = _host(A), _host(B)
    q, absab = oracle_mod.quad_entries(Ah, Bh, rows, cols)
    got = _host(C)[rows, cols]
    # componentwise: |C - AB| <= c * u * (|A||B|)_ij with c well below K
    assert np.all(np.abs(got - q) <= 64 * 2 ** -53 * absab)


@pytest.mark.parametrize("steps", [1, 2])
def test_tall_skinny_workload_full_size(fm, oracle_mod, steps):
    """grey424_26_206 on the tall-skinny workload, 131072 x 2400 x 2400 FP64,
    uniform(-1,1) (benchmarks/matmul_benchmarks.cpp:275-326,
    data/all_sequential_algs/tssquare.2400): leaves of 32768 x 1200 x 600
    (one level) or 8192 x 600 x 150 (two), checked at full size through
    sampled quad-precision entries and C x = A (B x)."""
    m, k = 131072, 2400
    g = torch.Generator(device=DEV)
    g.manual_seed(0xFA57 + 2)
    A = fm.col_major(torch.rand((m, k), generator=g, device=DEV, dtype=torch.float64) * 2 - 1)
    B = fm.col_major(torch.rand((k, k), generator=g, device=DEV, dtype=torch.float64) * 2 - 1)
    C = fm.fastmatmul("grey424_26_206", A, B, None, steps)
    x = torch.rand((k, 1), generator=g, device=DEV, dtype=torch.float64)
    lhs = C @ x
    rhs = A @ (B @ x)
    assert float((lhs - rhs).abs().max()) <= 1e-10 * k
    rng = np.random.default_rng(2 + steps)
    rows = rng.integers(0, m, 96)
    cols = rng.integers(0, k, 96)
    Ah, Bh = _host(A), _host(B)
    q, absab = oracle_mod.quad_entries(Ah, Bh, rows, cols)
    got = _host(C)[rows, cols]
    # componentwise: |C - AB| <= c * u * (|A||B|)_ij with c well below K
    assert np.all(np.abs(got - q) <= 64 * 2 ** -53 * absab)


def test_cpp_header_program(tmp_path):
    """The C++ drop-in header, run on the GPU: exact products through
    strassen/grey424_26_206/smirnov333_23_139/grey423_20_144, views included."""
    import subprocess
    from test_capi import _build_shim_example
    exe = _build_shim_example(tmp_path)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout


def test_concurrent_streams_use_private_workspaces(fm):
    """Calls without a caller workspace on two streams at once: each
    (device, stream) has its own cached workspace, so interleaved calls give
    the sequential results bit for bit."""
    n = 2048
    A1, B1, _ = make_inputs(n, n, n, 61, "u11")
    A2, B2, _ = make_inputs(n, n, n, 62, "u11")
    A1, B1, A2, B2 = (_to_dev(x, fm) for x in (A1, B1, A2, B2))
    ref1 = fm.fastmatmul("grey424_26_206", A1, B1, None, 2)
    ref2 = fm.fastmatmul("grey424_26_206", A2, B2, None, 2)
    C1, C2 = fm.empty(n, n), fm.empty(n, n)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(6):
        with torch.cuda.stream(s1):
            fm.fastmatmul("grey424_26_206", A1, B1, C1, 2)
        with torch.cuda.stream(s2):
            fm.fastmatmul("grey424_26_206", A2, B2, C2, 2)
    torch.cuda.synchronize()
    assert torch.equal(C1, ref1) and torch.equal(C2, ref2)


def test_plan_cache_eviction(fm):
    """alpha is part of a plan, so 300 calls with distinct alphas exceed the
    256-plan cache; evicted plans are rebuilt and every result stays exact."""
    A, B, _ = make_inputs(96, 64, 80, 63, "int")
    Ad, Bd = _to_dev(A, fm), _to_dev(B, fm)
    exact = A @ B
    for i in range(300):
        C = fm.fastmatmul("strassen", Ad, Bd, None, 1, alpha=float(i + 1)).cpu().numpy()
        assert np.array_equal(C, (i + 1) * exact), i
    C = fm.fastmatmul("strassen", Ad, Bd, None, 1, alpha=1.0).cpu().numpy()
    assert np.array_equal(C, exact)
