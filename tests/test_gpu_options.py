"""Per-call options and reusable plans on the GPU (fm_fastmatmul_ex,
fm_plan_create / fm_plan_run / fm_plan_destroy).

Every option changes only launch order, grouping, tile shape or workspace
placement, never the element-wise operations, so each result must equal the
default call's bit for bit (tolerance zero); on integer inputs both must equal
the exact product.
"""
import threading

import numpy as np
import pytest
import torch

from cases import make_inputs

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _dev(a, fm):
    return fm.col_major(torch.from_numpy(np.asfortranarray(a)).to(DEV))


OPTS = [
    dict(dfs_levels=1),
    dict(dfs_levels=2, split_remainder=0),
    dict(stage_fusion=0),
    dict(gemm_tile=1),                       # 64 x 64 leaf tiles
    dict(gemm_tile=12, stage_fusion=0),      # 128 x 64, 8 waves
    dict(dfs_levels=-1, workspace_limit=16 << 20),  # automatic cut-off: one depth-first level
]


@pytest.mark.parametrize("opts", OPTS, ids=[str(o) for o in OPTS])
def test_options_bit_identical(fm, opts):
    A, B, C0 = make_inputs(402, 500, 562, 31, "u11", np.float64, with_c=True)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    ref = fm.fastmatmul("grey424_26_206", Ad, Bd, _dev(C0, fm), 2, alpha=1.5, beta=-0.5)
    C = fm.fastmatmul("grey424_26_206", Ad, Bd, _dev(C0, fm), 2, alpha=1.5, beta=-0.5,
                      opts=opts)
    torch.cuda.synchronize()
    assert torch.equal(ref, C), float((ref - C).abs().max())


def test_options_do_not_touch_defaults(fm):
    before = fm.options()
    A, B, _ = make_inputs(256, 256, 256, 32, "int", np.float64)
    C = fm.fastmatmul("strassen", _dev(A, fm), _dev(B, fm), None, 2,
                      opts=dict(dfs_levels=1, gemm_tile=1))
    assert np.array_equal(C.cpu().numpy(), A @ B)
    after = fm.options()
    for f, _ in type(before)._fields_:
        assert getattr(before, f) == getattr(after, f), f


def test_two_threads_two_streams_different_options(fm):
    """Two host threads, each on its own stream with its own options, run
    interleaved multiplies; every result is exact (integer inputs)."""
    A1, B1, _ = make_inputs(512, 256, 512, 33, "int", np.float64)
    A2, B2, _ = make_inputs(300, 500, 100, 34, "int", np.float64)
    exact1, exact2 = A1 @ B1, A2 @ B2
    jobs = [("grey424_26_206", _dev(A1, fm), _dev(B1, fm), exact1, dict(dfs_levels=1, gemm_tile=1)),
            ("strassen", _dev(A2, fm), _dev(B2, fm), exact2, dict(stage_fusion=0, split_remainder=0))]
    torch.cuda.synchronize()
    errors = []

    def work(alg, Ad, Bd, exact, opts):
        try:
            s = torch.cuda.Stream(device=DEV)
            with torch.cuda.stream(s):
                outs = [fm.fastmatmul(alg, Ad, Bd, None, 2, opts=opts) for _ in range(6)]
            s.synchronize()
            for C in outs:
                if not np.array_equal(C.cpu().numpy(), exact):
                    errors.append(alg)
        except Exception as e:  # reported below, in the test's thread
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=j) for j in jobs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_plan_matches_call(fm, dtype):
    m, k, n = 257, 300, 263
    A, B, C0 = make_inputs(m, k, n, 35, "u11", dtype, with_c=True)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    ref = fm.fastmatmul("strassen", Ad, Bd, _dev(C0, fm), 2, alpha=2.0, beta=0.5)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    P = fm.Plan("strassen", m, k, n, 2, dtype=tdt, alpha=2.0, beta=0.5)
    for _ in range(2):
        C = P.run(Ad, Bd, _dev(C0, fm))
        assert torch.equal(ref, C)
    # classical plan, beta honoured
    Pc = fm.Plan(None, m, k, n, 0, dtype=tdt, alpha=2.0, beta=0.5)
    C = Pc.run(Ad, Bd, _dev(C0, fm))
    G = fm.gemm(Ad, Bd, _dev(C0, fm), 2.0, 0.5)
    assert torch.equal(G, C)


def test_plan_graph_replays(fm):
    """A plan run with a caller workspace captured into a hipGraph: every
    replay (with eager work and synchronisations between) equals the eager
    result."""
    m, k, n = 512, 512, 512
    A, B, _ = make_inputs(m, k, n, 36, "u11", np.float64)
    Ad, Bd = _dev(A, fm), _dev(B, fm)
    P = fm.Plan("grey424_26_206", m, k, n, 2)
    ws = torch.empty(max(1, P.workspace_bytes()), dtype=torch.uint8, device=DEV)
    C = fm.empty(m, n, torch.float64, DEV)
    eager = P.run(Ad, Bd, C, workspace=ws).clone()
    s = torch.cuda.Stream(device=DEV)
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        P.run(Ad, Bd, C, workspace=ws)
    for i in range(4):
        C.zero_()
        junk = torch.randn(1000, device=DEV).sum()  # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, eager), i
        del junk


def test_plan_wrong_shape_rejected(fm):
    P = fm.Plan("strassen", 64, 64, 64, 1)
    A = fm.empty(64, 32, torch.float64, DEV)
    B = fm.empty(32, 64, torch.float64, DEV)
    with pytest.raises(ValueError):
        P.run(A, B, fm.empty(64, 64, torch.float64, DEV))


def test_dfs_auto_under_limit(fm):
    bfs = fm.workspace_bytes("grey424_26_206", 1024, 512, 1024, 2)
    P = fm.Plan("grey424_26_206", 1024, 512, 1024, 2, opts=dict(dfs_levels=-1, workspace_limit=bfs - 1))
    assert P.workspace_bytes() < bfs
    P2 = fm.Plan("grey424_26_206", 1024, 512, 1024, 2, opts=dict(dfs_levels=-1))
    assert P2.workspace_bytes() == bfs  # fits a quarter of the free memory: breadth-first
