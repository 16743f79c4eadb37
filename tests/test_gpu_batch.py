"""Batched uneven multiplies (fm_batch_*): autotuned algorithm/depth per
problem, one grouped leaf launch, optional hipGraph replay.

Pins: integer inputs give every problem's exact product (bit-exact), for
explicit and autotuned choices, through graph replays (run several times
with host work in between, every replay checked); uniform(-1,1) FP32 inputs
stay within FP32-level error of the FP64 product."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _problems(fm, shapes, dtype, dist, seed):
    rng = np.random.default_rng(seed)
    probs, host = [], []
    for (m, k, n) in shapes:
        if dist == "int":
            A = rng.integers(-8, 9, (m, k)).astype(np.float64)
            B = rng.integers(-8, 9, (k, n)).astype(np.float64)
        else:
            A = rng.random((m, k)) * 2 - 1
            B = rng.random((k, n)) * 2 - 1
        C0 = rng.integers(-8, 9, (m, n)).astype(np.float64)
        t = lambda x: fm.col_major(torch.from_numpy(np.asfortranarray(x)).to(DEV, dtype))
        probs.append({"A": t(A), "B": t(B), "C": t(C0)})
        host.append((A, B, C0))
    return probs, host


def test_batch_explicit_and_autotuned_exact(fm):
    shapes = [(576, 512, 640), (704, 1024, 512), (512, 576, 960), (1088, 512, 512),
              (640, 640, 640), (512, 2048, 576)]
    probs, host = _problems(fm, shapes, torch.float64, "int", 1)
    probs[0].update(alg="strassen", num_steps=2)
    probs[1].update(alg="grey423_20_144", num_steps=1, alpha=2.0, beta=-1.0)
    probs[2].update(alg="smirnov333_23_139", num_steps=2)
    probs[3].update(alg=None, num_steps=0, beta=0.5)  # depth 0: the classical product
    b = fm.Batch(probs, graph=False)
    b.run()
    torch.cuda.synchronize()
    for i, (p, (A, B, C0)) in enumerate(zip(probs, host)):
        alpha, beta = p.get("alpha", 1.0), p.get("beta", 0.0)
        alg, steps = b.choice(i)
        want = alpha * (A @ B) + beta * C0
        got = p["C"].cpu().numpy()
        assert np.array_equal(got, want), (i, alg, steps)


def test_batch_graph_replays(fm):
    shapes = [(512 + 64 * i, 512 + 128 * (i % 3), 512 + 64 * (i % 5)) for i in range(12)]
    probs, host = _problems(fm, shapes, torch.float64, "int", 2)
    b = fm.Batch(probs, graph=True)
    for rep in range(3):
        for p in probs:
            p["C"].fill_(7.0)
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        _ = np.ones(1000).sum()  # host work between replays
        for p, (A, B, C0) in zip(probs, host):
            assert np.array_equal(p["C"].cpu().numpy(), A @ B)


def test_batch_autotuner_choices_and_fp32_accuracy(fm, oracle_mod):
    rng = np.random.default_rng(3)
    vals = list(range(512, 4097, 64))
    shapes = [tuple(int(v) for v in rng.choice(vals, 3)) for _ in range(16)]
    probs, host = _problems(fm, shapes, torch.float32, "u11", 4)
    b = fm.Batch(probs, graph=True)
    b.run()
    torch.cuda.synchronize()
    chosen = [b.choice(i) for i in range(len(probs))]
    grey20 = {"grey%s_20_144" % o for o in ("423", "432", "243", "234", "324", "342")}
    assert {a for a, s in chosen} <= {"classical", "strassen", "smirnov333_23_139"} | grey20
    for (p, (A, B, C0), (alg, st)) in zip(probs, host, chosen):
        m, n = A.shape[0], B.shape[1]
        rows, cols = rng.integers(0, m, 256), rng.integers(0, n, 256)
        q, absab = oracle_mod.quad_entries(A.astype(np.float32).astype(np.float64),
                                           B.astype(np.float32).astype(np.float64), rows, cols)
        got = p["C"].cpu().numpy()[rows, cols].astype(np.float64)
        assert np.all(np.abs(got - q) <= 64 * (1 + st) * 2.0 ** -24 * absab), (alg, st)



def test_batch_autotuner_picks_fast_form_where_it_pays(fm, oracle_mod):
    # FP64: one 8192^3 problem, where a fast algorithm's saved products
    # outweigh its stage bytes on this GPU, among 512^3 ones, where they do
    # not: the cost model picks a fast form for the large problem only
    shapes = [(8192, 8192, 8192)] + [(512, 512, 512)] * 3
    probs, host = _problems(fm, shapes, torch.float64, "u11", 5)
    b = fm.Batch(probs, graph=False)
    b.run()
    torch.cuda.synchronize()
    chosen = [b.choice(i) for i in range(len(probs))]
    assert chosen[0][0] != "classical" and chosen[0][1] >= 1, chosen[0]
    assert all(c == ("classical", 0) for c in chosen[1:]), chosen
    rng = np.random.default_rng(6)
    for (p, (A, B, C0), (alg, st)) in zip(probs, host, chosen):
        m, n = A.shape[0], B.shape[1]
        rows, cols = rng.integers(0, m, 128), rng.integers(0, n, 128)
        q, absab = oracle_mod.quad_entries(A, B, rows, cols)
        got = p["C"].cpu().numpy()[rows, cols]
        assert np.all(np.abs(got - q) <= 64 * (1 + st) * 2.0 ** -53 * absab), (alg, st)


def test_batch_split_planes_match_single_calls(fm):
    """split_bf16 batches: one preparation launch for every problem's
    operands, one grouped plane launch.  The plane kernels compute the
    same slice products in the same order whatever the tile, so every
    problem equals its own single call bit for bit (explicit algorithms with
    peeling, classical, alpha/beta), also through graph replays."""
    shapes = [(576, 512, 640), (700, 1000, 515), (512, 576, 960), (1088, 520, 512),
              (640, 640, 640)]
    probs, host = _problems(fm, shapes, torch.float32, "u11", 5)
    probs[0].update(alg="strassen", num_steps=2)
    probs[1].update(alg="grey423_20_144", num_steps=1, alpha=2.0, beta=-1.0)
    probs[2].update(alg="smirnov333_23_139", num_steps=1)
    probs[3].update(alg=None, num_steps=0, beta=0.5)
    probs[4].update(alg="strassen", num_steps=1)
    C0s = [p["C"].clone() for p in probs]
    want = []
    for p, C0 in zip(probs, C0s):
        C = C0.clone()
        alpha, beta = p.get("alpha", 1.0), p.get("beta", 0.0)
        if p.get("alg"):
            fm.fastmatmul(p["alg"], p["A"], p["B"], C, p["num_steps"], alpha=alpha, beta=beta,
                          mode="split_bf16")
        else:  # a batch's classical problem honours beta, as gemm does
            fm.gemm(p["A"], p["B"], C, alpha, beta, mode="split_bf16")
        want.append(C.cpu().numpy())
    b = fm.Batch(probs, mode="split_bf16", graph=True)
    for rep in range(2):
        for p, C0 in zip(probs, C0s):
            p["C"].copy_(C0)
        torch.cuda.synchronize()
        b.run()
        torch.cuda.synchronize()
        for i, (p, w) in enumerate(zip(probs, want)):
            got = p["C"].cpu().numpy()
            assert np.array_equal(got.view(np.uint32), w.view(np.uint32)), (rep, i, b.choice(i))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_grouped_stages_match_single_calls(fm, dtype):
    """The S/T/W stages of every problem that shares an algorithm and a
    recursion position run as one grouped launch (per-node shapes): each
    problem still equals its own single call bit for bit, with peeling,
    two-level stages, alpha/beta and three depths, and the batch issues far
    fewer stage launches than its problems have levels."""
    shapes = [(512 + 64 * i, 576 + 128 * (i % 3), 640 + 64 * (i % 4)) for i in range(10)]
    probs, host = _problems(fm, shapes, dtype, "u11", 6)
    algs = [("strassen", 2), ("strassen", 1), ("grey423_20_144", 2), ("strassen", 2),
            ("grey423_20_144", 1), ("smirnov333_23_139", 2), ("smirnov333_23_139", 1),
            ("strassen", 3), ("grey423_20_144", 2), ("strassen", 2)]
    for i, (p, (a, st)) in enumerate(zip(probs, algs)):
        p.update(alg=a, num_steps=st, alpha=1.0 + 0.25 * (i % 3), beta=(-0.5 if i % 2 else 0.0))
    C0s = [p["C"].clone() for p in probs]
    want = []
    for p, C0 in zip(probs, C0s):
        C = C0.clone()
        fm.fastmatmul(p["alg"], p["A"], p["B"], C, p["num_steps"], alpha=p["alpha"], beta=p["beta"])
        want.append(C)
    b = fm.Batch(probs, graph=False)
    fm.profile_read()
    fm.profile_enable(True)
    try:
        b.run()
        torch.cuda.synchronize()
    finally:
        fm.profile_enable(False)
    prof = fm.profile_read()
    for i, (p, w) in enumerate(zip(probs, want)):
        assert torch.equal(p["C"], w), (i, p["alg"], p["num_steps"])
    alone = 0  # stage launches the problems issue as single calls
    for p, (a, st) in zip(probs, algs):
        m, k = p["A"].shape
        info = fm.plan_info(a, m, k, p["B"].shape[1], st, dtype=dtype)
        alone += sum(info["launches"][s] for s in ("S", "T", "W"))
    grouped = prof["S"][1] + prof["T"][1] + prof["W"][1]
    assert grouped < alone, (grouped, alone)
    # problems of one algorithm, depth and beta class share every stage launch
    same = [dict(p, alg="strassen", num_steps=2, beta=0.0) for p in probs]
    b2 = fm.Batch(same, graph=False)
    fm.profile_enable(True)
    try:
        b2.run()
        torch.cuda.synchronize()
    finally:
        fm.profile_enable(False)
    prof = fm.profile_read()
    m, k = probs[0]["A"].shape
    one = fm.plan_info("strassen", m, k, probs[0]["B"].shape[1], 2, dtype=dtype)["launches"]
    assert all(prof[s][1] == one[s] for s in ("S", "T", "W")), (prof, one)


def test_batch_full_workload_256_uneven(fm, oracle_mod):
    """The full batched workload: 256 independent FP32 multiplies, M, K, N
    each uniform on {512, 576, ..., 4096}, uniform(-1,1), autotuned,
    captured in a hipGraph.  Every problem is checked on 32 sampled entries
    against the quad-precision oracle: |C - AB| <= 64 (1 + depth) 2^-24
    (|A||B|)_ij."""
    rng = np.random.default_rng(0xC3)
    vals = np.arange(512, 4097, 64)
    shapes = [tuple(int(v) for v in rng.choice(vals, 3)) for _ in range(256)]
    g = torch.Generator(device=DEV)
    g.manual_seed(0xC3)
    probs = []
    for (m, k, n) in shapes:
        A = fm.col_major(torch.rand((m, k), generator=g, device=DEV, dtype=torch.float32) * 2 - 1)
        B = fm.col_major(torch.rand((k, n), generator=g, device=DEV, dtype=torch.float32) * 2 - 1)
        probs.append({"A": A, "B": B, "C": fm.empty(m, n, torch.float32, DEV)})
    b = fm.Batch(probs, graph=True)
    b.run()
    b.run()
    torch.cuda.synchronize()
    for i, (p, (m, k, n)) in enumerate(zip(probs, shapes)):
        alg, st = b.choice(i)
        rows = torch.from_numpy(rng.integers(0, m, 32)).to(DEV)
        cols = torch.from_numpy(rng.integers(0, n, 32)).to(DEV)
        Ar = np.asfortranarray(p["A"][rows].double().cpu().numpy())
        Bc = np.asfortranarray(p["B"][:, cols].double().cpu().numpy())
        e = np.arange(32)
        q, absab = oracle_mod.quad_entries(Ar, Bc, e, e)
        got = p["C"][rows, cols].double().cpu().numpy()
        assert np.all(np.abs(got - q) <= 64 * (1 + st) * 2.0 ** -24 * absab), (i, alg, st)
