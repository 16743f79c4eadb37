#!/usr/bin/env python3
"""Benchmark: effective TFLOP/s (2MNK/t) of the fast multiply on one MI355X.

Workload (default): grey424_26_206 = <4,2,4;26> with CSE, numsteps = 2,
FP64, square M = K = N = 8192, A and B iid uniform(-1,1) from a fixed seed,
alpha = 1, beta = 0 -- inputs resident in HBM before the timed region.
A step is one C := A B through the library's C ABI (S/T stages, one grouped
MFMA launch of the 676 leaf products, W stages).

Printed as one JSON line: the effective rate, the same-GPU classical rates
(the library's own MFMA GEMM and torch/hipBLASLt), the error against the
classical product, the leaf kernel's roofline fraction (durations from HIP
events on the launch stream over the timed region) and the reference's CPU
path timed on this host's cores.

    python bench.py [--steps K] [--warmup W] [--workload c2sq|c2ts|c1] [--no-cpu]
                    [--configs c1,c2ts,c3,c4,c5 | --no-configs]

Each secondary configuration reports its own roofline: the leaf kernel
against its matrix-unit peak (FP64, FP32, BF16 x 6 slice products, INT8 x 36
digit products) and every stage, including the plane preparation, against
HBM bandwidth, from the library's per-launch events.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fast-matmul_amd"))

PEAK_FP64_TF = 78.6
PEAK_HBM_GBS = 8000.0
# HBM bytes of the default workload's leaf launches per step, from the two
# rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) summarised in this file,
# scaled by a 1 GiB copy kernel measured in the same runs; valid for the leaf
# plan recorded beside it
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_grey424_26_206_fp64_8192.md")


def measured_traffic(leaf_launches):
    """Leaf bytes per step from TRAFFIC_FILE if it was measured on this plan."""
    plan = "+".join("%dx%d" % (l["cfg"], l["products"]) for l in leaf_launches)
    val, rec = None, None
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_step (leaf GEMM):"):
                    val = float(line.split(":")[1])
                if line.startswith("leaf plan (tile config x products):"):
                    rec = line.split(":")[1].strip()
    except OSError:
        return None
    return val if rec == plan else None

WORKLOADS = {
    # name: (alg, steps, M, K, N)
    "c2sq": ("grey424_26_206", 2, 8192, 8192, 8192),
    "c2ts": ("grey424_26_206", 2, 131072, 2400, 2400),
    "c2ts1": ("grey424_26_206", 1, 131072, 2400, 2400),
    "c1": ("strassen", 1, 2048, 2048, 2048),
}


def leaf_flops(R, m, k, n, steps, M, K, N):
    mm, kk, nn = M, K, N
    for _ in range(steps):
        mm, kk, nn = mm // m, kk // k, nn // n
    return R ** steps * 2.0 * mm * kk * nn, (mm, kk, nn)


def timed(fn, steps, sync):
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def _rate(fn, flops, sync, steps=3):
    fn()
    sync()
    t = timed(fn, steps, sync)
    return flops / t / 1e12, t * 1e3


def _err(C, D, A, B):
    """(normwise, elementwise) error of C against the reference product D."""
    d = (C.double() - D).abs()
    nw = float(d.max()) / (float(A.abs().max()) * float(B.abs().max()) * A.shape[1])
    return nw, float((d / D.abs()).max())


# published dense peaks of MI355X: TFLOP/s (TOP/s for int8), HBM GB/s
PEAKS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "int8": 5000.0, "hbm_GBps": PEAK_HBM_GBS}
# slice products per leaf product: split-bf16 (3 slices, i + j <= 2) and
# int8 digits (8 digits, s + t <= 7)
SPLIT_PRODUCTS = 6
OZAKI_PRODUCTS = 36


def profiled(fastmm, fn, steps, sync):
    """Per-step device time and launch count of each stage kind over `steps`
    calls of fn (events the library records around each launch)."""
    fn()
    sync()
    fastmm.profile_read()
    fastmm.profile_enable(True)
    try:
        for _ in range(steps):
            fn()
        sync()
    finally:
        fastmm.profile_enable(False)
    p = fastmm.profile_read()
    return {k: (v[0] / steps, v[1] / steps) for k, v in p.items()}


def roofline_block(prof, plan, leaf_ops, leaf_peak, leaf_unit, leaf_kind):
    """Leaf kernel against its matrix-unit peak, each stage against HBM:
    algorithmic work per step over the measured per-step time."""
    out = {}
    ms, n = prof["LEAF"]
    if ms > 0:
        a = leaf_ops / (ms * 1e-3) / 1e12
        out["leaf"] = {"kernel": leaf_kind, "ms": ms, "launches": n, "achieved": a, "peak": leaf_peak,
                       "unit": leaf_unit, "frac": a / leaf_peak, "ops": leaf_ops}
    for k in ("S", "T", "W"):
        ms, n = prof[k]
        if ms > 0 and plan["stage_bytes"][k] > 0:
            gbs = plan["stage_bytes"][k] / (ms * 1e-3) / 1e9
            out[k] = {"ms": ms, "launches": n, "algorithmic_GB": plan["stage_bytes"][k] / 1e9,
                      "GB_per_s": gbs, "hbm_frac": gbs / PEAK_HBM_GBS}
    ms, n = prof.get("PREP", (0.0, 0))
    if ms > 0:
        blk = {"ms": ms, "launches": n}
        if plan.get("prep_bytes"):
            gbs = plan["prep_bytes"] / (ms * 1e-3) / 1e9
            blk.update({"algorithmic_GB": plan["prep_bytes"] / 1e9, "GB_per_s": gbs,
                        "hbm_frac": gbs / PEAK_HBM_GBS})
        out["PREP"] = blk
    if prof["PEEL"][0] > 0:
        out["PEEL"] = {"ms": prof["PEEL"][0], "launches": prof["PEEL"][1]}
    return out


def stage_patterns(t, steps, M, K, N, stage_bytes):
    """Slab counts of the two-level S, T and W stages (both levels in one
    launch): (source slabs, output slabs, elements per slab).  S reads the
    (m k)^2 sub-blocks of A and writes the rest of its bytes as slabs of the
    same size; T likewise on B; W reads the R^2 products and writes C's
    (m n)^2 sub-blocks."""
    if steps != 2:
        return {}
    m, k, n = t.m, t.k, t.n
    if M % (m * m) or K % (k * k) or N % (n * n):
        return {}
    sl = {"S": (M // (m * m)) * (K // (k * k)), "T": (K // (k * k)) * (N // (n * n)),
          "W": (M // (m * m)) * (N // (n * n))}
    reads = {"S": (m * k) ** 2, "T": (k * n) ** 2, "W": t.R ** 2}
    out = {}
    for key in ("S", "T", "W"):
        total = stage_bytes[key] / (8.0 * sl[key])
        if abs(total - round(total)) < 1e-6 and round(total) > reads[key]:
            out[key] = (reads[key], int(round(total)) - reads[key], sl[key])
    return out


def measured_peaks(fastmm, torch, dev):
    """The matrix units' sustained rates (fm_probe_mfma_peak) and the HBM
    bandwidth of plain streaming kernels on this GPU."""
    out = {}
    for kind in ("fp64", "fp32", "bf16", "int8"):
        try:
            out[kind] = fastmm.probe_mfma_peak(kind, 0.05)
        except Exception as e:  # reported, never fatal
            out[kind] = "error: %s" % e
    x = torch.empty(1 << 29, dtype=torch.float64, device=dev).fill_(1.0)
    y = torch.empty_like(x)
    sync = torch.cuda.synchronize
    # one read and one write stream, as the stage kernels have
    out["hbm_elementwise_GBps"] = 2.0 * x.numel() * 8 / timed(lambda: torch.abs(x, out=y), 5, sync) / 1e9
    # a write stream alone (the stages write far more than they read)
    out["hbm_write_GBps"] = x.numel() * 8 / timed(lambda: y.fill_(2.0), 5, sync) / 1e9
    del x, y
    torch.cuda.empty_cache()
    out["note"] = ("tera-ops/s of back-to-back MFMAs on every SIMD (no memory traffic); GB/s of an "
                   "elementwise 4 GiB read + 4 GiB write and of a 4 GiB fill (torch kernels); "
                   "the rooflines use the published peaks")
    return out


def other_configs(fastmm, torch, dev, names=("c1", "c2ts", "c3", "c4", "c5")):
    """Secondary measurements of the other configurations (those in
    `names`), each guarded."""
    sync = torch.cuda.synchronize
    out = {}
    g = torch.Generator(device=dev)
    g.manual_seed(0xC0FFEE)

    def rnd(m, n, dt):
        return fastmm.col_major(torch.rand((m, n), generator=g, device=dev, dtype=dt) * 2 - 1)

    def guarded(name, fn):
        try:
            out[name] = fn()
        except Exception as e:  # report, never fail the headline number
            out[name] = {"error": "%s: %s" % (type(e).__name__, e)}
        torch.cuda.empty_cache()

    def fast_roofline(alg, steps, M, K, N, fn, mode="plain", dtype=torch.float64, psteps=3):
        t = fastmm.tables.get_table(alg)
        F, _ = leaf_flops(t.R, t.m, t.k, t.n, steps, M, K, N)
        plan = fastmm.plan_info(alg, M, K, N, steps, dtype=dtype, mode=mode)
        prof = profiled(fastmm, fn, psteps, sync)
        if mode == "split_bf16":
            return roofline_block(prof, plan, F * SPLIT_PRODUCTS, PEAKS["bf16"], "TFLOP/s",
                                  "split-bf16 slice products (6 per leaf, bf16 MFMA)")
        if mode == "ozaki_int8":
            return roofline_block(prof, plan, F * OZAKI_PRODUCTS, PEAKS["int8"], "TOP/s",
                                  "int8 digit products (36 per leaf, int8 MFMA)")
        key = "fp64" if dtype == torch.float64 else "fp32"
        return roofline_block(prof, plan, F, PEAKS[key], "TFLOP/s", "%s MFMA leaf GEMM" % key.upper())

    def c1():
        n = 2048
        A, B = rnd(n, n, torch.float64), rnd(n, n, torch.float64)
        C = fastmm.empty(n, n, torch.float64, dev)
        fl = 2.0 * n ** 3
        fn = lambda: fastmm.fastmatmul("strassen", A, B, C, 1)  # noqa: E731
        v, ms = _rate(fn, fl, sync, 10)
        D = torch.matmul(B.t(), A.t()).t()
        vb, _ = _rate(lambda: torch.matmul(B.t(), A.t()), fl, sync, 10)
        vo, _ = _rate(lambda: fastmm.gemm(A, B, C), fl, sync, 10)
        fn()
        nw, ew = _err(C, D, A, B)
        return {"workload": "strassen numsteps=1 FP64 2048^3", "tflops": v, "ms": ms,
                "torch_matmul_tflops": vb, "fastmm_gemm_tflops": vo, "normwise_err": nw,
                "max_rel_diff": ew,
                "roofline": fast_roofline("strassen", 1, n, n, n, fn, psteps=10)}

    def c2ts():
        M, K, N = 131072, 2400, 2400
        A, B = rnd(M, K, torch.float64), rnd(K, N, torch.float64)
        C = fastmm.empty(M, N, torch.float64, dev)
        fl = 2.0 * M * K * N
        r = {"workload": "grey424_26_206 FP64 131072x2400x2400"}
        D = torch.matmul(B.t(), A.t()).t()
        for st in (1, 2):
            fn = lambda: fastmm.fastmatmul("grey424_26_206", A, B, C, st)  # noqa: E731
            r["numsteps%d_tflops" % st], r["numsteps%d_ms" % st] = _rate(fn, fl, sync)
            r["numsteps%d_normwise_err" % st], r["numsteps%d_max_rel_diff" % st] = _err(C, D, A, B)
            r["numsteps%d_roofline" % st] = fast_roofline("grey424_26_206", st, M, K, N, fn)
        r["fastmm_gemm_tflops"], _ = _rate(lambda: fastmm.gemm(A, B, C), fl, sync)
        r["torch_matmul_tflops"], _ = _rate(lambda: torch.matmul(B.t(), A.t()), fl, sync)
        return r

    def c4():
        n = 16384
        r = {"workload": "strassen numsteps=2 FP32 data 16384^3, split-bf16 leaves (6 products)"}
        A, B = rnd(n, n, torch.float32), rnd(n, n, torch.float32)
        for dist in ("uniform", "scaled"):
            if dist == "scaled":
                rs = 10.0 ** (torch.rand((n, 1), generator=g, device=dev, dtype=torch.float32) * 6 - 3)
                cs = 10.0 ** (torch.rand((1, n), generator=g, device=dev, dtype=torch.float32) * 6 - 3)
                A = fastmm.col_major((A * rs).contiguous())
                B = fastmm.col_major((B * cs).contiguous())
            C = fastmm.empty(n, n, torch.float32, dev)
            fl = 2.0 * n ** 3
            fn = lambda: fastmm.fastmatmul("strassen", A, B, C, 2, mode="split_bf16")  # noqa: E731
            v, ms = _rate(fn, fl, sync)
            # FP64 product as the oracle, on a 2048-row slice
            rows = slice(0, 2048)
            D = torch.matmul(B.double().t(), A[rows].double().t()).t()
            nw, ew = _err(C[rows], D, A[rows].double(), B.double())
            r[dist] = {"tflops": v, "ms": ms, "normwise_err_vs_fp64": nw, "max_rel_diff_vs_fp64": ew}
            if dist == "uniform":
                r["roofline"] = fast_roofline("strassen", 2, n, n, n, fn, mode="split_bf16",
                                              dtype=torch.float32)
            fastmm.fastmatmul("strassen", A, B, C, 2)
            nwp, _ = _err(C[rows], D, A[rows].double(), B.double())
            r[dist]["plain_fp32_leaves_normwise_err"] = nwp
            del D
        fn = lambda: fastmm.fastmatmul("strassen", A, B, C, 2)  # noqa: E731
        r["plain_fp32_strassen2_tflops"], _ = _rate(fn, 2.0 * n ** 3, sync)
        r["plain_fp32_roofline"] = fast_roofline("strassen", 2, n, n, n, fn, dtype=torch.float32)
        r["torch_fp32_matmul_tflops"], _ = _rate(lambda: torch.matmul(B.t(), A.t()), 2.0 * n ** 3, sync)
        return r

    def c5():
        n = 13824
        A, B = rnd(n, n, torch.float64), rnd(n, n, torch.float64)
        C = fastmm.empty(n, n, torch.float64, dev)
        fl = 2.0 * n ** 3
        r = {"workload": "smirnov333_23_139 numsteps=2 FP64 13824^3, int8-digit leaves (8 digits)"}
        fn = lambda: fastmm.fastmatmul("smirnov333_23_139", A, B, C, 2, mode="ozaki_int8")  # noqa: E731
        r["ozaki_tflops"], r["ozaki_ms"] = _rate(fn, fl, sync)
        D = torch.matmul(B.t(), A.t()).t()
        r["ozaki_normwise_err"], r["ozaki_max_rel_diff"] = _err(C, D, A, B)
        r["ozaki_roofline"] = fast_roofline("smirnov333_23_139", 2, n, n, n, fn, mode="ozaki_int8")
        fn = lambda: fastmm.fastmatmul("smirnov333_23_139", A, B, C, 2)  # noqa: E731
        r["plain_fp64_tflops"], r["plain_fp64_ms"] = _rate(fn, fl, sync)
        r["plain_normwise_err"], _ = _err(C, D, A, B)
        r["plain_fp64_roofline"] = fast_roofline("smirnov333_23_139", 2, n, n, n, fn)
        r["torch_matmul_tflops"], _ = _rate(lambda: torch.matmul(B.t(), A.t()), fl, sync)
        return r

    def c3():
        import numpy as np
        rng = np.random.default_rng(0xC3)
        vals = np.arange(512, 4097, 64)
        shapes = [tuple(int(v) for v in rng.choice(vals, 3)) for _ in range(256)]
        probs = []
        fl = 0.0
        for (m, k, n) in shapes:
            probs.append({"A": rnd(m, k, torch.float32), "B": rnd(k, n, torch.float32),
                          "C": fastmm.empty(m, n, torch.float32, dev)})
            fl += 2.0 * m * k * n

        def batch_run(mode):
            b = fastmm.Batch(probs, mode=mode, graph=True)
            v, ms = _rate(b.run, fl, sync)
            chosen = {}
            leaf_fl = 0.0
            for i, (m, k, n) in enumerate(shapes):
                a, st = b.choice(i)
                chosen["%s/%d" % (a, st)] = chosen.get("%s/%d" % (a, st), 0) + 1
                if st == 0:
                    leaf_fl += 2.0 * m * k * n
                else:
                    t = fastmm.tables.get_table(a)
                    leaf_fl += leaf_flops(t.R, t.m, t.k, t.n, st, m, k, n)[0]
            errs = []
            for p in probs[:8]:
                D = torch.matmul(p["B"].double().t(), p["A"].double().t()).t()
                errs.append(_err(p["C"], D, p["A"].double(), p["B"].double())[0])
            del b
            # the same batch without the graph, with the library's per-launch events
            be = fastmm.Batch(probs, mode=mode, graph=False)
            prof = profiled(fastmm, be.run, 3, sync)
            del be
            plan = {"stage_bytes": {"S": 0.0, "T": 0.0, "W": 0.0}}
            if mode == "split_bf16":
                roof = roofline_block(prof, plan, leaf_fl * SPLIT_PRODUCTS, PEAKS["bf16"], "TFLOP/s",
                                      "split-bf16 slice products (6 per leaf, bf16 MFMA)")
            else:
                roof = roofline_block(prof, plan, leaf_fl, PEAKS["fp32"], "TFLOP/s",
                                      "FP32 MFMA grouped leaf GEMM")
            return v, ms, chosen, max(errs), roof

        # FP32-accurate leaves from split-bf16 slice products (6 bf16 MFMA
        # products per leaf, FP32 accumulation)
        vs, mss, chosen_s, err_s, roof_s = batch_run("split_bf16")
        v, ms, chosen, err_p, roof_p = batch_run("plain")
        # classical reference: one torch matmul per problem
        vt, _ = _rate(lambda: [torch.matmul(p["B"].t(), p["A"].t()) for p in probs], fl, sync)
        return {"workload": "256 FP32 problems, M,K,N uniform on {512,...,4096}/64, autotuned, hipGraph",
                "total_gflop": fl / 1e9, "tflops": v, "ms": ms, "choices": chosen,
                "torch_per_problem_matmul_tflops": vt, "normwise_err_max_first8": err_p,
                "roofline": roof_p,
                "split_bf16": {"tflops": vs, "ms": mss, "choices": chosen_s,
                               "normwise_err_max_first8": err_s, "roofline": roof_s}}

    for name, fn in (("c1", c1), ("c2ts", c2ts), ("c3", c3), ("c4", c4), ("c5", c5)):
        if name in names:
            guarded(name, fn)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2sq", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the host CPU baseline")
    ap.add_argument("--cpu-n", type=int, default=8192,
                    help="square size of the all-core CPU modes (the GPU workload's size)")
    ap.add_argument("--cpu-n1", type=int, default=4096, help="square size of the one-core CPU mode")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed CPU runs per mode (median)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the secondary measurements of the other workloads")
    ap.add_argument("--configs", default="c1,c2ts,c3,c4,c5",
                    help="comma-separated secondary workloads to measure (c1, c2ts, c3, c4, c5)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 1 GiB read + 1 GiB write elementwise kernel, the "
                         "known-size reference for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the fast multiply runs on a single MI355X "
              "(multi-GPU sharding is designed, not built); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import fastmm

    alg, steps, M, K, N = WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(0xFA57)
    A = fastmm.col_major(torch.rand((M, K), generator=g, device=dev, dtype=torch.float64) * 2 - 1)
    B = fastmm.col_major(torch.rand((K, N), generator=g, device=dev, dtype=torch.float64) * 2 - 1)
    C = fastmm.empty(M, N, torch.float64, dev)
    sync = torch.cuda.synchronize

    def step():
        fastmm.fastmatmul(alg, A, B, C, steps)

    for _ in range(args.warmup):
        step()
    sync()
    fastmm.profile_read()
    fastmm.profile_enable(True)
    t = timed(step, args.steps, sync)
    fastmm.profile_enable(False)
    prof = fastmm.profile_read()
    flops = 2.0 * M * N * K
    value = flops / t / 1e12

    # same-GPU classical products
    Cg = fastmm.empty(M, N, torch.float64, dev)
    csteps = max(2, args.steps // 2)
    fastmm.gemm(A, B, Cg)
    t_own = timed(lambda: fastmm.gemm(A, B, Cg), csteps, sync)
    At, Bt = A.t(), B.t()  # row-major views for torch: C^T = B^T A^T
    D = torch.matmul(Bt, At)
    t_blas = timed(lambda: torch.matmul(Bt, At, out=D), csteps, sync)
    Dc = D.t()
    err_abs = (C - Dc).abs()
    normwise = float(err_abs.max()) / (float(A.abs().max()) * float(B.abs().max()) * K)
    elementwise = float((err_abs / Dc.abs()).max())
    del err_abs

    t_tab = fastmm.tables.get_table(alg)
    F_leaf, leaf_dims = leaf_flops(t_tab.R, t_tab.m, t_tab.k, t_tab.n, steps, M, K, N)
    # the leaf products of one step may run as two launches (full rounds of
    # large tiles, then the part-empty last round on small tiles): the leaf
    # kernel's rate is the step's leaf flops over the step's leaf kernel time
    leaf_ms, leaf_n = prof["LEAF"]
    leaf_ms_step = leaf_ms / args.steps
    achieved = F_leaf / (leaf_ms_step * 1e-3) / 1e12 if leaf_n else None
    plan = fastmm.plan_info(alg, M, K, N, steps)
    # stage rates against HBM: algorithmic bytes of each stage (every source
    # element read once, every output written once) over its measured time
    plan_bytes = fastmm.plan_info(alg, M, K, N, steps)["stage_bytes"]
    stages = {}
    for k, v in prof.items():
        ms = v[0] / args.steps
        st = {"ms_per_step": ms, "launches_per_step": v[1] / args.steps}
        if k in ("S", "T", "W") and ms > 0:
            gbs = plan_bytes[k] / (ms * 1e-3) / 1e9
            st.update({"algorithmic_GB": plan_bytes[k] / 1e9, "GB_per_s": gbs,
                       "hbm_frac": gbs / PEAK_HBM_GBS})
        stages[k] = st
    # each stage against the measured rate of its own access pattern
    # (fm_probe_stream: one element position of every source slab read, of
    # every output slab written, per lane)
    for k, pat in stage_patterns(t_tab, steps, M, K, N, plan_bytes).items():
        if stages.get(k, {}).get("GB_per_s"):
            try:
                pgbs = fastmm.probe_stream(*pat)
                stages[k].update({"pattern": {"reads": pat[0], "writes": pat[1], "slab_elems": pat[2]},
                                  "pattern_GBps": pgbs, "pattern_frac": stages[k]["GB_per_s"] / pgbs})
            except Exception as e:  # reported, never fatal
                stages[k]["pattern_GBps"] = "error: %s" % e

    if args.copy_probe:
        x = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)
        y = torch.empty_like(x)
        sync()
        torch.mul(x, 1.0, out=y)
        sync()
        del x, y

    peaks = measured_peaks(fastmm, torch, dev)

    configs = None
    if not args.no_configs and args.workload == "c2sq":
        del Cg, D, At, Bt
        torch.cuda.empty_cache()
        configs = other_configs(fastmm, torch, dev, tuple(args.configs.split(",")))

    cpu = None
    if not args.no_cpu:
        try:
            r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"),
                                "--alg", alg, "--steps", str(steps), "--n", str(args.cpu_n),
                                "--n1", str(args.cpu_n1), "--reps", str(args.cpu_reps)],
                               stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                               timeout=600)
            if r.returncode == 0:
                cpu = json.loads(r.stdout.strip().splitlines()[-1])
            else:
                cpu = {"error": r.stderr.strip()[-400:]}
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": str(e)}

    out = {
        "metric": "effective TFLOP/s (2MNK/t) vs classical GEMM on MI355X at stated max rel err",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A, B iid uniform(-1,1), torch Philox seed 0xFA57, resident in HBM",
        "config": {"workload": "%s numsteps=%d FP64 %dx%dx%d (C := A B, alpha=1, beta=0)"
                               % (alg, steps, M, K, N),
                   "algorithm": alg, "numsteps": steps, "M": M, "K": K, "N": N,
                   "leaf_products": t_tab.R ** steps, "leaf_shape": list(leaf_dims)},
        "classical": {"fastmm_mfma_gemm_tflops": flops / t_own / 1e12,
                      "torch_matmul_tflops": flops / t_blas / 1e12,
                      "speedup_vs_best_classical": value / max(flops / t_own, flops / t_blas) * 1e12},
        "error_vs_classical": {"max_rel_diff_elementwise": elementwise,
                               "normwise_max_abs_over_maxA_maxB_K": normwise},
        "stages": stages,
        "roofline": {"bound": "mfma", "kernel": "leaf grouped GEMM (fm_gemm_kernel<double>)",
                     "achieved": achieved, "peak": PEAK_FP64_TF, "unit": "TFLOP/s",
                     "frac": (achieved / PEAK_FP64_TF) if achieved else None,
                     "flops_per_step": F_leaf, "leaf_ms_per_step": leaf_ms_step,
                     "leaf_launches_per_step": leaf_n / args.steps,
                     "leaf_launches": plan["leaf"],
                     "traffic": measured_traffic(plan["leaf"]) if args.workload == "c2sq" else None,
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT)
                     if args.workload == "c2sq" else None},
        "cpu_baseline": cpu,
        "measured_peaks": peaks,
        "configs": configs,
    }
    # whole multiply against its roofline: leaf flops at the FP64 MFMA peak
    # plus the stages' algorithmic bytes at the HBM peak (stages run one
    # after another, so the bounds add)
    b_add = sum(stages[k].get("algorithmic_GB", 0.0) for k in ("S", "T", "W")) * 1e9
    t_roof = F_leaf / (PEAK_FP64_TF * 1e12) + b_add / (PEAK_HBM_GBS * 1e9)
    out["multiply_roofline"] = {"t_roof_ms": t_roof * 1e3, "t_measured_ms": t * 1e3,
                                "frac": t_roof / t, "leaf_flops": F_leaf, "stage_bytes": b_add}
    # the same bound with each stage at the measured rate of its own access
    # pattern (fm_probe_stream) instead of the HBM nameplate
    pat = [stages[k] for k in ("S", "T", "W") if stages[k].get("algorithmic_GB")]
    if pat and all(isinstance(s.get("pattern_GBps"), float) for s in pat):
        t_pat = F_leaf / (PEAK_FP64_TF * 1e12) + sum(
            s["algorithmic_GB"] / s["pattern_GBps"] for s in pat)
        out["multiply_roofline"].update({"t_roof_pattern_ms": t_pat * 1e3,
                                         "frac_pattern": t_pat / t})
    print(json.dumps(out))


if __name__ == "__main__":
    main()
