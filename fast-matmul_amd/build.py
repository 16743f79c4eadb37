#!/usr/bin/env python3
"""Build libfastmm.so for gfx950: run the generator, compile every HIP/C++
source with hipcc in parallel, link one shared library into fastmm/.

    python fast-matmul_amd/build.py [-j N]

Object files go to fast-matmul_amd/_build/ (git-ignored); the library lands
at fast-matmul_amd/fastmm/libfastmm.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
OUT = os.path.join(PKG, "fastmm", "libfastmm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wno-unused-command-line-argument",
]
# stage kernels must keep every multiply and add separately rounded
NOCONTRACT = ["-ffp-contract=off"]


def sources():
    srcs = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".cpp")):
            srcs.append((os.path.join(CSRC, f), []))
    gdir = os.path.join(CSRC, "gen")
    for f in sorted(os.listdir(gdir)):
        if f.endswith((".hip", ".cpp")):
            srcs.append((os.path.join(gdir, f), NOCONTRACT))
    return srcs


# host-only translation units (C ABI, plan builder, registry)
HOSTCXX = os.environ.get("FM_HOSTCXX", "/opt/rocm/llvm/bin/clang++")
HOST = ["-O2", "-fPIC", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wall", "-Wno-unused-function"]


def _up_to_date(obj, dep):
    """True when obj and its dependency list exist and no dependency is newer."""
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return False
    t = os.path.getmtime(obj)
    with open(dep) as f:
        text = f.read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else []
    return all(os.path.exists(d) and os.path.getmtime(d) <= t for d in files)


def compile_one(src, extra, incremental=False):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "__")
    obj = os.path.join(BUILD, rel + ".o")
    dep = obj + ".d"
    if incremental and _up_to_date(obj, dep):
        return obj, 0.0
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + extra + ["-MMD", "-MF", dep, "-x", "hip", "-c", src, "-o", obj]
    else:
        cmd = [HOSTCXX] + HOST + ["-MMD", "-MF", dep, "-c", src, "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return obj, time.time() - t0


def build(jobs: int = 0, verbose: bool = False, incremental: bool = False) -> str:
    """Generate and compile everything (incremental: skip objects whose
    sources and headers are unchanged since they were built)."""
    sys.path.insert(0, os.path.join(PKG, "gen"))
    import hipgen  # noqa: E402

    hipgen.main(["hipgen"])
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(16, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(compile_one, s, e, incremental) for s, e in srcs]
        for f in futs:
            obj, dt = f.result()
            objs.append(obj)
            if verbose:
                print("  %-60s %.1fs" % (os.path.basename(obj), dt))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=0)
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources whose dependencies changed")
    a = ap.parse_args()
    t0 = time.time()
    print(build(a.j, a.v, a.incremental), "%.1fs" % (time.time() - t0))
