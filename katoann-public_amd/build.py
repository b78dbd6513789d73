"""Builds libkatoann.so (gfx950 kernels + host library + C ABI) in place.

    python katoann-public_amd/build.py [-j N]

Device code: hipcc for gfx950, no fast-math, FP contraction off (every fused
multiply-add is explicit), fp32 denormals kept, packed-fp32 instructions off.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libkatoann.so")
CLI = os.path.join(HERE, "bin", "katoann")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-I" + os.path.join(HERE, "..", "include")]
DEVICE = ["--offload-arch=gfx950", "-fno-gpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
SOURCES = ["kernels.hip", "rank.hip", "rank_gemm.hip", "mlp.hip", "index.cpp", "server.cpp", "noalloc.hip", "build.hip"]


def _obj(src: str) -> str:
    return os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")


def _compile(src: str) -> str:
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    obj = _obj(src)
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return obj


def build(jobs: int = 4) -> str:
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s) for s in SOURCES]
        objs = [f.result() for f in futs]
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, OUT)
    _build_cli()
    return OUT


def _build_cli() -> str:
    """bin/katoann, the command-line front end, linked against libkatoann.so
    found next to bin/ at run time."""
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "cli", "katoann.cpp"), "-o", tmp, "-L" + HERE, "-lkatoann",
           "-Wl,-rpath,$ORIGIN/..", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    j = int(sys.argv[sys.argv.index("-j") + 1]) if "-j" in sys.argv else 4
    print(build(j))
