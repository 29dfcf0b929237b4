"""Builds libjc_hip.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python jump-cannon_amd/build.py

Every source compiles with -ffp-contract=off (no a*b+c -> fma: the f32 results
must match the restated reference path bit for bit) and with packed-FP32
instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libjc_hip.so")
SOURCES = ["prims.hip", "fa2bh.hip", "batch.hip", "pagerank.hip", "spmv.hip", "sgd.hip", "traverse.hip", "mpnn.hip", "host.cpp", "multilevel.cpp"]
# the reference's one-shot analytics binaries (crates/graph-compute/src/bin/)
CLIS = {"graph-pagerank": "graph_pagerank.cpp", "graph-bfs": "graph_bfs.cpp", "graph-cc": "graph_cc.cpp"}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("JC_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l
                     and "not a recognized feature" not in l)
    if warn.strip():
        sys.stderr.write(warn + "\n")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_cli(verbose)
    return OUT


def build_cli(verbose: bool = False) -> list:
    """graph-pagerank, graph-bfs and graph-cc: the reference's analytics CLIs
    over the C ABI (host code only, linked against libjc_hip.so next to them)."""
    outs = []
    for name, src in CLIS.items():
        out = os.path.join(HERE, name)
        tmp = out + ".tmp"
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, src), "-o", tmp, f"-L{HERE}",
               "-ljc_hip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CLI build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, out)
        if verbose:
            print(f"built {out}")
        outs.append(out)
    return outs


if __name__ == "__main__":
    build(verbose=True)
