"""Builds the HIP library libgmk.so for gfx950 (in-tree, under gmkbe_amd/_lib/)."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_lib")
LIB = os.path.join(OUT_DIR, "libgmk.so")
SOURCES = ["bb_kernel.hip", "bb_run.hip", "partition_kernels.hip", "gmk_engine.hip", "partition.hip", "thin_kernels.hip", "thin_host.hip", "seeds.cpp", "mlp.hip", "schedule.hip", "reduce.hip", "seed_kernels.hip", "edge_parse.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compiles every source (in parallel, `jobs` at a time; 0: one per source
    up to 8) and links the library and the CLI."""
    os.makedirs(OUT_DIR, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        if src.endswith(".cpp"):
            cmd = ["g++", "-O3", "-std=c++20", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj]
        else:
            cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=jobs or min(8, len(cmds))) as ex:
        for f in [ex.submit(subprocess.run, cmd, check=True) for cmd in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp, "-lpthread"],
                   check=True)
    os.replace(tmp, LIB)
    # command-line driver over the C ABI
    cli = os.path.join(OUT_DIR, "gmkbpe")
    subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", os.path.join(CSRC, "gmkbpe.cpp"), "-o", cli + ".tmp",
                    "-L" + OUT_DIR, "-lgmk", "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(cli + ".tmp", cli)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
