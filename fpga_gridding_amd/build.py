"""Build libfpga_gridding_amd.so in-tree with hipcc for gfx950.

Usage: ``python -m fpga_gridding_amd.build`` (or ``build()`` from Python).
The library is written next to this file so that it travels with the
source tree; it is not installed anywhere.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libfpga_gridding_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GW_OFFLOAD_ARCH", "gfx950")

SOURCES = ["gridder.hip", "weights.hip", "wkernel.hip", "capi.cpp",
           "server.cpp", "io.cpp"]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-pthread",
    "-fvisibility=hidden",
    # Packed FP32 instructions stay off (see the device code's notes).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # IEEE sqrt and division, denormals kept: the index math must round
    # exactly like the host's (oskar_grid_wproj.c:169-196).
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Wall",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    """Compile every source into the shared library; returns its path."""
    objs, cmds = [], []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *FLAGS, "-I", INCLUDE, "-I", CSRC, "-c",
               os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    # The sources compile independently: run them side by side.
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "4"))))
    with ThreadPoolExecutor(jobs) as ex:
        for r in ex.map(lambda c: subprocess.run(c), cmds):
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, r.args)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread",
           "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # The example program (host C, links the library in place).
    exe = os.path.join(HERE, "examples", "grid_vis")
    cmd = ["gcc", "-O2", "-std=gnu99", "-Wall", "-I", INCLUDE,
           os.path.join(HERE, "examples", "grid_vis.c"), "-o", exe,
           "-L", HERE, "-lfpga_gridding_amd", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
