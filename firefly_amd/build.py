"""Build the gfx950 HIP library (firefly_amd/libff_hip.so).

Run ``python -m firefly_amd.build`` or ``__graft_entry__.build()``.  Compiles
every source in firefly_amd/csrc with hipcc for gfx950 in-tree, so the built
library travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libff_hip.so")
HOST_LIB = os.path.join(HERE, "libfirefly.so")
DEMO = os.path.join(HERE, "render_demo")

SOURCES = ["ff_kernels.hip", "ff_lbvh.hip", "ff_abi.cpp", "ff_bvh.cpp"]

# -ffp-contract=off: every float/double operation rounds on its own, as the
# reference's arithmetic is written (bit-exact hits against the host oracle).
# -packed-fp32-ops off: no packed FP32 instructions are generated.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build(verbose: bool = False) -> str:
    """Compile libff_hip.so; raises on failure.  Returns the library path."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    cc = hipcc()
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    for src in SOURCES:
        obj = os.path.join(build_dir, src + ".o")
        cmd = [cc] + [f for f in HIPCC_FLAGS if f != "-shared"] + inc + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


def build_host(verbose: bool = False) -> None:
    """C++ host interface (libfirefly.so, links libff_hip.so) and the demo."""
    cc = hipcc()
    host = os.path.join(HERE, "host")
    common = ["-O2", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", host,
              "-Wl,-rpath,$ORIGIN", "-L", HERE]
    steps = [
        ([cc] + common + ["-shared", os.path.join(host, "firefly.cpp"), os.path.join(host, "gltf.cpp"),
                          os.path.join(host, "png.cpp"), "-lff_hip",
                          "-o", HOST_LIB + ".tmp"],
         HOST_LIB),
        ([cc] + common + [os.path.join(host, "render_demo.cpp"), "-lfirefly", "-lff_hip", "-o", DEMO + ".tmp"],
         DEMO),
    ]
    for cmd, out in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)


def build_oracle(verbose: bool = False) -> str:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)
    return os.path.join(odir, "liboracle.so")


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_oracle(verbose="-v" in sys.argv)
    print(LIB)
