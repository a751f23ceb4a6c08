"""Build libclotho_amd.so for gfx950 with hipcc (in-tree, no JIT cache)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib", "libclotho_amd.so")

SOURCES = ["capi.hip", "pipeline.hip", "batch.hip", "traits.hip", "traits_mfma.hip",
           "traits_noalloc.hip", "stream_copy.hip"]

HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # bit-exact parity with the host restatement: no FMA contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    objdir = os.path.join(HERE, "lib", "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = ["hipcc", *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = OUT + ".tmp"
    subprocess.check_call(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, OUT)
    build_host_demo(verbose)
    return OUT


DEMO = os.path.join(HERE, "lib", "engine_demo")
GENERIC_DEMO = os.path.join(HERE, "lib", "generic_demo")


def build_host_demo(verbose: bool = False) -> str:
    """The C++ adapter drivers: plain g++ against the C ABI library."""
    root = os.path.dirname(HERE)
    for src, out in (("engine_demo.cpp", DEMO), ("generic_demo.cpp", GENERIC_DEMO)):
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(root, "include"),
               os.path.join(HERE, "host", src), "-o", out + ".tmp",
               "-L", os.path.dirname(OUT), "-lclotho_amd", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
    return DEMO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
