"""Build the HIP tracer library (gfx950) in-tree: libmolflow_gpu.so."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmolflow_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the ray/facet test must round exactly like the CPU
# restatement (explicit fma calls are kept). -packed-fp32-ops: no packed
# FP32 instructions in any generated code.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
SOURCES = ["mfg_trace.hip", "mfg_host.cpp"]


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.rsplit(".", 1)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = OUT + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"], check=True)
    os.replace(tmp, OUT)
    for o in objs:
        os.remove(o)
    return OUT


MICROBENCH = os.path.join(HERE, "libmfg_microbench.so")


def build_microbench(verbose: bool = False) -> str:
    """The atomic-throughput measurement helper (include/mfg_microbench.h)."""
    src = os.path.join(CSRC, "mfg_microbench.hip")
    tmp = MICROBENCH + ".tmp"
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-shared", src, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, MICROBENCH)
    return MICROBENCH


ROOT = os.path.dirname(HERE)
DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "run_model.cpp")
DRIVER = os.path.join(ROOT, "tests", "cpp", "run_model")


def build_driver(verbose: bool = False) -> str:
    """The C++ host driver over include/molflow_gpu.hpp (used by the GPU
    tests); links the in-tree library through an $ORIGIN-relative rpath."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), DRIVER_SRC,
           "-L", HERE, "-lmolflow_gpu", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
           "-Wl,-rpath,$ORIGIN/../../molflow_sources_amd", "-o", DRIVER]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_driver(verbose=True))
    print(build_microbench(verbose=True))
