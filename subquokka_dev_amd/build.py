"""Builds the native libraries in-tree for gfx950.

    libsubquokka_hip.so   HIP engine + C++ host formats behind include/subquokka_hip.h
    quokka                CLI with the reference's -i <ini> -c <circuit> interface

Compiled with hipcc directly (no CMake, no JIT cache): the outputs sit next to
the sources so they travel with the repository to the GPU host.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsubquokka_hip.so")
CLI = os.path.join(HERE, "quokka")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops off: hipcc is not allowed to form packed FP32 instructions.
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-value",
]
INCLUDES = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

# Per-source flags.  sq_pass.hip: the register-pass op loop dispatches on
# wave-uniform op records; left structurized, each uniform switch becomes flow
# blocks whose merges copy the whole register-resident state once or twice per
# op.  Skipping structurization of uniform regions keeps them as plain scalar
# branches (divergent regions are still structurized).
SOURCE_FLAGS = {"sq_pass.hip": ["-mllvm", "-structurizecfg-skip-uniform-regions"],
                # MFMA results in vector registers: no accumulator-register reads per batch
                "sq_fused.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
                # k_adj_multi branches on each gate's kind (uniform): plain scalar
                # branches, not merged regions holding every kind's temporaries
                "sq_adjoint.hip": ["-mllvm", "-structurizecfg-skip-uniform-regions"]}

HIP_SOURCES = ["sq_engine.hip", "sq_blocks.hip", "sq_fused.hip", "sq_adjoint.hip", "sq_stream.hip", "sq_pass.hip",
               "sq_sample.hip"]
HOST_SOURCES = ["host/formats.cpp", "host/simulator.cpp", "host/planner.cpp"]
CLI_SOURCES = ["host/quokka_main.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _objdir():
    d = os.path.join(HERE, "_build")
    os.makedirs(d, exist_ok=True)
    return d


def build(verbose=True, jobs=8):
    """Compiles every source and links the shared library and the CLI."""
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    objdir = _objdir()
    procs, objs = [], []
    for src in HIP_SOURCES + HOST_SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        flags = list(HIP_FLAGS) + SOURCE_FLAGS.get(src, [])
        if src.endswith(".cpp"):
            flags = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-x", "c++"]
            cmd = ["/usr/bin/g++"] + flags + INCLUDES + ["-c", os.path.join(CSRC, src), "-o", obj]
        else:
            cmd = [HIPCC] + flags + INCLUDES + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError("compile failed")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, LIB)
    if CLI_SOURCES:
        srcs = [os.path.join(CSRC, s) for s in CLI_SOURCES]
        _run(["/usr/bin/g++", "-std=c++17", "-O2", "-o", CLI] + INCLUDES + ["-I" + os.path.join(CSRC, "host")] +
             srcs + ["-L" + HERE, "-lsubquokka_hip", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
