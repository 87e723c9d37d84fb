"""Builds the HIP library and the command-line tool in-tree for gfx950.

    python sadcs_amd/build.py          # (python -m sadcs_amd.build imports the package, which needs a
                                       #  library matching its symbols)

Outputs sadcs_amd/libsadcs_hip.so (the C ABI of include/sadcs_hip.h) and sadcs_amd/bin/sadcs
(the command line). Every HIP compile disables packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsadcs_hip.so")
WORK = os.path.join(HERE, "libsadcs_work.so")
BIN = os.path.join(HERE, "bin", "sadcs")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SADCS_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    tmp = LIB + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(CSRC, "engine.hip"), "-o", tmp])
    os.replace(tmp, LIB)


def build_workloads():
    tmp = WORK + ".tmp"
    run(["g++", "-std=c++20", "-O2", "-fopenmp", "-fPIC", "-shared", "-Wall", os.path.join(CSRC, "workloads.cpp"), "-o", tmp])
    os.replace(tmp, WORK)


def build_cli():
    sources = [os.path.join(CSRC, f) for f in ("gpusa.cpp", "checker.cpp", "fileio.cpp")]
    sources += [os.path.join(CSRC, "cli", f) for f in ("main.cpp", "selftest.cpp")]
    sources = [s for s in sources if os.path.exists(s)]
    if len(sources) < 5:
        return
    tmp = BIN + ".tmp"
    run(["g++", "-std=c++20", "-O2", "-fopenmp", "-Wall", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
         *sources, "-o", tmp, f"-L{HERE}", "-lsadcs_hip", "-Wl,-rpath,$ORIGIN/..",
         "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, BIN)


def main():
    build_library()
    build_workloads()
    build_cli()


if __name__ == "__main__":
    sys.exit(main())
