"""Build libloopgpu.so (gfx950 kernels + C-ABI + worker) and the loopgpu_run
driver in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libloopgpu.so")
EXE = os.path.join(HERE, "loopgpu_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
]
SOURCES = ["lg_kernels.hip", "lg_api.hip", "lg_noalloc.hip", "worker.cpp", "correlation.cpp", "lg_server.cpp"]


def run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    run([HIPCC] + FLAGS + ["-shared"] + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", OUT + ".tmp"], verbose)
    os.replace(OUT + ".tmp", OUT)
    run(["g++", "-O2", "-std=c++17", os.path.join(CSRC, "loop_main.cpp"), "-o", EXE + ".tmp",
         "-L" + HERE, "-lloopgpu", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(EXE + ".tmp", EXE)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
