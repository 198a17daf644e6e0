"""Build libbruno_hip.so in-tree for gfx950 with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", n) for n in ("bruno_hip.hip", "server.hip", "fastq.hip")]
OUT = os.path.join(HERE, "libbruno_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed fp32 instructions anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + SRCS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libbruno_hip.so")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_driver(verbose)
    return OUT


DRIVER_SRC = os.path.join(HERE, "host", "compact_dbg.cpp")
DRIVER = os.path.join(HERE, "host", "compact_dbg")


def build_driver(verbose=False):
    """The C++ driver over the C ABI (links libbruno_hip.so, rpath $ORIGIN/..)."""
    cmd = ["g++", "-O2", "-std=c++17", "-o", DRIVER, DRIVER_SRC, "-L" + HERE, "-lbruno_hip",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the compact_dbg driver")
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
