"""Build the HIP library libtracer_hip.so in-tree for gfx950.

The library is a plain C-ABI shared object (include/tracer_hip.h) compiled
with hipcc; no PyTorch extension machinery is involved.  Device code is
compiled with -ffp-contract=off so that the FP64 path evaluates every
operation in the order the reference Fortran writes it, and with packed FP32
instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtracer_hip.so")
SOURCES = ["tc_api.hip", "tc_mg.hip", "tc_coarse.hip", "tc_fields.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    extra = [os.path.join(CSRC, s) for s in sorted(os.listdir(CSRC))
             if s.endswith(".hip") and s not in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs + extra
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtracer_hip.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
