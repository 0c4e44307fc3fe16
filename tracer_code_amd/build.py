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
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FDIR, "_build")
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


def build_fortran(verbose=False):
    """Compile the ISO_C_BINDING module and its two programs with flang
    (tc_layout: struct layout check, no GPU; tc_solve_demo: a Fortran caller
    of the solver).  They find libtracer_hip.so through an $ORIGIN rpath.
    Returns None when no Fortran compiler is installed."""
    if not os.path.exists(FLANG):
        return None
    os.makedirs(FBUILD, exist_ok=True)
    mod_o = os.path.join(FBUILD, "tracer_hip.o")
    link = ["-L", HERE, "-ltracer_hip", "-Wl,-rpath,$ORIGIN/../.."]
    cmds = [[FLANG, "-O2", "-c", "-J", FBUILD, os.path.join(FDIR, "tracer_hip.f90"), "-o", mod_o]]
    for prog in ("tc_layout", "tc_solve_demo"):
        cmds.append([FLANG, "-O2", "-I", FBUILD, os.path.join(FDIR, prog + ".f90"), mod_o] + link +
                    ["-o", os.path.join(FBUILD, prog)])
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed building the Fortran interface")
    return FBUILD


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
