"""The Fortran ISO_C_BINDING interface (tracer_code_amd/fortran/tracer_hip.f90).

* CPU: flang's layout of the interoperable types equals the C layout
  (sizes and field offsets of tc_params / tc_stats), and the module has an
  interface for every entry point of include/tracer_hip.h.
* GPU: a Fortran program that calls the C-ABI in the order of the
  reference's start-up and first solve gives bit-identical SF, u, w and one
  time step's SF and tracers to the Python mirror of the same calls.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from tracer_code_amd import Solver, _abi, build, config, ic

FB = build.FBUILD


def test_module_binds_every_entry_point():
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "tracer_hip.h")).read()
    mod = open(os.path.join(root, "tracer_code_amd", "fortran", "tracer_hip.f90")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(tc_\w+)\(", hdr, re.M))
    bound = set(re.findall(r'bind\(C, name="(tc_\w+)"\)', mod))
    assert declared and not (declared - bound), sorted(declared - bound)
    assert not (bound - declared), sorted(bound - declared)


def need(prog):
    path = os.path.join(FB, prog)
    if not os.path.exists(path):
        pytest.skip(f"{prog} not built (no Fortran compiler)")
    return path


def test_fortran_type_layout_matches_c():
    out = subprocess.run([need("tc_layout")], capture_output=True, text=True, check=True).stdout
    got = {l.split()[0]: int(l.split()[1]) for l in out.splitlines() if l.strip()}
    P, S = _abi.TcParams, _abi.TcStats
    want = {"sizeof_params": ctypes.sizeof(P), "off_multi": P.multi.offset, "off_fixes": P.fixes.offset,
            "off_aspect": P.aspect.offset, "off_D1": P.D1.offset, "off_Dint": P.Dint.offset,
            "off_Trrss_base": P.Trrss_base.offset, "sizeof_stats": ctypes.sizeof(S),
            "off_residual_last": S.residual_last.offset, "off_mg_cycles": S.mg_cycles.offset,
            "off_q_step": S.q_step.offset}
    E, R = _abi.TcEqualizeInfo, _abi.TcStatsRow
    want.update({"sizeof_equalize": ctypes.sizeof(E), "off_deficit": E.deficit.offset, "off_count": E.count.offset,
                 "sizeof_stats_row": ctypes.sizeof(R), "off_empty_cells": R.empty_cells.offset,
                 "off_ent": R.ent.offset})
    assert got == want


@pytest.mark.gpu
def test_fortran_caller_matches_python_mirror(tmp_path):
    demo = need("tc_solve_demo")
    params, cfg = config.preset_params("cfg0", nx=32, nz=32)
    T, r, s, ty = ic.workload(cfg, seed=5)
    nx, nz, ntr = params.nx, params.nz, len(r)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(bytes(params))
        f.write(np.int64(ntr).tobytes())
        f.write(np.ascontiguousarray(T, dtype=np.float64).tobytes())
        f.write(r.astype(np.float64).tobytes())
        f.write(s.astype(np.float64).tobytes())
        f.write(ty.astype(np.int32).tobytes())
    print(f"[fortran demo 32x32 ntr={ntr}]", flush=True)
    res = subprocess.run([demo, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    raw = open(fout, "rb").read()
    plane = (nz + 5) * (nx + 5)
    q = int(np.frombuffer(raw[:4], np.int32)[0])
    d = np.frombuffer(raw[4:], np.float64)
    resid, resref = d[0], d[1]
    SF, u, w = (d[2 + j * plane: 2 + (j + 1) * plane].reshape(nz + 5, nx + 5) for j in range(3))
    o = 2 + 3 * plane
    SF1 = d[o:o + plane].reshape(nz + 5, nx + 5)
    r1, s1 = d[o + plane:o + plane + ntr], d[o + plane + ntr:o + plane + 2 * ntr]
    with Solver(params, ntr) as g:
        g.initialize(T, r, s, ty)
        qg, rg, rrg = g.multigrid()
        g.compute_velocities()
        assert (q, resid, resref) == (qg, rg, rrg)
        assert np.array_equal(SF, g.get("SF"))
        assert np.array_equal(u, g.get("u")) and np.array_equal(w, g.get("w"))
        g.time_step(1)
        assert np.array_equal(SF1, g.get("SF"))
        gr, gs = g.get_tracers("x")
        assert np.array_equal(r1, gr) and np.array_equal(s1, gs)
