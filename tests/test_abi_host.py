"""The C-ABI library loads without a GPU and exports every entry point the
header declares; the ctypes struct layout equals the C layout; the host-side
generators are deterministic and follow the reference's initialisation rules.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from tracer_code_amd import _abi, config, ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tracer_hip.h")


def declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tc_\w+)\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.SIGNATURES)


def test_params_struct_layout_matches_c(tmp_path):
    prog = tmp_path / "sz.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "tracer_hip.h"\n'
                    'int main(){printf("%zu %zu %zu %zu\\n", sizeof(tc_params), offsetof(tc_params, Trrss_base),'
                    ' sizeof(tc_stats), offsetof(tc_stats, mg_cycles));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == ctypes.sizeof(_abi.TcParams)
    assert int(out[1]) == _abi.TcParams.Trrss_base.offset
    assert int(out[2]) == ctypes.sizeof(_abi.TcStats)
    assert int(out[3]) == _abi.TcStats.mg_cycles.offset


def test_record_struct_layouts_match_c(tmp_path):
    prog = tmp_path / "rec.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "tracer_hip.h"\n'
                    'int main(){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(tc_equalize_info),'
                    ' offsetof(tc_equalize_info, deficit), offsetof(tc_equalize_info, count),'
                    ' sizeof(tc_stats_row), offsetof(tc_stats_row, empty_cells), offsetof(tc_stats_row, ent));'
                    'return 0;}\n')
    exe = tmp_path / "rec"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    E, R = _abi.TcEqualizeInfo, _abi.TcStatsRow
    assert out == [ctypes.sizeof(E), E.deficit.offset, E.count.offset, ctypes.sizeof(R), R.empty_cells.offset,
                   R.ent.offset]


def test_init_rejects_unsupported_parameters():
    lib = _abi.load_library()
    p = config.make_params(32, 32)
    p.order = 4
    h = ctypes.c_void_p()
    assert lib.tc_init(ctypes.byref(p), 0, ctypes.byref(h)) == _abi.TC_ERR_UNSUPPORTED
    p = config.make_params(32, 32)
    p.ivis = 3
    assert lib.tc_init(ctypes.byref(p), 0, ctypes.byref(h)) == _abi.TC_ERR_UNSUPPORTED


@pytest.mark.parametrize("name,nx,nz,grids,ngc", [("cfg0", 128, 128, 6, 25), ("cfg1", 768, 256, 6, 225),
                                                  ("cfg2", 2048, 1024, 6, 2145)])
def test_presets(name, nx, nz, grids, ngc):
    p, cfg = config.preset_params(name)
    assert (p.nx, p.nz, p.grids) == (nx, nz, grids)
    assert (nx // 32 + 1) * (nz // 32 + 1) == ngc
    assert p.tolerance == 1e-8 and p.Nmulti == 70 and p.torder == 4 and p.RK_update == 1


def test_tracer_generator_deterministic_and_in_domain():
    r1, s1, t1 = ic.tracers(16, 8, 2.0, 5, 0.2, seed=9)
    r2, s2, t2 = ic.tracers(16, 8, 2.0, 5, 0.2, seed=9)
    assert np.array_equal(r1, r2) and np.array_equal(s1, s2) and np.array_equal(t1, t2)
    assert len(r1) == 17 * 9 * 5
    assert np.all((r1 > 0) & (r1 < 2.0) & (s1 > 0) & (s1 < 1.0))
    # creation order: tracer id -> node (i major, k minor); each within half a cell
    node = np.arange(len(r1)) // 5
    i, k = node // 9, node % 9
    assert np.all(np.abs(r1 - i * 2.0 / 16) <= 0.5 * 2.0 / 16)
    assert np.all(np.abs(s1 - k / 8) <= 0.5 / 8)
    assert np.array_equal(t1, (s1 <= 0.2).astype(np.int32))


def test_temperature_generator_boundary_conditions():
    T = ic.initial_temperature(16, 8, 2.0, "fourier4", seed=3)
    assert np.all(T[2, :] == 1.0) and np.all(T[2 + 8, :] == 0.0)
    assert np.array_equal(T[2:11, 1], T[2:11, 3])
    assert np.allclose(T[1, :], 2.0 - T[3, :], rtol=0, atol=1e-15)
    assert np.array_equal(T[11, :], -T[9, :])
