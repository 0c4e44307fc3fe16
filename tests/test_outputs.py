"""The output writers against the reference's own sample outputs.

tests/golden/{stats,viscosity_smoother,equalize}.dat are the files the
reference ships (one 18-step run of its main.f90 with comp=0, tracer=1,
equalize=1).  Every record is parsed into its values and written again;
the result must equal the file byte for byte, which pins the G and I edit
descriptors (F editing inside [0.1, 10**d), E editing outside, zero,
trailing blanks) and the column order of each writer.  The E-with-exponent
form of the snapshots (g14.5E3) and the edges of the F/E switch are
checked against the rules of ISO Fortran 2008 10.7.5.2.2.
"""
import os

import pytest

from tracer_code_amd import _abi
from tracer_code_amd.outputs import (e_format, equalize_line, g_format, i_format, snapshot_name, stats_line,
                                     viscosity_line)

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def records(name):
    with open(os.path.join(GOLD, name)) as f:
        return [l.rstrip("\n") for l in f]


def cut(line, widths):
    out, p = [], 0
    for w in widths:
        out.append(line[p:p + w])
        p += w
    assert p == len(line)
    return out


def test_stats_dat_records():
    lines = records("stats.dat")
    assert len(lines) == 18
    for line in lines:
        f = cut(line, [20] * 12 + [6] * 3)
        v = [float(x) for x in f[:12]]
        rec = _abi.TcStatsRow()
        (rec.time, rec.Tavg, rec.vrms, rec.Nu, rec.Nu_bot, rec.RaT_avg, rec.mobility, rec.strain_max) = v[:8]
        for j in range(4):
            rec.q[j] = v[8 + j]
        rec.empty_cells, rec.tracer_min, rec.tracer_max = (int(x) for x in f[12:])
        assert stats_line(rec, comp=0, ntype=0, tracer=1) == line


def test_viscosity_smoother_records():
    for line in records("viscosity_smoother.dat"):
        t, n_eta, eig, n_temp = cut(line, [20, 5, 20, 5])
        assert viscosity_line(float(t), int(n_eta), float(eig), int(n_temp), [], comp=0) == line


def test_equalize_dat_records():
    for line in records("equalize.dat"):
        f = cut(line, [20] + [10] * 6)
        info = _abi.TcEqualizeInfo()
        tpc = int(f[1])
        info.minexcess[0] = int(f[2])
        info.deficit, info.moved, info.nfree = int(f[3]), int(f[4]), int(f[5])
        info.count[0] = int(f[6])
        assert equalize_line(float(f[0]), info, comp=0, ntype=0, tpc=tpc) == line


@pytest.mark.parametrize("x,w,d,e,want", [
    (0.0, 20, 8, None, "       0.0000000    "),
    (1.0, 14, 5, 3, "   1.0000     "),
    (-2.5, 14, 5, 3, "  -2.5000     "),
    (0.1, 20, 8, None, "      0.10000000    "),
    (99999.123456, 14, 5, 3, "   99999.     "),
    (1e-3, 14, 5, 3, "  0.10000E-002"),
    (-1.5e7, 14, 5, 3, " -0.15000E+008"),
    (123456789.0, 20, 8, None, "      0.12345679E+09"),
    (99999999.5, 20, 8, None, "      0.10000000E+09"),
    (0.099999999999, 20, 8, None, "      0.10000000    "),
    (9.9999999999, 20, 8, None, "       10.000000    "),
    (1e-120, 20, 8, None, "      0.10000000-119"),
])
def test_g_editing_rules(x, w, d, e, want):
    assert g_format(x, w, d, e) == want


def test_i_editing_and_names():
    assert i_format(17, 6) == "    17"
    assert i_format(-3, 5) == "   -3"
    assert i_format(1234567, 6) == "******"
    assert snapshot_name(0) == "T10000" and snapshot_name(37) == "T10037"
    assert e_format(0.5, 8, 2) == "0.50E+00"
