"""Text outputs of the reference, byte for byte: stats.dat (basics.f90:236-324),
viscosity_smoother.dat (basics.f90:316-322), equalize.dat
(composition.f90:566-574) and the T1xxxx snapshots (main.f90:234-268).

The reference writes them with Fortran edit descriptors; a Fortran host
calling the C-ABI keeps its own write statements.  This module is for a
Python host: Gw.d / Gw.dEe output editing (ISO Fortran 2008 10.7.5.2.2:
F editing inside 0.1 <= |x| < 10**d after rounding, E editing outside,
n = 4 or e+2 trailing blanks) and Iw editing, with round-to-nearest on the
exact binary value, as gfortran prints them.  The values come from the
device (Solver.stats_record, the equalize record, staged snapshot fields).
"""
from decimal import ROUND_HALF_EVEN, Decimal

_HALF = Decimal("0.5")


def _digits(mag, places):
    """mag (a non-negative Decimal) rounded to `places` decimals, as text
    with a leading zero before the point."""
    q = mag.quantize(Decimal(1).scaleb(-places), rounding=ROUND_HALF_EVEN)
    txt = format(q, "f")
    if places == 0:
        txt += "."
    return txt


def e_format(x, w, d, e=None):
    """Ew.d / Ew.dEe: [-]0.d1..dd E+xx (exponent of 3 digits without the E
    letter when |exp| > 99 and e is not given)."""
    v = Decimal(float(x))
    neg = v.is_signed()
    mag = abs(v)
    if mag == 0:
        mant, exp = "0." + "0" * d, 0
    else:
        exp = mag.adjusted() + 1
        m = (mag.scaleb(-exp)).quantize(Decimal(1).scaleb(-d), rounding=ROUND_HALF_EVEN)
        if m >= 1:
            exp += 1
            m = (mag.scaleb(-exp)).quantize(Decimal(1).scaleb(-d), rounding=ROUND_HALF_EVEN)
        mant = format(m, "f")
    sign_e = "-" if exp < 0 else "+"
    if e is None:
        ex = f"E{sign_e}{abs(exp):02d}" if abs(exp) <= 99 else f"{sign_e}{abs(exp):03d}"
    else:
        ex = f"E{sign_e}{abs(exp):0{e}d}"
        if len(ex) > e + 2:
            return "*" * w
    s = ("-" if neg else "") + mant + ex
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def g_format(x, w, d, e=None):
    """Gw.d (e None) or Gw.dEe output editing of one real."""
    n = 4 if e is None else e + 2
    v = Decimal(float(x))
    neg = v.is_signed()
    mag = abs(v)
    if mag == 0:
        body = _digits(mag, d - 1)
    else:
        lo = Decimal("0.1") - _HALF * Decimal(1).scaleb(-d - 1)
        hi = Decimal(10) ** d - _HALF
        if mag < lo or mag >= hi:
            return e_format(x, w, d, e)
        j = 0
        while not (mag < Decimal(10) ** j - _HALF * Decimal(1).scaleb(j - d)):
            j += 1
        body = _digits(mag, d - j)
    s = ("-" if neg else "") + body
    if len(s) > w - n:
        return "*" * w
    return s.rjust(w - n) + " " * n


def i_format(v, w):
    s = str(int(v))
    return "*" * w if len(s) > w else s.rjust(w)


def stats_line(rec, comp, ntype, tracer=0, tstep=None, RaT=1.0):
    """One stats.dat record (basics.f90:284-312) from a TcStatsRow; tstep
    replaces time in the local-time-step forms (i10)."""
    q = list(rec.q)
    g = lambda x: g_format(x, 20, 8)
    if RaT != 0.0:
        head = [rec.Tavg, rec.vrms, rec.Nu, rec.Nu_bot, rec.RaT_avg, rec.mobility, rec.strain_max] + q
        first = i_format(tstep, 10) if tstep is not None else g(rec.time)
        out = first + "".join(g(x) for x in head)
        counts = "".join(i_format(v, 6) for v in (rec.empty_cells, rec.tracer_min, rec.tracer_max))
        if comp == 0:
            return out + (counts if tracer == 1 else "")
        if tstep is not None:
            return out + counts
        return out + counts + "".join(g(x) for x in list(rec.Cavg)[:ntype] + list(rec.ent)[:ntype])
    first = i_format(tstep, 10) if tstep is not None else g(rec.time)
    out = first + "".join(g(x) for x in (rec.vrms, rec.mobility, rec.strain_max))
    if comp == 0 or tstep is not None:
        return out
    counts = "".join(i_format(v, 6) for v in (rec.empty_cells, rec.tracer_min, rec.tracer_max))
    return out + counts + "".join(g(x) for x in list(rec.Cavg)[:ntype] + list(rec.ent)[:ntype])


def viscosity_line(time, n_eta, eig_fine, n_temp, n_comp, comp, RaT=1.0):
    """One viscosity_smoother.dat record (basics.f90:316-322); None when the
    reference writes none (RaT=0, comp=0)."""
    g = lambda x: g_format(x, 20, 8)
    if RaT != 0.0 and comp == 0:
        return g(time) + i_format(n_eta, 5) + g(eig_fine) + i_format(n_temp, 5)
    if RaT != 0.0:
        return (g(time) + i_format(n_eta, 5) + g(eig_fine) + i_format(n_temp, 5) +
                "".join(i_format(v, 5) for v in n_comp))
    if comp == 1:
        return g(time) + i_format(n_eta, 5) + g(eig_fine) + "".join(i_format(v, 5) for v in n_comp)
    return None


def equalize_line(time, info, comp, ntype, tpc=0):
    """One equalize.dat record (composition.f90:566-574) from a
    TcEqualizeInfo: time, tpc (comp=0) or tpc_type(0:ntype), minexcess,
    the deficit, the tracers moved, nfree and the per-type counts."""
    nt = ntype if comp == 1 else 0
    ints = [tpc] if comp == 0 else list(info.tpc_type)[:nt + 1]
    ints += list(info.minexcess)[:nt + 1] + [info.deficit, info.moved, info.nfree] + list(info.count)[:nt + 1]
    return g_format(time, 20, 8) + "".join(i_format(v, 10) for v in ints)


def snapshot_name(fcount):
    """'T' followed by 10000+fcount in i5 (main.f90:240)."""
    return "T" + i_format(10000 + fcount, 5)


def snapshot_lines(xg, zg, fields, comp, ntype, tracer=0, RaT=1.0):
    """The records of one snapshot (main.f90:241-266): for each i, one line
    per k, then the list-directed blank record.  fields holds interior
    arrays indexed [k, i]: T, SF, u, w, strain, vis, Tvis, nempty, and
    Cnew, Cvis as [type, k, i]."""
    g = lambda x: g_format(x, 14, 5, 3)
    nz, nx = fields["SF"].shape[0] - 1, fields["SF"].shape[1] - 1
    f = fields
    lines = []
    for i in range(nx + 1):
        for k in range(nz + 1):
            if RaT != 0.0:
                base = [xg[i], zg[k], f["T"][k, i], f["SF"][k, i], f["u"][k, i], f["w"][k, i]]
                if comp == 0 and tracer == 0:
                    lines.append("".join(g(x) for x in base + [f["vis"][k, i], f["Tvis"][k, i]]))
                    continue
                row = "".join(g(x) for x in base + [f["strain"][k, i], f["vis"][k, i], f["Tvis"][k, i]])
                row += i_format(f["nempty"][k, i], 6)
                if comp == 0:
                    lines.append(row)
                    continue
            else:
                row = "".join(g(x) for x in (xg[i], zg[k], f["SF"][k, i], f["u"][k, i], f["w"][k, i],
                                             f["strain"][k, i], f["vis"][k, i]))
                row += i_format(f["nempty"][k, i], 6)
            row += "".join(g(f["Cnew"][t, k, i]) for t in range(ntype))
            row += "".join(g(f["Cvis"][t, k, i]) for t in range(ntype))
            lines.append(row)
        lines.append("  ")
    return lines


def staged_snapshot_fields(solver):
    """The staged snapshot fields of a Solver (after snapshot_stage) as the
    interior arrays snapshot_lines takes."""
    sl = (slice(2, -2), slice(2, -2))
    f = {k: solver.snapshot_get(k)[sl] for k in ("T", "SF", "u", "w", "strain", "vis", "Tvis", "nempty")}
    nt = solver.params.ntype if solver.params.comp == 1 else 0
    if nt:
        for k in ("Cnew", "Cvis"):
            f[k] = solver.snapshot_get(k)[2:-2, 2:-2, :].transpose(2, 0, 1)
    return f
