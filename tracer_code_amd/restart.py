"""The reference's restart files through the library's codec
(tc_restart_*): Trestart (xg, zg, T, Cnew, SF in one Fortran unformatted
record, main.f90:301-332 / grid.f90:56-95) and restart_tracers (rtr, str,
Ttype).  Grid fields are numpy arrays (nz+5, nx+5) of the ghosted
(-2:nx+2, -2:nz+2) layout; Cnew is (nz+5, nx+5, ntype)."""
import numpy as np

from .model import library


def _p(a):
    return a.ctypes.data if a is not None else None


def read_trestart(path, nx, nz, has_T=True, ntype_c=0):
    lib = library()
    xg = np.empty(nx + 7)
    zg = np.empty(nz + 7)
    T = np.empty((nz + 5, nx + 5)) if has_T else None
    C = np.empty((nz + 5, nx + 5, ntype_c)) if ntype_c else None
    SF = np.empty((nz + 5, nx + 5))
    st = lib.tc_restart_read(str(path).encode(), nx, nz, int(bool(has_T)), ntype_c, _p(xg), _p(zg), _p(T), _p(C),
                             _p(SF))
    if st != 0:
        raise ValueError(f"tc_restart_read({path}): status {st}")
    return dict(xg=xg, zg=zg, T=T, Cnew=C, SF=SF)


def write_trestart(path, xg, zg, T, Cnew, SF):
    lib = library()
    nz, nx = SF.shape[0] - 5, SF.shape[1] - 5
    arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    xg, zg, T, Cnew, SF = map(arr, (xg, zg, T, Cnew, SF))
    st = lib.tc_restart_write(str(path).encode(), nx, nz, int(T is not None), 0 if Cnew is None else Cnew.shape[-1],
                              _p(xg), _p(zg), _p(T), _p(Cnew), _p(SF))
    if st != 0:
        raise ValueError(f"tc_restart_write({path}): status {st}")


def read_tracers(path, ntr):
    lib = library()
    r, s, t = np.empty(ntr), np.empty(ntr), np.empty(ntr, dtype=np.int32)
    st = lib.tc_restart_tracers_read(str(path).encode(), int(ntr), _p(r), _p(s), _p(t))
    if st != 0:
        raise ValueError(f"tc_restart_tracers_read({path}): status {st}")
    return r, s, t


def write_tracers(path, r, s, t):
    lib = library()
    r = np.ascontiguousarray(r, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.int32)
    st = lib.tc_restart_tracers_write(str(path).encode(), len(r), _p(r), _p(s), _p(t))
    if st != 0:
        raise ValueError(f"tc_restart_tracers_write({path}): status {st}")
