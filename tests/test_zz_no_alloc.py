"""tc_bin_tracers_noalloc (binning on caller-owned device buffers, no
allocation) against tc_tracers_to_corners with smooth=0 on the same tracers:
count_type, nempty and Cnew bit-identical.  Two calls in a row through the
C-ABI, the second with a smaller grid and fewer tracers, so that counts left
in the per-GPU work array by the first call must not leak into the second;
both with the counts kept in the work array (count_type NULL) and in a
caller plane; and the out-of-domain status."""
import ctypes

import numpy as np
import pytest

from tracer_code_amd import Solver, _abi, config, ic

pytestmark = pytest.mark.gpu


def _lib():
    return _abi.load_library()


def _reference(params, cfg, seed):
    """Binning by the context entry point: (r, s, ty, Cnew before, count_type, nempty, Cnew after)."""
    T, r, s, ty = ic.workload(cfg, seed=seed)
    with Solver(params, len(r)) as g:
        g.initialize(T, r, s, ty)
        # move the tracers a little so that some nodes empty and some change type mix
        rng = np.random.default_rng(seed)
        r2 = np.clip(r + (rng.random(len(r)) - 0.5) * 1.5 * params.aspect / params.nx, 1e-6, params.aspect - 1e-6)
        s2 = np.clip(s + (rng.random(len(s)) - 0.5) * 1.5 / params.nz, 1e-6, 1 - 1e-6)
        keep = rng.random(len(r)) > 0.3
        r2, s2 = np.where(keep, r, r2), np.where(keep, s, s2)
        g.put_tracers(r2, s2, ty)
        cnew_before = g.get("Cnew")
        g.tracers_to_corners("x", smooth=False)
        return r2, s2, ty, cnew_before, g.get("count_type"), g.get("nempty"), g.get("Cnew")


class _Hip:
    """Device buffers for the test through the HIP runtime (the test, not the
    library, allocates them)."""

    def __init__(self):
        import os
        self.rt = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1))) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        self.live.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _noalloc(params, r, s, ty, cnew_before, with_counts, shift=None):
    nx, nz, nt = params.nx, params.nz, params.ntype
    plane = (nz + 5) * (nx + 5)
    hip = _Hip()
    if shift is not None:
        r = r.copy()
        r[shift] = -0.1
    R, S, TY = hip.put(r.astype(np.float64)), hip.put(s.astype(np.float64)), hip.put(ty.astype(np.int32))
    cnt0 = np.full((nt + 1) * plane, 7, dtype=np.int32)
    ne0 = np.full(plane, 9, dtype=np.int32)
    cn0 = np.ascontiguousarray(np.moveaxis(cnew_before, -1, 0))  # type-fastest host layout -> one plane per type
    CNT = hip.put(cnt0) if with_counts else None
    NE, CN = hip.put(ne0), hip.put(cn0)
    st = _lib().tc_bin_tracers_noalloc(nx, nz, params.aspect, nt, params.comp, R, S, TY, len(r), CNT, NE, CN, None)
    out_cnt = None
    if with_counts:
        out_cnt = np.moveaxis(hip.get(CNT, cnt0).reshape(nt + 1, nz + 5, nx + 5), 0, -1)
    ne = hip.get(NE, ne0).reshape(nz + 5, nx + 5)
    cn = np.moveaxis(hip.get(CN, cn0).reshape(nt, nz + 5, nx + 5), 0, -1)
    hip.free()
    return st, out_cnt, ne, cn


CASES = [("cfg2", dict(nx=256, nz=128, tpc=6), 3), ("cfg0", dict(nx=64, nz=32, aspect=2.0, tpc=3), 4)]


@pytest.mark.parametrize("with_counts", [False, True])
def test_noalloc_binning_matches_context_binning(with_counts):
    for preset, ov, seed in CASES:   # second call: smaller grid, fewer tracers
        params, cfg = config.preset_params(preset, **ov)
        r, s, ty, cnew_before, ct_ref, ne_ref, cn_ref = _reference(params, cfg, seed)
        print(f"[noalloc {params.nx}x{params.nz} ntr={len(r)} counts={'caller' if with_counts else 'scratch'}]",
              flush=True)
        st, ct, ne, cn = _noalloc(params, r, s, ty, cnew_before, with_counts)
        assert st == 0
        assert np.array_equal(ne, ne_ref)
        assert np.array_equal(cn, cn_ref)
        if with_counts:
            assert np.array_equal(ct, ct_ref)


def test_noalloc_binning_out_of_domain_and_limits():
    params, cfg = config.preset_params("cfg0", nx=64, nz=32, aspect=2.0, tpc=3)
    r, s, ty, cnew_before, _, _, _ = _reference(params, cfg, 5)
    st, _, _, _ = _noalloc(params, r, s, ty, cnew_before, False, shift=17)
    assert st == _abi.TC_ERR_TRACER_DOMAIN
    # a grid whose counts do not fit the work array is refused before any launch
    big = int(np.sqrt(_abi.TC_BIN_NOALLOC_MAX_COUNTS / 2)) + 64
    assert _lib().tc_bin_tracers_noalloc(big, big, 1.0, 1, 1, None, None, None, 0, None, None, None,
                                         None) == _abi.TC_ERR_ARG  # nempty NULL: argument error first
    dummy = ctypes.c_int32(0)
    assert _lib().tc_bin_tracers_noalloc(big, big, 1.0, 1, 0, None, None, None, 0, None,
                                         ctypes.addressof(dummy), None, None) == _abi.TC_ERR_NOMEM
    # after the refused calls the array still serves a valid call
    st, _, ne, _ = _noalloc(params, r, s, ty, cnew_before, False)
    assert st == 0 and int(ne[2:-2, 2:-2].sum()) == len(r)
