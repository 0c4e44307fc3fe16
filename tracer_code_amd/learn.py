"""The learned initial guess: network handle, weights and the train/evaluate
workflow.

    python -m tracer_code_amd.learn train --config cfg1 --nx 768 --nz 256 --level 4 \\
        --seeds 20 --steps 25 --epochs 30 --ra-range 3e5 3e6 --out tracer_code_amd/weights/unet_cfg1_L4.npz
    python -m tracer_code_amd.learn evaluate --config cfg1 --nx 768 --nz 256 \\
        --weights tracer_code_amd/weights/unet_cfg1_L4.npz --seed 100 --steps 6 --ra 1e6

The network (a 4-level U-Net, channels 16/32/64/128, BF16 on the MFMA
units) runs and trains entirely in the HIP library (tc_nn_*, tc_learn_*);
this module only drives it.  It works on one multigrid level (`--level`,
its grid is that level's): before each solve the inputs (T, C, log10 eta,
the two previous solutions, the change of T, all restricted to the level)
are recorded on the device, after it the increment of the restricted
solution; as a guess, the predicted increment is prolonged to the fine grid.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))

# (cin, cout) per 3x3 layer in forward order; the 1x1 head 16 -> 1 follows
LAYERS = [(6, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (192, 64), (64, 64),
          (96, 32), (32, 32), (48, 16), (16, 16)]


def param_count():
    return sum(co * 9 * ci + co for ci, co in LAYERS) + 17


def init_weights(seed=0):
    """He-normal weights, zero biases, and a zero head (the untrained
    network predicts no increment)."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in LAYERS:
        parts.append(rng.standard_normal(co * 9 * ci) * np.sqrt(2.0 / (9 * ci)))
        parts.append(np.zeros(co))
    parts.append(np.zeros(17))
    return np.concatenate(parts).astype(np.float32)


def save_weights(path, w, meta):
    np.savez(path, weights=np.asarray(w, dtype=np.float32), meta=json.dumps(meta))


def load_weights(path):
    with np.load(path, allow_pickle=False) as f:
        return f["weights"].astype(np.float32), json.loads(str(f["meta"]))


class Network:
    """Owning handle of a tc_nn (H x W node grid)."""

    def __init__(self, H, W, trainable=True, lib=None):
        from .model import library
        self.lib = lib or library()
        self.H, self.W = int(H), int(W)
        h = ctypes.c_void_p()
        st = self.lib.tc_nn_create(self.H, self.W, 1 if trainable else 0, ctypes.byref(h))
        if st != 0:
            raise RuntimeError(f"tc_nn_create failed: status {st}")
        self.h = h
        n = ctypes.c_int64()
        self._check(self.lib.tc_nn_param_count(self.h, ctypes.byref(n)), "param_count")
        self.nparam = n.value

    def _check(self, st, what):
        if st != 0:
            msg = self.lib.tc_nn_last_error(self.h)
            raise RuntimeError(f"tc_nn {what}: status {st}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.tc_nn_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.tc_nn_set_weights(self.h, w.ctypes.data, w.size), "set_weights")

    def get_weights(self):
        w = np.empty(self.nparam, dtype=np.float32)
        self._check(self.lib.tc_nn_get_weights(self.h, w.ctypes.data, w.size), "get_weights")
        return w

    def get_grads(self):
        g = np.empty(self.nparam, dtype=np.float32)
        self._check(self.lib.tc_nn_get_grads(self.h, g.ctypes.data, g.size), "get_grads")
        return g

    def set_sample(self, x, target=None, mask=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.shape == (self.H, self.W, 6)
        t = None if target is None else np.ascontiguousarray(target, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        self._check(self.lib.tc_nn_set_sample(self.h, x.ctypes.data, None if t is None else t.ctypes.data,
                                              None if m is None else m.ctypes.data), "set_sample")

    def forward(self):
        out = np.empty((self.H, self.W), dtype=np.float32)
        self._check(self.lib.tc_nn_forward(self.h, out.ctypes.data), "forward")
        return out

    def backward(self):
        loss = ctypes.c_double()
        self._check(self.lib.tc_nn_backward(self.h, ctypes.byref(loss)), "backward")
        return loss.value

    def adam(self, lr, grad_scale=1.0):
        self._check(self.lib.tc_nn_adam(self.h, float(lr), float(grad_scale)), "adam")


def level_shape(nx, nz, level):
    """(H, W) of the network on multigrid level `level` (1 = the fine grid)."""
    return nz >> (level - 1), nx >> (level - 1)


def _solver(config, nx, nz, seed, tpc=None, ra=None):
    """A Solver on the named workload; `ra` replaces RaT and scales RaC with
    it (the buoyancy ratio stays)."""
    from . import Solver, config as cfgmod, ic
    ov = dict(nx=nx, nz=nz)
    if tpc:
        ov["tpc"] = tpc
    if ra:
        base = cfgmod.PRESETS[config]
        ov["RaT"] = float(ra)
        ov["RaC"] = tuple(c * float(ra) / base["RaT"] for c in base["RaC"])
    params, cfg = cfgmod.preset_params(config, **ov)
    T, r, s, ty = ic.workload(cfg, seed=seed)
    g = Solver(params, len(r))
    g.initialize(T, r, s, ty)
    return params, g


def draw_ra(ra_range, seed):
    """Ra log-uniform in ra_range (None: the preset's)."""
    if not ra_range:
        return None
    rng = np.random.default_rng(1000 + seed)
    lo, hi = np.log(ra_range[0]), np.log(ra_range[1])
    return float(np.exp(rng.uniform(lo, hi)))


def train(config, nx, nz, seeds, steps, epochs, lr, batch, out, mode="mixed", tpc=None, ra_range=None, level=4,
          dc_sweeps=0, log=print):
    """Record the solves of `steps` time steps from each of `seeds` seeded
    starts (Ra drawn per seed from ra_range), then train the network on all
    of them on the GPU -- each epoch visits every run's samples in a fresh
    shuffled order, with a cosine learning-rate decay to lr/20 -- and save
    the weights."""
    net = Network(*level_shape(nx, nz, level), trainable=True)
    net.set_weights(init_weights(0))
    total = 0
    t0 = time.perf_counter()
    solvers, ras = [], []
    for sd in range(seeds):
        ra = draw_ra(ra_range, sd)
        _, g = _solver(config, nx, nz, seed=sd, tpc=tpc, ra=ra)
        g.set_mode(mode)
        g.set_dc_sweeps(dc_sweeps)
        g.learn_attach(net)
        g.learn_record(4 * steps)
        g.time_step(steps)
        n = g.learn_samples()
        total += n
        solvers.append(g)
        ras.append(ra)
        log(f"[learn] run {sd} (Ra {ra if ra else 'preset'}): {n} samples ({time.perf_counter() - t0:.1f}s)")
    losses = []
    for ep in range(epochs):
        lr_ep = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * ep / max(epochs - 1, 1))))
        order = np.random.default_rng(ep).permutation(len(solvers))
        ls = [solvers[j].learn_train(1, lr_ep, batch, seed=ep * 1000 + int(j)) for j in order]
        losses.append(float(np.mean(ls)))
        log(f"[learn] epoch {ep}: lr {lr_ep:.2e}, mean loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f}s)")
    for g in solvers:
        g.learn_record(0)
        g.learn_attach(None)
        g.close()
    w = net.get_weights()
    net.close()
    meta = dict(config=config, nx=nx, nz=nz, seeds=seeds, steps=steps, epochs=epochs, lr=lr, batch=batch,
                samples=total, losses=losses, mode=mode, tpc=tpc, ra=ras, level=level, dc_sweeps=dc_sweeps)
    if out:
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        save_weights(out, w, meta)
    return w, meta


def evaluate(config, nx, nz, weights, seed, steps, mode="mixed", tpc=None, warm=3, ra=None, level=4, dc_sweeps=0):
    """Multigrid cycles and time per step from the same seeded state with
    each initial-guess mode: "plain" (the reference's warm start), "extrap"
    (plain or the stage increment: plain plus the last step's change between
    the same two RK stages) and "learned" (also the stage's extrapolation in
    time plus the network's correction of it).  `warm` steps first build
    the per-stage history.  Reports which start won how often and the median
    residual after the first level-1 smoothing of each candidate relative to
    the plain start (over the solves where both were judged)."""
    res = {}
    for gm in ("plain", "extrap", "learned"):
        params, g = _solver(config, nx, nz, seed=seed, tpc=tpc, ra=ra)
        g.set_mode(mode)
        g.set_dc_sweeps(dc_sweeps)
        net = None
        if gm == "learned":
            net = Network(*level_shape(nx, nz, level), trainable=False)
            net.set_weights(weights)
            g.learn_attach(net)
        g.guess_enable(gm)
        g.time_step(warm)
        st0 = g.stats()
        g.sync()
        t0 = time.perf_counter()
        rel = {1: [], 2: []}
        for _ in range(steps):
            g.time_step(1)
            st = g.stats()
            for c in (1, 2):
                if st.guess_res[0] > 0 and st.guess_res[c] > 0:
                    rel[c].append(st.guess_res[c] / st.guess_res[0])
        g.sync()
        el = time.perf_counter() - t0
        st1 = g.stats()
        r = dict(cycles_per_step=(st1.mg_cycles - st0.mg_cycles) / steps, ms_per_step=1e3 * el / steps,
                 starts={k: int(st1.guesses_used[i] - st0.guesses_used[i])
                         for i, k in enumerate(("plain", "extrap", "learned"))})
        for c, k in ((1, "extrap"), (2, "learned")):
            if rel[c]:
                r[f"median_{k}_vs_plain_first_smoothing"] = float(np.median(rel[c]))
        res[gm] = r
        if net is not None:
            g.learn_attach(None)
            net.close()
        g.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "evaluate"])
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--tpc", type=int, default=None)
    ap.add_argument("--seeds", type=int, default=4)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--mode", default="mixed")
    ap.add_argument("--ra-range", type=float, nargs=2, default=None)
    ap.add_argument("--ra", type=float, default=None)
    ap.add_argument("--level", type=int, default=4)
    ap.add_argument("--dc-sweeps", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--weights", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        _, meta = train(a.config, a.nx, a.nz, a.seeds, a.steps, a.epochs, a.lr, a.batch, a.out, a.mode, a.tpc,
                        a.ra_range, a.level, a.dc_sweeps, log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(meta))
    else:
        w, meta = load_weights(a.weights)
        print(json.dumps(evaluate(a.config, a.nx, a.nz, w, a.seed, a.steps, a.mode, a.tpc, ra=a.ra,
                                  level=meta.get("level", a.level), dc_sweeps=a.dc_sweeps)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
