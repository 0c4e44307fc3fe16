"""The learned initial guess: network handle, weights and the train/evaluate
workflow.

    python -m tracer_code_amd.learn train --config cfg2 --nx 512 --nz 256 \\
        --seeds 4 --steps 12 --epochs 8 --out tracer_code_amd/weights/unet_cfg2_512x256.npz
    python -m tracer_code_amd.learn evaluate --config cfg2 --nx 512 --nz 256 \\
        --weights tracer_code_amd/weights/unet_cfg2_512x256.npz --seed 100 --steps 6

The network (a 4-level U-Net, channels 16/32/64/128, BF16 on the MFMA
units) runs and trains entirely in the HIP library (tc_nn_*, tc_learn_*);
this module only drives it.  Training data are the solver's own solves:
before each solve the inputs (T, C, log10 eta, the two previous solutions)
are recorded on the device, after it the increment of the solution.
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
LAYERS = [(5, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (192, 64), (64, 64),
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
        assert x.shape == (self.H, self.W, 5)
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


def _solver(config, nx, nz, seed, tpc=None):
    from . import Solver, config as cfgmod, ic
    ov = dict(nx=nx, nz=nz)
    if tpc:
        ov["tpc"] = tpc
    params, cfg = cfgmod.preset_params(config, **ov)
    T, r, s, ty = ic.workload(cfg, seed=seed)
    g = Solver(params, len(r))
    g.initialize(T, r, s, ty)
    return params, g


def train(config, nx, nz, seeds, steps, epochs, lr, batch, out, mode="mixed", tpc=None, log=print):
    """Record the solves of `steps` time steps from each of `seeds` seeded
    starts, train the network on them on the GPU and save the weights."""
    net = Network(nz, nx, trainable=True)
    net.set_weights(init_weights(0))
    total = 0
    t0 = time.perf_counter()
    loss = None
    solvers = []
    for sd in range(seeds):
        params, g = _solver(config, nx, nz, seed=sd, tpc=tpc)
        g.set_mode(mode)
        g.learn_attach(net)
        g.learn_record(4 * steps)
        g.time_step(steps)
        n = g.learn_samples()
        total += n
        log(f"[learn] seed {sd}: {n} samples ({time.perf_counter() - t0:.1f}s)")
        loss = g.learn_train(epochs, lr, batch, seed=sd)
        log(f"[learn] seed {sd}: trained {epochs} epochs, mean loss {loss:.4e} ({time.perf_counter() - t0:.1f}s)")
        g.learn_record(0)
        g.learn_attach(None)
        g.close()
    w = net.get_weights()
    net.close()
    meta = dict(config=config, nx=nx, nz=nz, seeds=seeds, steps=steps, epochs=epochs, lr=lr, batch=batch,
                samples=total, final_loss=loss, mode=mode)
    if out:
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        save_weights(out, w, meta)
    return w, meta


def evaluate(config, nx, nz, weights, seed, steps, mode="mixed", tpc=None, warm=2):
    """Multigrid cycles per solve and time per step with and without the
    learned guess from the same seeded state (`warm` steps first build the
    solution history the guess needs)."""
    res = {}
    for use in (False, True):
        params, g = _solver(config, nx, nz, seed=seed, tpc=tpc)
        g.set_mode(mode)
        net = Network(nz, nx, trainable=False)
        net.set_weights(weights)
        g.learn_attach(net)
        g.time_step(warm)
        g.guess_enable(use)
        st0 = g.stats()
        g.sync()
        t0 = time.perf_counter()
        used = 0
        for _ in range(steps):
            g.time_step(1)
            used += g.stats().guess_used
        g.sync()
        el = time.perf_counter() - t0
        st1 = g.stats()
        cyc = st1.mg_cycles - st0.mg_cycles
        res["learned" if use else "plain"] = dict(cycles_per_step=cyc / steps, ms_per_step=1e3 * el / steps,
                                                  last_guess_used=used)
        g.learn_attach(None)
        g.close()
        net.close()
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
    ap.add_argument("--out", default=None)
    ap.add_argument("--weights", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        _, meta = train(a.config, a.nx, a.nz, a.seeds, a.steps, a.epochs, a.lr, a.batch, a.out, a.mode, a.tpc,
                        log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(meta))
    else:
        w, _ = load_weights(a.weights)
        print(json.dumps(evaluate(a.config, a.nx, a.nz, w, a.seed, a.steps, a.mode, a.tpc)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
