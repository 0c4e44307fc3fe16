"""GPU checks of the learned-guess U-Net (BF16 on MFMA) against the CPU
reference oracle/unet.py, and of its use inside the solver.

* forward output and parameter gradients equal the reference evaluated with
  the same bfloat16 roundings (tolerances below), and stay close to the
  plain float64 reference;
* one Adam step equals the textbook update;
* in a defect-correction solve the guess never changes the answer: the
  solve still ends at the tolerance, and the solver reports whether the
  guess or the plain warm start was used.
"""
import numpy as np
import pytest

from oracle import unet
from tracer_code_amd import Solver, config, ic, learn

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.parametrize("H,W", [(16, 16), (32, 64)])
def test_forward_and_gradients_match_reference(H, W):
    th = unet.init_params(3)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (H, W, 6))
    t = rng.standard_normal((H, W))
    m = np.ones((H, W))
    m[0, :] = 0
    m[:, 0] = 0
    print(f"[nn grad check {H}x{W}]", flush=True)
    with learn.Network(H, W) as net:
        net.set_weights(th.astype(np.float32))
        net.set_sample(x, t, m)
        out = net.forward()
        loss = net.backward()
        g = net.get_grads()
    lb, ob, gb = unet.UNetRef(th.astype(np.float32), bf=True).loss_and_grad(x, t, m)
    l64, o64, g64 = unet.UNetRef(th.astype(np.float32)).loss_and_grad(x, t, m)
    e_out, e_out64 = rel(out, ob), rel(out, o64)
    e_g, e_g64 = rel(g, gb), rel(g, g64)
    print(f"out rel err vs bf16-ref {e_out:.2e} vs fp64-ref {e_out64:.2e}; loss {loss:.6e} ref {lb:.6e}; "
          f"grad rel err vs bf16-ref {e_g:.2e} vs fp64-ref {e_g64:.2e}", flush=True)
    assert e_out <= 1e-2 and e_out64 <= 5e-2
    assert abs(loss - lb) <= 2e-2 * abs(lb)
    assert e_g <= 3e-2 and e_g64 <= 1e-1
    # layer by layer: every block of parameters carries the gradient
    o = 0
    for ci, co, _ in unet.LAYERS:
        n = co * 9 * ci + co
        assert rel(g[o:o + n], gb[o:o + n]) <= 6e-2, (ci, co)
        o += n


def test_adam_step():
    H, W = 16, 16
    th = unet.init_params(5).astype(np.float32)
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (H, W, 6))
    t = rng.standard_normal((H, W))
    m = np.ones((H, W))
    with learn.Network(H, W) as net:
        net.set_weights(th)
        net.set_sample(x, t, m)
        net.forward()
        net.backward()
        g = net.get_grads().astype(np.float64)
        net.adam(1e-3, 0.5)
        w1 = net.get_weights().astype(np.float64)
    gs = 0.5 * g
    mm = 0.1 * gs
    vv = 0.001 * gs * gs
    want = th - 1e-3 * (mm / 0.1) / (np.sqrt(vv / 0.001) + 1e-8)
    assert np.max(np.abs(w1 - want)) <= 1e-6


@pytest.mark.parametrize("level", [1, 3])
def test_guess_in_solver_keeps_tolerance(level):
    params, cfg = config.preset_params("cfg0", nx=64, nz=64, tpc=4)
    T, r, s, ty = ic.workload(cfg, seed=11)
    print(f"[learned guess 64x64, network on level {level}]", flush=True)
    with Solver(params, len(r)) as g, learn.Network(*learn.level_shape(params.nx, params.nz, level)) as net:
        net.set_weights(learn.init_weights(0))
        g.initialize(T, r, s, ty)
        g.set_mode("mixed")
        g.learn_attach(net)
        g.learn_record(64)
        # samples need two earlier solves of the same RK stage and one
        # measured extrapolation error: steps 4.. give 4 each
        g.time_step(6)
        n = g.learn_samples()
        assert n == 12
        loss0 = g.learn_train(1, 0.0, 2, seed=1)   # lr 0: the untrained network's loss
        loss1 = g.learn_train(20, 1e-3, 2, seed=2)
        print(f"samples {n}, loss untrained {loss0:.3e}, after 20 epochs {loss1:.3e}", flush=True)
        assert np.isfinite(loss1) and loss1 < loss0
        g.guess_enable("learned")
        st0 = g.stats()
        judged = 0
        for _ in range(3):
            g.time_step(1)
            st = g.stats()
            assert st.residual_last <= params.tolerance * st.res_ref_last
            # the learned mode judges the plain start, the stage increment
            # and the learned guess; the plain start is always judged
            judged += st.guess_res[2] > 0
            if st.guess_res[1] > 0 or st.guess_res[2] > 0:
                assert st.guess_res[0] > 0
        starts = [st.guesses_used[i] - st0.guesses_used[i] for i in range(3)]
        print(f"starts over the last 3 steps (plain, stage increment, learned): {starts}; last solves judged: {judged}",
              flush=True)
        assert sum(starts) == 12 and judged >= 1
        g.learn_attach(None)


def test_trained_network_gain_on_held_out_run():
    """What the trained network gains: the committed weights (768x256 grid,
    network on multigrid level 4, trained on 720 solves of 20 seeded runs
    with Ra log-uniform in [3e5, 3e6]) against the plain warm start and the
    stage increment alone, on a held-out seed.  Every start is judged by the
    residual after the first level-1 smoothing, so the learned form can only
    be chosen where it is better there; the test reports the cycles and
    times of the three forms and checks that the learned form needs no more
    cycles than the plain one (up to 2%).  Every solve meets the tolerance
    (a solve that does not raises)."""
    import os
    path = os.path.join(os.path.dirname(learn.__file__), "weights", "unet_cfg1_L4.npz")
    w, meta = learn.load_weights(path)
    print(f"[trained network: {meta['samples']} samples, level {meta['level']}]", flush=True)
    res = learn.evaluate("cfg1", 768, 256, w, seed=100, steps=2, tpc=10, ra=1e6, level=meta["level"], warm=3)
    for k, v in res.items():
        print(f"{k}: {v}", flush=True)
    assert res["learned"]["cycles_per_step"] <= 1.02 * res["plain"]["cycles_per_step"]
    assert res["extrap"]["cycles_per_step"] <= 1.02 * res["plain"]["cycles_per_step"]
    assert sum(res["learned"]["starts"].values()) == 8
