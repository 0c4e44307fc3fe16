"""CPU reference of the learned-guess U-Net (test infrastructure).

Same architecture and conventions as tracer_code_amd/csrc/tc_nn.hip:
* NHWC activations, pixel p = k*W + i;
* 3x3 "same" convolutions with zero padding, tap = (dy+1)*3 + (dx+1);
* tanh-form GELU;
* 2x2 average pooling and nearest 2x upsampling, concatenation [up | skip];
* a 1x1 head 16 -> 1;
* masked MSE loss.

Parameters are one flat vector: per layer W[cout][9][cin] then b[cout], and
last the head w[16], b.  `bf16=True` rounds weights, inputs, activations,
pre-activations and activation gradients to bfloat16 at the points where the
GPU stores them, so that the two can be compared closely; `bf16=False` is
the plain float64 network that finite differences pin.
"""
import numpy as np

LAYERS = [(6, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (64, 128, 3), (128, 128, 3),
          (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1), (48, 16, 0), (16, 16, 0)]


def param_count():
    return sum(co * 9 * ci + co for ci, co, _ in LAYERS) + 17


def split_params(theta):
    out, o = [], 0
    for ci, co, _ in LAYERS:
        w = theta[o:o + co * 9 * ci].reshape(co, 9, ci)
        o += co * 9 * ci
        b = theta[o:o + co]
        o += co
        out.append((w, b))
    head_w, head_b = theta[o:o + 16], theta[o + 16]
    return out, head_w, head_b


def bf16(x):
    """Round to bfloat16 (nearest even) and return float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co, _ in LAYERS:
        parts.append(rng.standard_normal(co * 9 * ci) * np.sqrt(2.0 / (9 * ci)) * scale)
        parts.append(np.zeros(co))
    parts.append(rng.standard_normal(16) * np.sqrt(1.0 / 16) * scale)
    parts.append(np.zeros(1))
    return np.concatenate(parts)


GA, GB = 0.7978845608028654, 0.044715


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GA * (z + GB * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GA * (z + GB * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GA * (1.0 + 3.0 * GB * z * z)


def _patches(x):
    """x (H, W, C) -> (H, W, 9, C) neighbours with zero padding."""
    H, W, C = x.shape
    xp = np.zeros((H + 2, W + 2, C))
    xp[1:-1, 1:-1] = x
    out = np.empty((H, W, 9, C))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out[:, :, t] = xp[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
    return out


def conv(x, w, b):
    return np.einsum("hwtc,otc->hwo", _patches(x), w) + b


def conv_back(x, w, dz):
    """Gradients of z = conv(x, w, b) given dz: (dx, dw, db)."""
    P = _patches(x)
    dw = np.einsum("hwo,hwtc->otc", dz, P)
    db = dz.sum(axis=(0, 1))
    H, W, _ = x.shape
    dP = np.einsum("hwo,otc->hwtc", dz, w)
    dxp = np.zeros((H + 2, W + 2, x.shape[2]))
    for t in range(9):
        dy, dx_ = t // 3 - 1, t % 3 - 1
        dxp[1 + dy:1 + dy + H, 1 + dx_:1 + dx_ + W] += dP[:, :, t]
    return dxp[1:-1, 1:-1], dw, db


def pool(x):
    H, W, C = x.shape
    return x.reshape(H // 2, 2, W // 2, 2, C).mean(axis=(1, 3))


def pool_back(d):
    return np.repeat(np.repeat(d, 2, axis=0), 2, axis=1) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up_back(d):
    H, W, C = d.shape
    return d.reshape(H // 2, 2, W // 2, 2, C).sum(axis=(1, 3))


class UNetRef:
    def __init__(self, theta, bf=False):
        self.bf = bf
        self.r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
        layers, hw, hb = split_params(np.asarray(theta, dtype=np.float64))
        self.layers = [(self.r(w), b) for w, b in layers]
        self.hw, self.hb = hw, hb

    def _layer(self, j, x):
        # the stored pre-activation is rounded; the activation is taken of
        # the unrounded sum
        w, b = self.layers[j]
        zf = conv(x, w, b)
        return self.r(zf), self.r(gelu(zf))

    def forward(self, x):
        """x (H, W, 6) -> output (H, W); keeps what backward needs."""
        r = self.r
        s = {"x0": r(x)}
        s["z0"], s["a1"] = self._layer(0, s["x0"])
        s["z1"], s["sk1"] = self._layer(1, s["a1"])
        s["p1"] = r(pool(s["sk1"]))
        s["z2"], s["a2"] = self._layer(2, s["p1"])
        s["z3"], s["sk2"] = self._layer(3, s["a2"])
        s["p2"] = r(pool(s["sk2"]))
        s["z4"], s["a3"] = self._layer(4, s["p2"])
        s["z5"], s["sk3"] = self._layer(5, s["a3"])
        s["p3"] = r(pool(s["sk3"]))
        s["z6"], s["a4"] = self._layer(6, s["p3"])
        s["z7"], s["b4"] = self._layer(7, s["a4"])
        s["c3"] = np.concatenate([up(s["b4"]), s["sk3"]], axis=2)
        s["z8"], s["d3a"] = self._layer(8, s["c3"])
        s["z9"], s["d3b"] = self._layer(9, s["d3a"])
        s["c2"] = np.concatenate([up(s["d3b"]), s["sk2"]], axis=2)
        s["z10"], s["d2a"] = self._layer(10, s["c2"])
        s["z11"], s["d2b"] = self._layer(11, s["d2a"])
        s["c1"] = np.concatenate([up(s["d2b"]), s["sk1"]], axis=2)
        s["z12"], s["d1a"] = self._layer(12, s["c1"])
        s["z13"], s["d1b"] = self._layer(13, s["d1a"])
        self.s = s
        return s["d1b"] @ self.hw + self.hb

    def loss_and_grad(self, x, target, mask):
        """Masked MSE (sum over mask / count) and the flat parameter gradient."""
        out = self.forward(x)
        s, r = self.s, self.r
        n = mask.sum()
        d = (out - target) * mask
        loss = float((d * d).sum() / n)
        dout = 2.0 * d / n
        grads = [None] * len(LAYERS)
        ghw = np.einsum("hw,hwc->c", dout, s["d1b"])
        ghb = dout.sum()
        g = dout[:, :, None] * self.hw[None, None, :]

        def back(j, gy, xin):
            dz = r(gy * gelu_grad(s[f"z{j}"]))
            dx, dw, db = conv_back(xin, self.layers[j][0], dz)
            grads[j] = (dw, db)
            return dx

        g = back(13, g, s["d1a"])
        gc1 = back(12, g, s["c1"])
        g_sk1 = gc1[:, :, 32:]
        g = up_back(gc1[:, :, :32])
        g = back(11, g, s["d2a"])
        gc2 = back(10, g, s["c2"])
        g_sk2 = gc2[:, :, 64:]
        g = up_back(gc2[:, :, :64])
        g = back(9, g, s["d3a"])
        gc3 = back(8, g, s["c3"])
        g_sk3 = gc3[:, :, 128:]
        g = up_back(gc3[:, :, :128])
        g = back(7, g, s["a4"])
        g = back(6, g, s["p3"])
        g_sk3 = g_sk3 + pool_back(g)
        g = back(5, g_sk3, s["a3"])
        g = back(4, g, s["p2"])
        g_sk2 = g_sk2 + pool_back(g)
        g = back(3, g_sk2, s["a2"])
        g = back(2, g, s["p1"])
        g_sk1 = g_sk1 + pool_back(g)
        g = back(1, g_sk1, s["a1"])
        back(0, g, s["x0"])
        flat = []
        for dw, db in grads:
            flat.append(dw.ravel())
            flat.append(db)
        flat.append(ghw)
        flat.append(np.array([ghb]))
        return loss, out, np.concatenate(flat)
