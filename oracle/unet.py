"""The learned preconditioner's network, restated in numpy (f64) as the checker of the HIP forward and backward
kernels. TEST INFRASTRUCTURE ONLY: the shipped library never imports it.

The reference has no learned part (its FLD preconditioners are Jacobi and the multigrid of
crates/sr-engine/shaders/heat.wgsl:1160-1526); the network is this library's own, specified by its workload: a 3-level
U-Net with 16/32/64 channels and 3x3 convolutions, inputs (residual, log D, diagonal), output the correction.

Layout: activations are channels-last arrays (H, W, C); a 3x3 convolution's weights are (3, 3, Cin, Cout) with
zero padding at the image border; biases (Cout,). The network (names match the HIP kernels):

    e1a = relu(conv(x, 3 -> 16));   e1 = relu(conv(e1a, 16 -> 16))            full resolution
    p1  = pool(e1)                                                           2x2 mean
    e2a = relu(conv(p1, 16 -> 32)); e2 = relu(conv(e2a, 32 -> 32))            1/2
    p2  = pool(e2)
    e3a = relu(conv(p2, 32 -> 64)); e3 = relu(conv(e3a, 64 -> 64))            1/4
    d2a = relu(conv(cat(up(e3), e2), 96 -> 32)); d2 = relu(conv(d2a, 32 -> 32))
    d1a = relu(conv(cat(up(d2), e1), 48 -> 16)); d1 = relu(conv(d1a, 16 -> 16))
    out = d1 . w_out + b_out                                                 1x1, 16 -> 1

up() is nearest-neighbour 2x; cat() puts the upsampled channels first. Each function's gradient is checked against
central finite differences in tests/test_oracle_unet.py.
"""
from __future__ import annotations

import numpy as np

LAYERS = [  # name, cin, cout
    ("e1a", 3, 16), ("e1", 16, 16),
    ("e2a", 16, 32), ("e2", 32, 32),
    ("e3a", 32, 64), ("e3", 64, 64),
    ("d2a", 96, 32), ("d2", 32, 32),
    ("d1a", 48, 16), ("d1", 16, 16),
]


def init_params(seed=0, dtype=np.float64):
    """He-normal weights, zero biases; the 1x1 output layer small so the untrained network starts near zero."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, cin, cout in LAYERS:
        p[name + ".w"] = (rng.normal(size=(3, 3, cin, cout)) * np.sqrt(2.0 / (9 * cin))).astype(dtype)
        p[name + ".b"] = np.zeros(cout, dtype)
    p["out.w"] = (rng.normal(size=(16,)) * 0.01).astype(dtype)
    p["out.b"] = np.zeros(1, dtype)
    return p


def conv3(x, w, b):
    """3x3 convolution, zero padding: out[y, x, o] = b[o] + sum_{dy, dx, i} w[dy, dx, i, o] x[y + dy - 1, x + dx - 1, i]."""
    H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    out = np.broadcast_to(b, (H, W, w.shape[3])).astype(np.result_type(x, w)).copy()
    for dy in range(3):
        for dx in range(3):
            out += xp[dy:dy + H, dx:dx + W, :] @ w[dy, dx]
    return out


def conv3_backward(x, w, g):
    """Gradients of conv3 given g = dL/dout: (dL/dx, dL/dw, dL/db)."""
    H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    dw = np.zeros_like(w, dtype=np.result_type(x, g))
    dxp = np.zeros(xp.shape, dtype=np.result_type(w, g))
    for dy in range(3):
        for dx in range(3):
            patch = xp[dy:dy + H, dx:dx + W, :]
            dw[dy, dx] = patch.reshape(-1, patch.shape[2]).T @ g.reshape(-1, g.shape[2])
            dxp[dy:dy + H, dx:dx + W, :] += g @ w[dy, dx].T
    return dxp[1:H + 1, 1:W + 1, :], dw, g.sum(axis=(0, 1))


def pool(x):
    H, W, C = x.shape
    return x.reshape(H // 2, 2, W // 2, 2, C).mean(axis=(1, 3))


def pool_backward(g):
    return np.repeat(np.repeat(g, 2, axis=0), 2, axis=1) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up_backward(g):
    H, W, C = g.shape
    return g.reshape(H // 2, 2, W // 2, 2, C).sum(axis=(1, 3))


def relu(x):
    return np.maximum(x, 0.0)


def forward(p, x):
    """x: (H, W, 3) with H, W multiples of 4. Returns (out (H, W), cache)."""
    c = {"x": x}

    def layer(name, inp):
        c[name + ".in"] = inp
        y = relu(conv3(inp, p[name + ".w"], p[name + ".b"]))
        c[name] = y
        return y

    e1 = layer("e1", layer("e1a", x))
    e2 = layer("e2", layer("e2a", pool(e1)))
    e3 = layer("e3", layer("e3a", pool(e2)))
    d2 = layer("d2", layer("d2a", np.concatenate([up(e3), e2], axis=2)))
    d1 = layer("d1", layer("d1a", np.concatenate([up(d2), e1], axis=2)))
    out = d1 @ p["out.w"] + p["out.b"][0]
    return out, c


def backward(p, c, gout):
    """Gradients of every parameter given gout = dL/dout (H, W); also dL/dx."""
    g = {}
    d1 = c["d1"]
    g["out.w"] = np.einsum("hwc,hw->c", d1, gout)
    g["out.b"] = np.array([gout.sum()])
    gd1 = gout[:, :, None] * p["out.w"][None, None, :]

    def layer_back(name, gy):
        ga = gy * (c[name] > 0)
        gx, gw, gb = conv3_backward(c[name + ".in"], p[name + ".w"], ga)
        g[name + ".w"], g[name + ".b"] = gw, gb
        return gx

    gcat1 = layer_back("d1a", layer_back("d1", gd1))
    ge1 = gcat1[:, :, 32:]
    gd2 = up_backward(gcat1[:, :, :32])
    gcat2 = layer_back("d2a", layer_back("d2", gd2))
    ge2 = gcat2[:, :, 64:]
    ge3 = up_backward(gcat2[:, :, :64])
    gp2 = layer_back("e3a", layer_back("e3", ge3))
    ge2 = ge2 + pool_backward(gp2)
    gp1 = layer_back("e2a", layer_back("e2", ge2))
    ge1 = ge1 + pool_backward(gp1)
    gx = layer_back("e1a", layer_back("e1", ge1))
    return g, gx


def features(own, kx, ky, r):
    """The network's three input channels for one system (own: diagonal term c_i; kx, ky: face couplings; r: the
    residual), each cell's scale removed so the network sees dimensionless inputs:
      0: the residual scaled by the inverse square root of the diagonal d_i = own_i + sum of its couplings, divided
         by that scaled residual's RMS (s);
      1: log D: the log of the cell's mean face coupling, minus its grid mean;
      2: the diagonal term's share, log(own_i / d_i).
    Returns (x (H, W, 3), s, dinv_sqrt (H, W))."""
    h, w = own.shape
    ksum = kx[:, :w] + kx[:, 1:] + ky[:h, :] + ky[1:, :]
    d = own + ksum
    dis = 1.0 / np.sqrt(d)
    rt = r * dis
    s = np.sqrt(np.mean(rt * rt))
    s = s if s > 0 else 1.0
    logk = np.log(np.maximum(0.25 * ksum, 1e-30))
    x = np.stack([rt / s, logk - logk.mean(), np.log(own / d)], axis=2)
    return x, s, dis


def apply(p, own, kx, ky, r):
    """The preconditioner: z = D^-1/2 s N(x)."""
    x, s, dis = features(own, kx, ky, r)
    out, _ = forward(p, x)
    return out * s * dis


def adam_step(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba 2015, Algorithm 1) of every parameter, in place; t counts from 1."""
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        p[k] -= lr * mh / (np.sqrt(vh) + eps)


def matvec(own, kx, ky, v):
    """A v for the face-form operator: own_i v_i + sum_f k_f (v_i - v_j) (couplings on the grid's rim zero)."""
    h, w = own.shape
    out = own * v
    out[:, 1:] += kx[:, 1:w] * (v[:, 1:] - v[:, :-1])
    out[:, :-1] += kx[:, 1:w] * (v[:, :-1] - v[:, 1:])
    out[1:, :] += ky[1:h, :] * (v[1:, :] - v[:-1, :])
    out[:-1, :] += ky[1:h, :] * (v[:-1, :] - v[1:, :])
    return out


def energy_loss(p, own, kx, ky, r, e):
    """L = (z - e)^T A (z - e) / e^T A e with z the preconditioner applied to r and e = A^-1 r, and dL/dparams."""
    x, s, dis = features(own, kx, ky, r)
    out, cache = forward(p, x)
    z = out * s * dis
    err = z - e
    aerr = matvec(own, kx, ky, err)
    norm = float(np.sum(e * matvec(own, kx, ky, e)))
    loss = float(np.sum(err * aerr)) / norm
    gout = 2.0 * aerr / norm * s * dis
    g, _ = backward(p, cache, gout)
    return loss, g
