"""The network oracle (oracle/unet.py) pinned by central finite differences: every layer's backward, and the whole
network's parameter gradients through the energy-norm loss. A transposed weight, a swapped concat half or a missing
ReLU mask fails these."""
import numpy as np
import pytest

from oracle import unet


def fd(f, a, idx, h=1e-6):
    old = a[idx]
    a[idx] = old + h
    fp = f()
    a[idx] = old - h
    fm = f()
    a[idx] = old
    return (fp - fm) / (2 * h)


def test_conv3_against_direct_sum():
    rng = np.random.default_rng(0)
    x, w, b = rng.normal(size=(5, 6, 3)), rng.normal(size=(3, 3, 3, 4)), rng.normal(size=4)
    out = unet.conv3(x, w, b)
    y, xx, o = 0, 5, 2  # a corner-adjacent cell: the zero padding matters
    s = b[o]
    for dy in range(3):
        for dx in range(3):
            yy, xs = y + dy - 1, xx + dx - 1
            if 0 <= yy < 5 and 0 <= xs < 6:
                s += w[dy, dx, :, o] @ x[yy, xs]
    assert out[y, xx, o] == pytest.approx(s, rel=1e-13)


def test_conv3_backward_by_finite_differences():
    rng = np.random.default_rng(1)
    x, w, b = rng.normal(size=(6, 5, 3)), rng.normal(size=(3, 3, 3, 4)), rng.normal(size=4)
    gy = rng.normal(size=(6, 5, 4))
    f = lambda: float(np.sum(unet.conv3(x, w, b) * gy))
    gx, gw, gb = unet.conv3_backward(x, w, gy)
    for idx in [(0, 0, 0), (3, 4, 2), (5, 2, 1)]:
        assert gx[idx] == pytest.approx(fd(f, x, idx), rel=1e-7)
    for idx in [(0, 0, 0, 0), (2, 1, 2, 3), (1, 2, 0, 1)]:
        assert gw[idx] == pytest.approx(fd(f, w, idx), rel=1e-7)
    assert gb[2] == pytest.approx(fd(f, b, (2,)), rel=1e-7)


def test_pool_and_up_are_adjoint_pairs():
    rng = np.random.default_rng(2)
    a, c = rng.normal(size=(8, 6, 3)), rng.normal(size=(4, 3, 3))
    assert np.sum(unet.pool(a) * c) == pytest.approx(np.sum(a * unet.pool_backward(c)), rel=1e-13)
    assert np.sum(unet.up(c) * a) == pytest.approx(np.sum(c * unet.up_backward(a)), rel=1e-13)


def small_system(n=8, seed=3):
    rng = np.random.default_rng(seed)
    own = np.exp(rng.normal(size=(n, n)))
    kx = np.exp(rng.normal(size=(n, n + 1)))
    ky = np.exp(rng.normal(size=(n + 1, n)))
    kx[:, 0] = kx[:, n] = 0.0
    ky[0, :] = ky[n, :] = 0.0
    return own, kx, ky


def test_matvec_agrees_with_the_c_oracle():
    import oracle
    own, kx, ky = small_system(12, 4)
    v = np.random.default_rng(5).normal(size=(12, 12))
    assert np.allclose(unet.matvec(own, kx, ky, v), oracle.fld_matvec(own, kx, ky, None, v), rtol=1e-13)


def test_network_gradients_by_finite_differences():
    own, kx, ky = small_system(8, 6)
    rng = np.random.default_rng(7)
    r = rng.normal(size=(8, 8))
    e = np.linalg.solve(dense(own, kx, ky), r.ravel()).reshape(8, 8)
    p = unet.init_params(8)
    for k in p:  # nonzero biases so that every ReLU mask is exercised both ways
        if k.endswith(".b"):
            p[k] = rng.normal(size=p[k].shape) * 0.1
    p["out.w"] = rng.normal(size=16) * 0.3
    loss, g = unet.energy_loss(p, own, kx, ky, r, e)
    f = lambda: unet.energy_loss(p, own, kx, ky, r, e)[0]
    for name, idx in [("e1a.w", (1, 1, 0, 3)), ("e1a.w", (0, 2, 2, 5)), ("e1.b", (4,)), ("e2a.w", (2, 0, 7, 11)),
                      ("e3.w", (1, 1, 10, 20)), ("e3a.b", (3,)), ("d2a.w", (1, 2, 5, 1)), ("d2a.w", (0, 1, 70, 9)),
                      ("d1a.w", (2, 2, 3, 0)), ("d1a.w", (1, 0, 40, 7)), ("d1.w", (1, 1, 2, 2)), ("out.w", (5,)),
                      ("out.b", (0,))]:
        num = fd(f, p[name], idx)
        assert g[name][idx] == pytest.approx(num, rel=1e-5, abs=1e-9), name


def dense(own, kx, ky):
    n = own.size
    A = np.zeros((n, n))
    for i in range(n):
        v = np.zeros(n)
        v[i] = 1.0
        A[:, i] = unet.matvec(own, kx, ky, v.reshape(own.shape)).ravel()
    return A


def test_adam_step_matches_the_algorithm():
    p = {"a": np.array([1.0, -2.0])}
    g = {"a": np.array([0.5, 0.25])}
    m = {"a": np.zeros(2)}
    v = {"a": np.zeros(2)}
    unet.adam_step(p, g, m, v, 1, lr=0.1)
    # the first step of Adam moves each parameter by lr sign(g) (bias correction makes m/sqrt(v) = g/|g|)
    assert np.allclose(p["a"], [0.9, -2.1], atol=1e-6)
