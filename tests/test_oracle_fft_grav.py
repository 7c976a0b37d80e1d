"""The oracle's FFT, convolution and gravity pinned by the reference's own closed-form checks.

Mirrors crates/sr-physics/src/reference/fft.rs:340-548 (tolerances as there) and
crates/sr-engine/tests/gpu/grav.rs:40-437 (kernel values, Maclaurin disk force, self-force).
The direct DFT and direct convolution sums below are independent brute-force computations.
"""
import math

import numpy as np
import pytest

import oracle


def pseudo_random(count, seed):
    """fft.rs:292-300: the LCG mapped to [-1, 1)."""
    return oracle.lcg(seed, count) * 2.0 - 1.0


def pseudo_complex(count, seed):
    v = pseudo_random(2 * count, seed)
    return v[0::2] + 1j * v[1::2]


def direct_dft(x, inverse):
    n = len(x)
    k = np.arange(n)
    sign = 1.0 if inverse else -1.0
    ang = sign * 2.0 * np.pi * ((np.outer(k, k) % n) / n)
    out = (np.cos(ang) + 1j * np.sin(ang)) @ x
    return out / n if inverse else out


def rel_err(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_lcg_matches_the_reference_generator():
    # the recurrence by Python's own integers, independent of the C code
    s, out = 7, []
    for _ in range(5):
        s = (s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        out.append((s >> 11) / float(1 << 53))
    assert np.array_equal(oracle.lcg(7, 5), np.array(out))


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128])
def test_fft_matches_the_direct_dft(n):
    x = pseudo_complex(n, 7 + n)
    for inverse in (False, True):
        assert rel_err(oracle.fft(x, inverse), direct_dft(x, inverse)) < 1e-12


def test_fft_of_a_delta_and_a_wave_are_exact():
    d = np.zeros(16, complex)
    d[0] = 1
    y = oracle.fft(d)
    assert np.all(np.abs(y.real - 1) < 1e-15) and np.all(np.abs(y.imag) < 1e-15)
    n = np.arange(16)
    wave = np.exp(2j * np.pi * 3 * n / 16)
    y = oracle.fft(wave)
    want = np.where(n == 3, 16.0, 0.0)
    assert np.all(np.abs(y.real - want) < 1e-12) and np.all(np.abs(y.imag) < 1e-12)


def test_inverse_undoes_forward_in_1d_and_2d():
    for n in (32, 256, 2048):
        x = pseudo_complex(n, 3)
        assert rel_err(oracle.fft(oracle.fft(x), True), x) < 1e-13
    x = pseudo_complex(64 * 32, 5).reshape(32, 64)
    assert rel_err(oracle.fft2(oracle.fft2(x), True), x) < 1e-13


def test_fft2_matches_the_direct_2d_dft_and_obeys_parseval():
    nx, ny = 16, 8
    x = pseudo_complex(nx * ny, 11).reshape(ny, nx)
    y = oracle.fft2(x)
    direct = np.zeros((ny, nx), complex)
    for ky in range(ny):
        for kx in range(nx):
            jy, jx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
            phase = ((kx * jx) % nx) / nx + ((ky * jy) % ny) / ny
            direct[ky, kx] = np.sum(x * np.exp(-2j * np.pi * phase))
    assert rel_err(y, direct) < 1e-12
    space = np.sum(np.abs(x) ** 2)
    freq = np.sum(np.abs(y) ** 2) / (nx * ny)
    assert abs((space - freq) / space) < 1e-13


def test_parseval_holds_at_the_largest_size():
    x = pseudo_complex(2048, 17)
    y = oracle.fft(x)
    space, freq = np.sum(np.abs(x) ** 2), np.sum(np.abs(y) ** 2) / 2048
    assert abs((space - freq) / space) < 1e-13


def asymmetric_kernel(dx, dy):
    x, y = float(dx), float(dy)
    return (1.0 + 0.5 * math.tanh(0.2 * x) + 0.3 * math.tanh(0.1 * y)) / math.sqrt(1.0 + x * x + 2.0 * y * y)


def direct_convolution(field, kern):
    h, w = field.shape
    ty, tx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    out = np.zeros((h, w))
    for sy in range(h):
        for sx in range(w):
            if field[sy, sx] != 0.0:
                out += field[sy, sx] * kern(tx - sx, ty - sy)
    return out


def asym_np(dx, dy):
    x, y = dx.astype(float), dy.astype(float)
    return (1.0 + 0.5 * np.tanh(0.2 * x) + 0.3 * np.tanh(0.1 * y)) / np.sqrt(1.0 + x * x + 2.0 * y * y)


def fft_convolve(field, kern_np, ring=0, max_padded=2048):
    """The oracle's convolution machinery with an arbitrary kernel, through fft2 (fft.rs:238-282)."""
    h, w = field.shape
    px, py = oracle.padded_size(w, max_padded), oracle.padded_size(h, max_padded)
    ix = np.where(np.arange(px) < px // 2, np.arange(px), np.arange(px) - px)
    iy = np.where(np.arange(py) < py // 2, np.arange(py), np.arange(py) - py)
    k = kern_np(ix[None, :], iy[:, None]) + 0j
    kh = oracle.fft2(k)
    p = np.zeros((py, px), complex)
    p[:h, :w] = field
    out = oracle.fft2(oracle.fft2(p) * kh, True).real
    ys = (np.arange(h + 2 * ring) + py - ring) % py
    xs = (np.arange(w + 2 * ring) + px - ring) % px
    return out[np.ix_(ys, xs)]


def test_convolution_matches_direct_summation_on_48_by_40_with_an_asymmetric_kernel():
    w, h = 48, 40
    assert asymmetric_kernel(5, 3) != asymmetric_kernel(-5, 3)
    field = (pseudo_random(w * h, 29) + 1.0).reshape(h, w)
    fast = fft_convolve(field, asym_np)
    direct = direct_convolution(field, asym_np)
    assert rel_err(fast, direct) < 1e-10
    rim = np.concatenate([(fast[0] - direct[0]) / direct[0], (fast[-1] - direct[-1]) / direct[-1],
                          (fast[:, 0] - direct[:, 0]) / direct[:, 0], (fast[:, -1] - direct[:, -1]) / direct[:, -1]])
    assert np.max(np.abs(rim)) < 1e-10


def test_mass_in_the_corners_does_not_wrap_around():
    w, h = 40, 33
    field = np.zeros((h, w))
    for x, y in ((0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1)):
        field[y, x] = 1.0
    assert rel_err(fft_convolve(field, asym_np), direct_convolution(field, asym_np)) < 1e-10


def test_a_single_source_reproduces_the_kernel_shifted():
    w, h, sx, sy = 20, 17, 6, 11
    field = np.zeros((h, w))
    field[sy, sx] = 2.5
    out = fft_convolve(field, asym_np)
    yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    want = 2.5 * asym_np(xx - sx, yy - sy)
    assert np.all(np.abs(out - want) < 1e-12 * np.maximum(np.abs(want), 1.0))


def test_padded_sizes_follow_the_box_rule_and_refuse_out_of_range_boxes():
    for n, want in ((1, 32), (16, 32), (17, 64), (32, 64), (48, 128), (64, 128), (300, 1024), (1024, 2048)):
        assert oracle.padded_size(n) == want
    with pytest.raises(oracle.OracleError):
        oracle.padded_size(0)
    with pytest.raises(oracle.OracleError):
        oracle.padded_size(1025)
    assert oracle.padded_size(2048, 4096) == 4096


# ---------------------------------------------------------------- gravity (grav.rs, tests/gpu/grav.rs)

SELF_CELL = 4.0 * math.log(1.0 + math.sqrt(2.0))


def square_potential_by_quadrature(n=2000):
    """int over [-1/2, 1/2]^2 of dA / r by Simpson in polar form (tests/gpu/grav.rs:28-38)."""
    a, b = 0.0, math.pi / 4
    f = lambda t: 0.5 / math.cos(t)
    h = (b - a) / n
    s = f(a) + f(b) + sum(f(a + k * h) * (4 if k % 2 else 2) for k in range(1, n))
    return 8.0 * s * h / 3.0


def test_kernel_is_the_sheet_kernel():
    assert abs(SELF_CELL - 3.525494348078172) < 1e-15
    assert abs(SELF_CELL - square_potential_by_quadrature()) < 1e-12
    assert oracle.grav_kernel(0, 0) == -3.525494348078172
    assert oracle.grav_kernel(0, 0, G=2.5) == -2.5 * 3.525494348078172
    for i, j, r in ((1, 0, 1.0), (0, 1, 1.0), (3, 4, 5.0), (-5, 12, 13.0), (8, -15, 17.0), (-7, -24, 25.0)):
        assert abs(oracle.grav_kernel(i, j) + 1.0 / r) < 1e-15
        assert oracle.grav_kernel(i, j) == oracle.grav_kernel(-i, j) == oracle.grav_kernel(i, -j) == oracle.grav_kernel(j, i)
    # a cell of side dx carries mass sigma dx^2 at distance r dx: K scales by dx
    assert abs(oracle.grav_kernel(3, 4, dx=0.25) + 0.25 / 5.0) < 1e-16


def test_one_unit_source_gives_the_self_cell_and_minus_g_over_r():
    ph = oracle.reference_physics(sigma_vac=0.5)
    s = np.zeros((33, 33))
    s[16, 16] = 1.0
    phi, gx, gy = oracle.gravity_field(ph, s)
    assert abs(phi[16, 16] + SELF_CELL) < 1e-12
    assert abs(phi[16, 19] + 1 / 3) < 1e-12 and abs(phi[20, 19] + 1 / 5) < 1e-12
    # vacuum is no source: the same cell under sigma_vac pulls nothing
    s[16, 16] = 0.4
    phi, _, _ = oracle.gravity_field(ph, s)
    assert np.all(phi == 0.0)


def maclaurin(a, centre, side):
    cx, cy = centre
    j, i = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    r2 = ((i - cx) ** 2 + (j - cy) ** 2) / (a * a)
    return np.where(r2 < 1.0, np.sqrt(np.clip(1.0 - r2, 0, None)), 0.0)


def disk_error(a, gx, gy, origin):
    k = math.pi ** 2 / (2.0 * a)
    h, w = gx.shape
    j, i = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    dx, dy = i - origin[0], j - origin[1]
    r = np.hypot(dx, dy)
    m = (r > 0) & (r <= 0.75 * a)
    err = np.sqrt((gx + k * dx) ** 2 + (gy + k * dy) ** 2)[m] / (k * r[m])
    return err.max()


def worst_error(a, offset, ph):
    side = int(math.ceil(2 * a)) + 17
    o = (side - 1) / 2.0
    sigma = maclaurin(a, (o + offset[0], o + offset[1]), side)
    _, gx, gy = oracle.gravity_field(ph, sigma)
    return disk_error(a, gx, gy, (o + offset[0], o + offset[1]))


def test_maclaurin_disk_force_is_within_two_percent_and_first_order():
    ph = oracle.reference_physics(sigma_vac=0.0)
    e40 = worst_error(40.0, (0.0, 0.0), ph)
    e20 = worst_error(20.0, (0.0, 0.0), ph)
    assert e40 <= 0.02
    assert 1.6 <= e20 / e40 <= 2.4
    assert worst_error(40.0, (0.5, 0.5), ph) <= 0.02


def test_net_self_force_vanishes():
    ph = oracle.reference_physics(sigma_vac=1e-6)
    rng = np.random.default_rng(3)
    sigma = rng.uniform(0.0, 1.0, (40, 52))
    sigma[:, :3] = 0.0
    _, gx, gy = oracle.gravity_field(ph, sigma)
    gross = np.sum(sigma * np.hypot(gx, gy))
    assert abs(np.sum(sigma * gx)) <= 1e-12 * gross and abs(np.sum(sigma * gy)) <= 1e-12 * gross


def test_gravity_matches_a_direct_sum_with_the_ring():
    ph = oracle.reference_physics(sigma_vac=1e-6, G=1.3, dx=0.5)
    rng = np.random.default_rng(5)
    sigma = rng.uniform(0.0, 1.0, (13, 19))
    phi, gx, gy, pr = oracle.gravity_field(ph, sigma, ringed=True)
    h, w = sigma.shape
    kern = lambda dx, dy: np.where((dx == 0) & (dy == 0), -1.3 * 0.5 * SELF_CELL,
                                   -1.3 * 0.5 / np.sqrt(np.maximum(dx * dx + dy * dy, 1)))
    big = np.zeros((h + 2, w + 2))
    big[1:-1, 1:-1] = sigma
    direct = direct_convolution(big, kern)
    assert rel_err(pr, direct) < 1e-12
    assert np.allclose(gx, -(pr[1:-1, 2:] - pr[1:-1, :-2]) / (2 * 0.5), rtol=0, atol=1e-12)
    assert np.allclose(gy, -(pr[2:, 1:-1] - pr[:-2, 1:-1]) / (2 * 0.5), rtol=0, atol=1e-12)
