"""The oracle's cold pressure (oracle/eos.c restating crates/sr-physics/src/eos.rs) pinned by what the mathematics
fixes, with the bars of the reference's own EOS tests (eos.rs:287-440): the blend and its two limits, the table's
nodes against the integral in closed form (itself checked against an independent quadrature, scipy's), the
interpolation between nodes, monotonicity, the low end; then the cold terms inside the step: the floors' vacuum
reset carries the cold energy, and a hydro step with cold pressure conserves mass and energy to round-off."""
import numpy as np
import pytest
from scipy import integrate

import oracle

SHIPPED = [(21.3, 134.0), (2.67, 58.2)]


def rel(a, b):
    return abs((a - b) / b)


def test_blend_and_limits():
    for k1, k2 in SHIPPED:
        assert rel(oracle.cold_pressure(1e-12, k1, k2), k1 * 1e-24) < 1e-6
        assert rel(oracle.cold_pressure(1e16, k1, k2), k2 * 1e16 ** 1.5) < 1e-6
        for x in (1e-6, 1e-2, 1.0, 30.0, 1e4):
            assert rel(oracle.cold_pressure(x, k1, k2), k1 * x * x / (1.0 + k1 * np.sqrt(x) / k2)) < 1e-12
        assert oracle.cold_pressure(0.0, k1, k2) == 0.0 and oracle.cold_pressure(-1.0, k1, k2) == 0.0
    assert oracle.cold_pressure(1.0, 0.0, 134.0) == 0.0


@pytest.mark.parametrize("k1,k2", [(21.3, 134.0), (2.67, 58.2), (7.0, 7.0)])
def test_closed_form_is_the_integral(k1, k2):
    """u(x) = x int_0^x P(s)/s^2 ds by scipy's adaptive quadrature in t = sqrt(s), independent of the table's Simpson."""
    for x in (1e-7, 1e-3, 0.5, 20.0, 5e3, 3e7):
        f = lambda t: oracle.cold_pressure(t * t, k1, k2) / t ** 4 * 2.0 * t if t > 0 else 0.0
        j, _ = integrate.quad(f, 0.0, np.sqrt(x), epsabs=0.0, epsrel=1e-12, limit=200)
        assert rel(x * j, oracle.cold_energy_exact(x, k1, k2)) < 1e-9, x


@pytest.mark.parametrize("k1,k2", SHIPPED + [(1.0, 100.0), (100.0, 1.0), (7.0, 7.0)])
def test_table_nodes_equal_the_closed_form(k1, k2):
    t = oracle.cold_table(k1, k2)
    worst = max(rel(np.exp(t[i]), oracle.cold_energy_exact(oracle.cold_node_x(i), k1, k2)) for i in range(512))
    assert worst < 1e-9


@pytest.mark.parametrize("k1,k2", SHIPPED + [(1.0, 1e-2), (1.0, 1.0), (1.0, 1e2), (1e3, 1.0), (1.0, 1e3)])
def test_interpolation_between_nodes(k1, k2):
    t = oracle.cold_table(k1, k2)
    step = (np.log(1e8) - np.log(1e-8)) / 511
    worst = 0.0
    for i in range(511):
        for frac in (1 / 3, 0.5, 2 / 3):
            x = np.exp(np.log(oracle.cold_node_x(i)) + frac * step)
            worst = max(worst, rel(oracle.cold_table_u(t, x), oracle.cold_energy_exact(x, k1, k2)))
    assert worst <= 1e-4


def test_geometry_monotone_and_low_end():
    assert rel(oracle.cold_node_x(0), 1e-8) < 1e-12 and rel(oracle.cold_node_x(511), 1e8) < 1e-12
    for k1, k2 in SHIPPED:
        t = oracle.cold_table(k1, k2)
        assert np.all(np.diff(t) > 0)
        xs = 1e-8 * (1e16) ** (np.arange(4000) / 3999)
        u = np.array([oracle.cold_table_u(t, x) for x in xs])
        assert np.all(np.diff(u) > 0) and np.all(np.isfinite(u))
        assert oracle.cold_table_u(t, 0.0) == 0.0 and oracle.cold_table_u(t, -1.0) == 0.0
        # below the table: the K1 x^2 power law from the first node
        assert rel(oracle.cold_table_u(t, 1e-10), np.exp(t[0]) * 1e-4) < 1e-12
        assert rel(oracle.cold_table_u(t, 1e-8), k1 * 1e-16) < 1e-4


def test_cold_terms_of_a_physics():
    ph = oracle.reference_physics(**oracle.SHIPPED_COLD)
    s, ye, xn = 0.7, 0.8673, 0.05
    assert rel(oracle.pi_cold(ph, s, ye, xn),
               oracle.cold_pressure(ye * s, 21.3, 134.0) + oracle.cold_pressure(xn * s, 2.67, 58.2)) < 1e-15
    want = (oracle.cold_energy_exact(ye * s, 21.3, 134.0) + oracle.cold_energy_exact(xn * s, 2.67, 58.2)) / s
    assert rel(oracle.eps_cold(ph, s, ye, xn), want) < 1e-4
    off = oracle.reference_physics()
    assert oracle.pi_cold(off, s, ye, xn) == 0.0 and oracle.eps_cold(off, s, ye, xn) == 0.0


def test_vacuum_reset_carries_the_cold_energy():
    """floors.rs:128-139: a vacuum cell becomes Sigma_floor at rest with E = Sigma_floor (T_floor / mu + eps_cold)."""
    ph = oracle.reference_physics(**oracle.SHIPPED_COLD)
    p = np.zeros((14, 8, 8))
    p[0] = 1.0
    p[3] = 3.0
    p[4] = 1.0
    p[0][3, 4] = 1e-9  # below Sigma_vac
    out = oracle.floors_apply(ph, p)[0]
    sf = float(np.float32(ph.sigma_floor))
    want = sf * (ph.t_floor * 2.0 + oracle.eps_cold(ph, sf, 1.0, 0.0))  # hydrogen: 1/mu = 2, Y_e = 1
    assert rel(out[3][3, 4], want) < 1e-12 and out[0][3, 4] == sf


def test_hydro_step_with_cold_pressure_conserves():
    """Bounce walls, no sources: mass and energy to round-off over steps of a blob in cold-pressure gas."""
    ph = oracle.reference_physics(**oracle.SHIPPED_COLD, edge=oracle.EDGE_BOUNCE)
    n = 32
    yy, xx = np.mgrid[0:n, 0:n]
    rho = 0.05 + 0.9 * np.exp(-((xx - 14.5) ** 2 + (yy - 17.5) ** 2) / 18.0)
    p = np.zeros((14, n, n))
    p[0] = rho
    p[4] = 0.7
    p[5] = 0.2
    p[13] = 0.1  # neutrons: both pairs on
    ye = 0.7 + 0.2 * 0.5
    eps_c = np.array([[oracle.eps_cold(ph, r, ye, 0.1) for r in row] for row in rho])
    p[3] = rho * (0.3 + eps_c)
    m0, e0 = p[0].sum(), p[3].sum()
    for _ in range(5):
        dt = oracle.hydro_dt(ph, p)
        p = oracle.hydro_step(ph, p, dt)
    assert rel(p[0].sum(), m0) < 1e-13 and rel(p[3].sum(), e0) < 1e-12
    assert np.all(np.isfinite(p))
