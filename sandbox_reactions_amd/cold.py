"""The cold pressure of the equation of state on the host (crates/sr-physics/src/eos.rs:24-140), for laying down a
scene's cells: P(x) = 1 / (1 / (K1 x^2) + 1 / (K2 x^(3/2))) with x = Y_e Sigma (electrons) or X_n Sigma (neutron
matter), and u(x) = x int_0^x P(s) / s^2 ds tabulated as ln u on 512 points log-spaced over [1e-8, 1e8] (Simpson's
rule in t = sqrt(s)), read by linear interpolation in (ln x, ln u) and the K1 x^2 power law below the table; the
device builds and reads the same table. A pair with a K of 0 is off."""
from __future__ import annotations

import numpy as np

POINTS = 512
X_MIN, X_MAX = 1e-8, 1e8
SHIPPED = {"k1e": 21.3, "k2e": 134.0, "k1n": 2.67, "k2n": 58.2}  # assets/physics.json
# Z / A per species in the planes' order H He C O Ne Mg Si S Fe n (assets/elements.json)
Z_OVER_A = np.array([1.0, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 26.0 / 56.0, 0.0])

_tables: dict = {}


def pressure(x, k1: float, k2: float):
    x = np.asarray(x, dtype=np.float64)
    if not (k1 > 0.0 and k2 > 0.0):
        return np.zeros_like(x)
    xs = np.where(x > 0.0, x, 1.0)
    return np.where(x > 0.0, 1.0 / (1.0 / (k1 * xs * xs) + 1.0 / (k2 * xs ** 1.5)), 0.0)


def nodes() -> np.ndarray:
    return np.exp(np.log(X_MIN) + np.arange(POINTS) / (POINTS - 1) * (np.log(X_MAX) - np.log(X_MIN)))


def table(k1: float, k2: float) -> np.ndarray:
    """ln u at the nodes (built once per pair)"""
    key = (float(k1), float(k2))
    if key not in _tables:
        x = nodes()
        t = np.sqrt(x)
        t0 = np.concatenate([[0.0], t[:-1]])
        ln_u = np.empty(POINTS)
        j = 0.0
        for i in range(POINTS):
            n = 512 if i == 0 else 64
            tt = np.linspace(t0[i], t[i], n + 1)
            s = tt * tt
            f = np.where(tt > 0.0, pressure(np.where(tt > 0, s, 1.0), k1, k2) / np.where(tt > 0, s * s, 1.0) * 2.0 * tt,
                         0.0)
            wgt = np.ones(n + 1)
            wgt[1:-1:2] = 4.0
            wgt[2:-1:2] = 2.0
            j += float(np.dot(wgt, f)) * (tt[1] - tt[0]) / 3.0
            ln_u[i] = np.log(x[i] * j)
        _tables[key] = ln_u
    return _tables[key]


def energy(x, k1: float, k2: float):
    """u(x), the cold energy per area of the pair at x"""
    x = np.asarray(x, dtype=np.float64)
    if not (k1 > 0.0 and k2 > 0.0):
        return np.zeros_like(x)
    ln_u = table(k1, k2)
    xs = np.where(x > 0.0, x, 1.0)
    step = (np.log(X_MAX) - np.log(X_MIN)) / (POINTS - 1)
    pos = (np.log(xs) - np.log(X_MIN)) / step
    i = np.clip(np.floor(pos).astype(np.int64), 0, POINTS - 2)
    frac = pos - i
    inside = np.exp(ln_u[i] + frac * (ln_u[i + 1] - ln_u[i]))
    below = np.exp(ln_u[0]) * (xs / X_MIN) ** 2
    return np.where(x > 0.0, np.where(pos < 0.0, below, inside), 0.0)


def y_e(fractions) -> np.ndarray:
    """Y_e of mass fractions (last axis the 10 species)"""
    return np.asarray(fractions, dtype=np.float64) @ Z_OVER_A


def pi_cold(c: dict, sigma, ye, xn):
    """Pi_e + Pi_n of cells of Sigma `sigma`, Y_e `ye` and X_n `xn` under the constants `c` (k1e, k2e, k1n, k2n)"""
    sigma = np.asarray(sigma, dtype=np.float64)
    return pressure(ye * sigma, c.get("k1e", 0.0), c.get("k2e", 0.0)) + pressure(xn * sigma, c.get("k1n", 0.0),
                                                                                     c.get("k2n", 0.0))


def eps_cold(c: dict, sigma, ye, xn):
    """(u_e + u_n) / Sigma"""
    sigma = np.asarray(sigma, dtype=np.float64)
    u = energy(ye * sigma, c.get("k1e", 0.0), c.get("k2e", 0.0)) + energy(xn * sigma, c.get("k1n", 0.0),
                                                                          c.get("k2n", 0.0))
    return u / sigma
