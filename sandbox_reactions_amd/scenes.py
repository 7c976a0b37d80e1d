"""Seeded synthetic scenes: the benchmark workloads' fields and the reference's own parity scenes as planes.

Plane layout: 14 f32 planes (Sigma, Sigma u, Sigma v, E, then 10 species fractions in the reference's order H He C O Ne
Mg Si S Fe n), row-major, y down (crates/sr-engine/src/state.rs:26-46). Random numbers come from the reference's
64-bit LCG (crates/sr-physics/src/reference/fft.rs:292-300), Gaussian deviates by Box-Muller on it, fields
from numpy's FFT (input generation, not the measured path).
"""
from __future__ import annotations

import numpy as np

NS = 10
SEED_BASE = 0x5EED0000


def lcg_uniform(seed: int, count: int) -> np.ndarray:
    """s <- s * 6364136223846793005 + 1442695040888963407 (mod 2^64), u = (s >> 11) / 2^53 — vectorised by
    jumping: the k-th state is A^k s + C_k, computed with Python integers in blocks."""
    a, c, m = 6364136223846793005, 1442695040888963407, 1 << 64
    out = np.empty(count, dtype=np.float64)
    block = min(count, 1 << 16) if count else 1
    # multipliers for one block: s_k = A_k s + C_k
    ak = np.empty(block, dtype=np.uint64)
    ck = np.empty(block, dtype=np.uint64)
    A, Cc = 1, 0
    for k in range(block):
        A = (A * a) % m
        Cc = (Cc * a + c) % m
        ak[k] = A
        ck[k] = Cc
    s = seed % m
    done = 0
    with np.errstate(over="ignore"):
        while done < count:
            nb = min(block, count - done)
            st = ak[:nb] * np.uint64(s) + ck[:nb]
            out[done:done + nb] = (st >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
            s = int(st[nb - 1])
            done += nb
    return out


def normal(seed: int, count: int) -> np.ndarray:
    u = lcg_uniform(seed, 2 * ((count + 1) // 2))
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:count]


def gaussian_random_field(n: int, seed: int, slope: float = -3.0) -> np.ndarray:
    """Unit-variance field with P(k) ~ k^slope for k >= 1, the k = 0 mode zeroed."""
    white = normal(seed, n * n).reshape(n, n)
    f = np.fft.fft2(white)
    k = np.sqrt(np.fft.fftfreq(n)[:, None] ** 2 + np.fft.fftfreq(n)[None, :] ** 2) * n
    amp = np.zeros_like(k)
    nz = k >= 1.0
    amp[nz] = k[nz] ** (slope / 2.0)
    g = np.fft.ifft2(f * amp).real
    g -= g.mean()
    return g / g.std()


def lognormal(n: int, sigma_ln: float, seed: int) -> np.ndarray:
    g = gaussian_random_field(n, seed)
    return np.exp(sigma_ln * g - 0.5 * sigma_ln * sigma_ln)


def planes_from(rho, T, u=None, v=None, gamma=5.0 / 3.0, mu=1.0, x_h=1.0) -> np.ndarray:
    """The 14 planes of an ideal gas: E = rho T / (mu (gamma - 1)) + rho |u|^2 / 2, pure hydrogen by default."""
    h, w = rho.shape
    u = np.zeros_like(rho) if u is None else u
    v = np.zeros_like(rho) if v is None else v
    p = np.zeros((14, h, w), dtype=np.float64)
    p[0] = rho
    p[1] = rho * u
    p[2] = rho * v
    p[3] = rho * T / (mu * (gamma - 1.0)) + 0.5 * rho * (u * u + v * v)
    p[4] = x_h
    if x_h != 1.0:
        p[5] = 1.0 - x_h
    return p


def config_physics_kwargs(n: int) -> dict:
    """Box L = 1 (dx = 1/n), ideal gas gamma = 5/3, mu = 1, G = 1, Kramers opacity kappa0 = 1, leave edges."""
    return dict(gamma=5.0 / 3.0, G=1.0, dx=1.0 / n, mu=1.0, opacity=1, kappa0=1.0, edge=1, max_padded=4096)


def blob_scene(n: int = 256, seed: int = SEED_BASE) -> np.ndarray:
    """The blob workload: rho = 1e-2 + exp(-r^2 / 2 sigma^2), sigma = 0.1 L, centred; T = 1 +- 1% seeded; v = 0."""
    c = (np.arange(n) + 0.5) / n - 0.5
    r2 = c[None, :] ** 2 + c[:, None] ** 2
    rho = 1e-2 + np.exp(-r2 / (2 * 0.1 ** 2))
    T = 1.0 + 0.01 * (2.0 * lcg_uniform(seed, n * n).reshape(n, n) - 1.0)
    return planes_from(rho, T)


def turbulent_scene(n: int = 1024, gamma: float = 5.0 / 3.0, seed: int = SEED_BASE) -> np.ndarray:
    """The turbulent full-step workload: rho lognormal (sigma_ln 1.0), T lognormal (0.5), velocities from two fields at rms Mach 0.5."""
    rho = lognormal(n, 1.0, seed + 1)
    T = lognormal(n, 0.5, seed + 2)
    u = gaussian_random_field(n, seed + 3)
    v = gaussian_random_field(n, seed + 4)
    cs = np.sqrt(gamma * T)  # p = rho T, c^2 = gamma p / rho
    mach = np.sqrt(np.mean((u * u + v * v) / (cs * cs)))
    u *= 0.5 / mach
    v *= 0.5 / mach
    return planes_from(rho, T, u, v, gamma=gamma)


# mass fractions in the planes' order H He C O Ne Mg Si S Fe n
BURN_FRESH = np.array([0.75, 0.25, 0, 0, 0, 0, 0, 0, 0, 0], float)
BURN_ASHES = np.array([0.05, 0.35, 0.2, 0.2, 0.08, 0.04, 0.04, 0.02, 0.02, 0.0], float)


def burn_scene(n: int = 1024, seed: int = SEED_BASE + 0x2000):
    """The burning workload: the reference's burning field (scenes/ref_burn.json: Sigma 30, T 100, H 0.75 | He 0.25,
    the cold pressure on, gamma 2, at rest) made uneven so that every regime of P6 occurs: Sigma = 30 x lognormal
    (sigma_ln 0.5), T = 100 x lognormal (0.5), the fractions a mix w X_ashes + (1 - w) X_fresh with w from a third
    smooth field (ashes: He, C, O, Ne, Mg, Si, S, Fe). Returns (planes (14, n, n) f32, the physics keywords)."""
    from . import cold
    sigma = 30.0 * lognormal(n, 0.5, seed)
    t = 100.0 * lognormal(n, 0.5, seed + 1)
    wgt = 0.5 * (1.0 + np.tanh(gaussian_random_field(n, seed + 2)))
    x = (1.0 - wgt)[None] * BURN_FRESH[:, None, None] + wgt[None] * BURN_ASHES[:, None, None]
    inv_mu = np.einsum("s,shw->hw", np.array([2.0, 0.75, 7 / 12, 9 / 16, 11 / 20, 13 / 24, 15 / 28, 17 / 32,
                                              27 / 56, 1.0]), x)
    ye = np.einsum("s,shw->hw", cold.Z_OVER_A, x)
    gamma = 2.0
    p = np.zeros((14, n, n))
    p[0] = sigma
    p[3] = sigma * (t * inv_mu / (gamma - 1.0) + cold.eps_cold(cold.SHIPPED, sigma, ye, x[9]))
    p[4:] = x
    return p.astype(np.float32), dict(gamma=gamma, **cold.SHIPPED)


def sod_scene(w: int = 400, h: int = 8) -> np.ndarray:
    """The reference's scenes/sod.json on a w x h strip: left half Sigma 1, Pi 1; right Sigma 0.125, Pi 0.1; H;
    gamma = 2 (Pi = Sigma eps, so E = Pi)."""
    rho = np.where(np.arange(w)[None, :] < w // 2, 1.0, 0.125) * np.ones((h, 1))
    pi = np.where(np.arange(w)[None, :] < w // 2, 1.0, 0.1) * np.ones((h, 1))
    p = np.zeros((14, h, w))
    p[0] = rho
    p[3] = pi
    p[4] = 1.0
    return p


def fld_system(n: int, seed: int = SEED_BASE, sigma_rho: float = 1.5, sigma_t: float = 0.8, stiffness: float = 1e3,
               a_r: float = 6.25e-7, c_sb: float = 12800.0, gamma: float = 5.0 / 3.0):
    """The stiff FLD system alone at n x n, L = 1. rho lognormal (1.5), T lognormal (0.8), Kramers kappa
    = rho T^-3.5, Levermore-Pomraning D on the faces; dt such that median_i (sum_f k_f / c_i) = stiffness;
    b = c_i E_r with E_r = a_r T^4 (1 + 0.1 (2u - 1)). Returns (own, kx, ky, b, dt) as float64 arrays: own the
    diagonal c_i (the grid's edge faces carry the vacuum's coupling into own)."""
    dx = 1.0 / n
    rho = lognormal(n, sigma_rho, seed + 11)
    T = lognormal(n, sigma_t, seed + 12)
    kap = rho * T ** -3.5
    ks = kap * rho
    er = a_r * T ** 4

    def face(era, erb, ksa, ksb):
        fer = 0.5 * (era + erb)
        k = 0.5 * (ksa + ksb)
        R = np.abs(erb - era) / (dx * k * fer)
        lam = (2.0 + R) / (6.0 + 3.0 * R + R * R)
        return c_sb * lam / k / dx ** 2

    kx = np.zeros((n, n + 1))
    ky = np.zeros((n + 1, n))
    kx[:, 1:-1] = face(er[:, :-1], er[:, 1:], ks[:, :-1], ks[:, 1:])
    ky[1:-1, :] = face(er[:-1, :], er[1:, :], ks[:-1, :], ks[1:, :])
    ksum = kx[:, :-1] + kx[:, 1:] + ky[:-1, :] + ky[1:, :]
    cap = rho / (gamma - 1.0)
    slope = 1.0 / (4.0 * a_r * T ** 3)
    # c_i = cap / dt * slope; choose dt so that the median of ksum / c_i is `stiffness`
    ratio = ksum / (cap * slope)
    dt = stiffness / np.median(ratio)
    c = cap / dt * slope
    u = lcg_uniform(seed + 13, n * n).reshape(n, n)
    b = c * er * (1.0 + 0.1 * (2.0 * u - 1.0))
    return c, kx, ky, b, dt


ENSEMBLE_SIDES = (128, 256, 512, 1024)


def ensemble_scenes(count: int = 64, seed: int = SEED_BASE + 0x1000, sides=ENSEMBLE_SIDES):
    """The batched-ensemble workload: `count` independent scenes, each a square of side drawn uniformly from `sides`
    (log-uniform sizes over 128..1024), 1-3 Gaussian blobs (peak rho U[0.1, 1], sigma U[0.03, 0.15] L, centres
    U[0.25, 0.75] L) on a lognormal background (sigma_ln 1.0, mean 1e-2), T = 1 +- 1% seeded, at rest.
    Returns [(side, planes)] with planes float32 (14, side, side)."""
    u = lcg_uniform(seed, count * 14).reshape(count, 14)
    out = []
    for k in range(count):
        side = int(sides[min(len(sides) - 1, int(u[k, 0] * len(sides)))])
        blobs = 1 + min(2, int(u[k, 1] * 3))
        c = (np.arange(side) + 0.5) / side
        rho = 1e-2 * lognormal(side, 1.0, seed + 16 + 2 * k)
        for b in range(blobs):
            peak, sig, cx, cy = u[k, 2 + 4 * b:6 + 4 * b]
            peak, sig = 0.1 + 0.9 * peak, 0.03 + 0.12 * sig
            cx, cy = 0.25 + 0.5 * cx, 0.25 + 0.5 * cy
            rho = rho + peak * np.exp(-((c[None, :] - cx) ** 2 + (c[:, None] - cy) ** 2) / (2.0 * sig * sig))
        T = 1.0 + 0.01 * (2.0 * lcg_uniform(seed + 17 + 2 * k, side * side).reshape(side, side) - 1.0)
        out.append((side, planes_from(rho, T).astype(np.float32)))
    return out
