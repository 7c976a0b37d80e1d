"""Shared test helpers: one set of constants for the oracle and the HIP library, comparisons."""
import numpy as np

import oracle

# the reference's physics.json values the path reads (assets/physics.json)
REFERENCE = dict(gamma=2.0, G=1.0, dx=1.0, sigma_vac=1e-6, sigma_floor=1e-8, t_floor=0.05, cfl=0.4, eta_g=0.3,
                 a_r=6.25e-7, c_sb=12800.0, kappa0=1.0, k_cond=0.0, kappa_dust=0.0, t_dust=0.0, opacity=0, mu=0.0,
                 edge=1, pin_temperature=0, constant_diffusivity=0.0, flux_limiter=1)


def both(**over):
    """(oracle physics, library physics) with the same constants."""
    import sandbox_reactions_amd as sr
    d = dict(REFERENCE)
    d.update(over)
    mp = d.pop("max_padded", 4096)
    return oracle.reference_physics(max_padded=float(mp), **d), sr.physics(max_padded=int(mp), **d)


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sum(np.abs(b))
    return float(np.sum(np.abs(a - b)) / den) if den > 0 else float(np.sum(np.abs(a - b)))


def f32(planes):
    """Planes as the device holds them (f32 everywhere; Sigma starts from the same f32 value)."""
    return np.asarray(planes, dtype=np.float32).astype(np.float64)


def mass_radius(sigma, vac=1e-6, fraction=0.9):
    """The radius about the centre of mass holding `fraction` of the mass, each cell of Sigma >= vac split into 4 x 4
    points of its area, the crossing interpolated between sorted radii (the reference's tracker,
    crates/sr-engine/src/calibrate.rs:333-368, and crates/sr-engine/tests/gpu/collapse.rs:81-113's r50, restated);
    and the mass."""
    ys, xs = np.nonzero(sigma >= vac)
    m = sigma[ys, xs].astype(np.float64)
    mass = m.sum()
    cx, cy = ((xs + 0.5) * m).sum() / mass, ((ys + 0.5) * m).sum() / mass
    off = (np.arange(4) + 0.5) / 4 - 0.5
    a, b = np.meshgrid(off, off, indexing="ij")
    px = (xs[:, None] + 0.5 + a.ravel()[None, :]).ravel()
    py = (ys[:, None] + 0.5 + b.ravel()[None, :]).ravel()
    pm = np.repeat(m / 16.0, 16)
    r = np.hypot(px - cx, py - cy)
    o = np.argsort(r, kind="stable")
    r, pm = r[o], pm[o]
    c = np.cumsum(pm)
    k = int(np.searchsorted(c, fraction * mass))
    below, rp = (c[k - 1], r[k - 1]) if k > 0 else (0.0, 0.0)
    return rp + (r[k] - rp) * (fraction * mass - below) / pm[k], mass
