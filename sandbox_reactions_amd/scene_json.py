"""Scene files of the reference engine (version 1 JSON, crates/sr-engine/src/scene.rs:328-575) turned into the 14
canonical planes this library uploads, with the world, edge rule, switches, test settings and physics overrides.

Covered: `rect` and `disk` objects (Maclaurin profile sampled at cell centres, scene.rs:545-556), each with `t` or
`pressure`, a `composition` of mass fractions summing to 1 (within 1e-6, scene.rs:46, :215-245) and a velocity `v`;
overlapping objects add their mass, momentum, energy and species mass (scene.rs:485-497); cells no object covers
hold the floor state, Sigma_floor of hydrogen at rest at T_floor (scene.rs:466-470). Refused with a message: anything
the path does not carry -- `preset` objects (they need the calibration file), the neutrinos switch, sinks with a
neutron pair other than the shipped one, and, with the cold pressure on, a start with a cell the hold would take
(Sigma >= sigma_hold and cold: the hold of floors.rs:200-260, which the path does not carry). The reactions switch
gives the world P6 with every shipped record (`Scene.react_params`), the sinks switch P3 (`Scene.sinks_params`).

The energy follows the EOS with this library's adiabatic index gamma (the reference's gamma = 2 by default) and the
cold pressure of the shipped constants (k1e, k2e, k1n, k2n; a scene's overrides may turn a pair off):
eps_th = T / (mu (gamma - 1)) for `t`, eps_th = (Pi - Pi_e - Pi_n) / ((gamma - 1) Sigma) for `pressure` (refused where
the cold pressure alone exceeds Pi, scene.rs:454-457), plus eps_cold and the kinetic |v|^2 / 2.
"""
from __future__ import annotations

import json
import math

import numpy as np

from . import cold as cold_eos

# assets/elements.json: key, A, Z (the order of the species planes)
ELEMENTS = [("H", 1, 1), ("He", 4, 2), ("C", 12, 6), ("O", 16, 8), ("Ne", 20, 10), ("Mg", 24, 12), ("Si", 28, 14),
            ("S", 32, 16), ("Fe", 56, 26), ("n", 1, 0)]
INDEX = {k: i for i, (k, _, _) in enumerate(ELEMENTS)}
COMPOSITION_TOLERANCE = 1e-6
# the assets/physics.json constants this path reads, and their shipped values
SHIPPED = {"sigma_floor": 1e-8, "sigma_vac": 1e-6, "t_floor": 0.05, "a_r": 6.25e-7, "c_sb": 12800.0, "kappa0": 1.0,
           "k_cond": 0.0, "cfl": 0.4, "eta_g": 0.3, "kappa_dust": 0.0, "t_dust": 0.0}
COLD = ("k1e", "k2e", "k1n", "k2n")
# m_tov_sb for the shipped constants: the sheet's neutron-matter ceiling as the reference computes it (sinks.rs:104-115)
M_TOV_SHIPPED = 4853.17
PHYSICS_KEYS = {"t_h", "t_he", "t_c", "t_ne", "t_o", "t_si", "q_h", "a_h_burn", "a_he_burn", "a_c_alpha", "a_c_burn",
                "a_ne_burn", "a_o_burn", "a_mg_burn", "a_si_burn", "a_s_burn", "a_n_fe", "nu_c_burn", "nu_ne_burn",
                "nu_o_burn", "nu_mg_burn", "nu_si_burn", "nu_s_burn", "sigma_n", "burn_dx", "burn_dt", "burn_k",
                "sigma_floor", "sigma_vac", "t_floor", "sigma_obj", "paint_sigma", "paint_t", "heat_factor",
                "cool_factor", "k1e", "k2e", "k1n", "k2n", "a_r", "c_sb", "kappa0", "k_cond", "a_nu", "t_nu", "m_nu",
                "f_dep", "sigma_bh", "cfl", "eta_g", "kappa_dust", "t_dust", "block", "sigma_hold", "eta_hold",
                "r_hold"}


class SceneError(ValueError):
    pass


def inv_mu(comp):
    """1/mu = sum_s X_s (Z_s + 1) / A_s (crates/sr-physics/src/registry.rs:146-154)."""
    return sum(x * (z + 1.0) / a for x, (_, a, z) in zip(comp, ELEMENTS))


def _fields(obj, at, allowed):
    if not isinstance(obj, dict):
        raise SceneError(f"{at}: must be an object")
    for k in obj:
        if k not in allowed:
            raise SceneError(f"{at}.{k}: unknown key")
    return obj


def _num(f, key, at, positive=False, required=True):
    if key not in f:
        if required:
            raise SceneError(f"{at}.{key}: missing")
        return None
    v = f[key]
    if isinstance(v, bool) or not isinstance(v, (int, float)) or not math.isfinite(v):
        raise SceneError(f"{at}.{key}: must be a number")
    if positive and not v > 0:
        raise SceneError(f"{at}.{key}: must be > 0, got {v}")
    return float(v)


def _whole(f, key, at):
    v = f.get(key)
    if isinstance(v, bool) or not isinstance(v, int):
        raise SceneError(f"{at}.{key}: must be a whole number")
    return v


def _gas(f, at):
    has_t, has_p = "t" in f, "pressure" in f
    if has_t and has_p:
        raise SceneError(f"{at}: takes t or pressure, not both")
    if not (has_t or has_p):
        raise SceneError(f"{at}: needs t or pressure")
    heat = ("t", _num(f, "t", at, True)) if has_t else ("pressure", _num(f, "pressure", at, True))
    comp_map = f.get("composition")
    if not isinstance(comp_map, dict):
        raise SceneError(f"{at}.composition: must be an object of mass fractions")
    comp = [0.0] * len(ELEMENTS)
    for k, v in comp_map.items():
        if k not in INDEX:
            raise SceneError(f"{at}.composition.{k}: unknown key")
        if isinstance(v, bool) or not isinstance(v, (int, float)) or not (math.isfinite(v) and v >= 0):
            raise SceneError(f"{at}.composition.{k}: must be a number >= 0")
        comp[INDEX[k]] = float(v)
    if abs(sum(comp) - 1.0) > COMPOSITION_TOLERANCE:
        raise SceneError(f"{at}.composition: the fractions sum to {sum(comp)}, not 1")
    v = f.get("v", [0.0, 0.0])
    if not (isinstance(v, list) and len(v) == 2 and all(isinstance(c, (int, float)) and math.isfinite(c) for c in v)):
        raise SceneError(f"{at}.v: must be [vx, vy], two finite numbers")
    return heat, comp, (float(v[0]), float(v[1]))


def _flag(f, key, at, default):
    v = f.get(key, default)
    if not isinstance(v, bool):
        raise SceneError(f"{at}.{key}: must be true or false")
    return v


class Scene:
    """A parsed scene: world (w, h), edge (0 bounce / 1 leave), switches, test settings, overrides, objects."""

    def __init__(self, root: dict):
        top = _fields(root, "scene", {"version", "world", "slowdown", "switches", "test", "objects", "overrides"})
        if top.get("version") != 1:
            raise SceneError("version: must be 1")
        w = _fields(top.get("world"), "world", {"width", "height", "edge"})
        self.width, self.height = _whole(w, "width", "world"), _whole(w, "height", "world")
        if not (8 <= self.width <= 4096 and 8 <= self.height <= 4096 and self.width % 8 == 0 and self.height % 8 == 0):
            raise SceneError(f"world: {self.width} x {self.height} is not a world size (multiples of 8 in [8, 4096])")
        # the file's own world: objects are placed relative to it, and a resized world shifts them (resize())
        self.file_width, self.file_height = self.width, self.height
        edge = w.get("edge", "leave")
        if edge not in ("leave", "bounce"):
            raise SceneError(f"world.edge: must be \"leave\" or \"bounce\", got {edge!r}")
        self.edge = 1 if edge == "leave" else 0
        self.slowdown = _flag(top, "slowdown", "scene", True)
        sw = _fields(top.get("switches", {}), "switches", {"gravity", "hydro", "heat", "reactions", "neutrinos",
                                                             "sinks"})
        self.switches = {k: _flag(sw, k, "switches", True) for k in ("gravity", "hydro", "heat", "reactions",
                                                                    "neutrinos", "sinks")}
        t = _fields(top.get("test", {}), "test", {"pin_temperature", "constant_diffusivity", "flux_limiter"})
        cd = t.get("constant_diffusivity")
        if cd is not None:
            cd = _num(t, "constant_diffusivity", "test", True)
        self.test = {"pin_temperature": _flag(t, "pin_temperature", "test", False),
                     "constant_diffusivity": cd, "flux_limiter": _flag(t, "flux_limiter", "test", True)}
        ov = top.get("overrides", {})
        if not isinstance(ov, dict):
            raise SceneError("overrides: must be an object of physics.json keys")
        self.overrides = {}
        for k, v in ov.items():
            if k not in PHYSICS_KEYS:
                raise SceneError(f"overrides.{k}: unknown key (physics.json keys only)")
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                raise SceneError(f"overrides.{k}: must be a number")
            self.overrides[k] = float(v)
        for a, b in (("k1e", "k2e"), ("k1n", "k2n")):
            ka, kb = (self.overrides.get(k, cold_eos.SHIPPED[k]) for k in (a, b))
            if not (isinstance(ka, (int, float)) and isinstance(kb, (int, float))) or ka < 0 or kb < 0 or \
                    (ka > 0) != (kb > 0):
                raise SceneError(f"overrides.{a}, {b}: a cold pressure pair is on or off as a whole")
        self.sigma_hold = float(self.overrides.get("sigma_hold", 10.0))
        objs = top.get("objects")
        if not isinstance(objs, list):
            raise SceneError("objects: must be a list")
        self.objects = []
        for i, o in enumerate(objs):
            at = f"objects[{i}]"
            kind = o.get("kind") if isinstance(o, dict) else None
            if kind == "rect":
                f = _fields(o, at, {"kind", "x0", "y0", "w", "h", "sigma", "t", "pressure", "composition", "v"})
                rw, rh = _whole(f, "w", at), _whole(f, "h", at)
                if rw < 1 or rh < 1:
                    raise SceneError(f"{at}: w and h must be >= 1, got {rw} x {rh}")
                self.objects.append(("rect", (_whole(f, "x0", at), _whole(f, "y0", at), rw, rh,
                                              _num(f, "sigma", at, True)), _gas(f, at)))
            elif kind == "disk":
                f = _fields(o, at, {"kind", "x", "y", "a", "sigma_c", "t", "pressure", "composition", "v"})
                self.objects.append(("disk", (_num(f, "x", at), _num(f, "y", at), _num(f, "a", at, True),
                                              _num(f, "sigma_c", at, True)), _gas(f, at)))
            elif kind == "preset":
                raise SceneError(f"{at}: preset objects need the calibration file; not carried by this path")
            else:
                raise SceneError(f"{at}.kind: must be \"preset\", \"disk\" or \"rect\"")

    def check_path(self):
        """Refuses a scene that switches on a pass this path does not carry: neutrinos (P7's heating, and P6's cooling
        that feeds it); and sinks with a neutron pair other than the shipped one or none (sinks_params)."""
        if self.switches["neutrinos"]:
            raise SceneError("switches.neutrinos: this path carries gravity, hydro, heat, reactions and sinks only "
                             "(set neutrinos false)")
        self.sinks_params()

    def sinks_params(self):
        """P3's (Sigma_BH, m_tov_sb) for this scene, None with the sinks switch off. m_tov_sb is the sheet's
        neutron-matter ceiling (sinks.rs:104-115), which this path does not compute: the reference's value for the
        shipped pair (4,853.17), infinite with the neutrons' cold pressure off (no sink from mass); another pair is
        refused."""
        if not self.switches["sinks"]:
            return None
        c = self.constants()
        if c["k1n"] == 0.0 and c["k2n"] == 0.0:
            m_tov = math.inf
        elif (c["k1n"], c["k2n"]) == (cold_eos.SHIPPED["k1n"], cold_eos.SHIPPED["k2n"]):
            m_tov = M_TOV_SHIPPED
        else:
            raise SceneError("switches.sinks: m_tov_sb for a neutron pair other than the shipped one (the sheet's "
                             "ceiling) is not carried by this path")
        return float(self.overrides.get("sigma_bh", 6000.0)), m_tov

    def react_params(self):
        """P6's parameters for this scene (Step::with_reactions's ReactSetup, cpu_gpu.rs:72-85): every shipped record,
        resolved against the shipped constants with the scene's overrides, its pin_temperature test flag and its
        neutrinos switch; None with the reactions switch off."""
        if not self.switches["reactions"]:
            return None
        from . import react
        return react.params(overrides=self.overrides, pin_temperature=self.test["pin_temperature"],
                            neutrinos=self.switches["neutrinos"])

    def constants(self) -> dict:
        """The physics constants this path reads, shipped values with the scene's overrides."""
        c = dict(SHIPPED, **cold_eos.SHIPPED)
        for k in c:
            if k in self.overrides:
                c[k] = float(self.overrides[k])
        return c

    def cold_on(self) -> bool:
        c = self.constants()
        return c["k1e"] > 0.0 or c["k1n"] > 0.0

    def resize(self, width: int, height: int):
        """The world `--world` stands over the scene's (headless.rs:182-187). The objects keep their place relative to
        the world's centre: each is shifted by ((W - W_file) / 2, (H - H_file) / 2) cells (scene.rs:428-435); both
        sizes are multiples of 8, so the shift is whole."""
        if not (8 <= width <= 4096 and 8 <= height <= 4096 and width % 8 == 0 and height % 8 == 0):
            raise SceneError(f"world: {width} x {height} is not a world size (multiples of 8 in [8, 4096])")
        self.width, self.height = width, height

    def shift(self):
        return (self.width - self.file_width) // 2, (self.height - self.file_height) // 2

    def cells(self, obj):
        """(ys, xs, sigma) of the cells an object covers in the current world, shifted as resize() says
        (scene.rs:521-563)."""
        kind, geo, _ = obj
        w, h = self.width, self.height
        sx, sy = self.shift()
        if kind == "rect":
            x0, y0, rw, rh, sigma = geo
            x0, y0 = x0 + sx, y0 + sy
            xs = np.arange(min(max(x0, 0), w), min(max(x0 + rw, 0), w))
            ys = np.arange(min(max(y0, 0), h), min(max(y0 + rh, 0), h))
            yy, xx = np.meshgrid(ys, xs, indexing="ij")
            return yy.ravel(), xx.ravel(), np.full(yy.size, sigma)
        x, y, a, sc = geo
        x, y = x + sx, y + sy
        i = np.arange(math.floor(x - a), math.ceil(x + a))
        j = np.arange(math.floor(y - a), math.ceil(y + a))
        jj, ii = np.meshgrid(j, i, indexing="ij")
        r2 = ((ii + 0.5 - x) ** 2 + (jj + 0.5 - y) ** 2) / (a * a)
        keep = (r2 < 1.0) & (ii >= 0) & (jj >= 0) & (ii < w) & (jj < h)
        return jj[keep], ii[keep], sc * np.sqrt(1.0 - r2[keep])

    def planes(self, gamma: float = 2.0) -> np.ndarray:
        """The 14 planes (f32, (14, h, w)) as the reference's Scene::planes makes them (scene.rs:428-471)."""
        w, h = self.width, self.height
        n = w * h
        sig, mx, my, en = (np.zeros(n) for _ in range(4))
        spm = np.zeros((len(ELEMENTS), n))
        c = self.constants()
        for oi, obj in enumerate(self.objects):
            ys, xs, s = self.cells(obj)
            (kind, value), comp, v = obj[2]
            idx = ys * w + xs
            ye, xn = float(cold_eos.y_e(comp)), float(comp[-1])
            if kind == "t":
                eps = value / ((gamma - 1.0) / inv_mu(comp)) * np.ones_like(s)
            else:
                thermal = value - cold_eos.pi_cold(c, s, ye, xn)
                if np.any(thermal <= 0.0):
                    k = int(np.argmax(thermal <= 0.0))
                    raise SceneError(f"objects[{oi}].pressure: {value} is below the cold pressure at cell "
                                     f"({xs[k]}, {ys[k]})")
                eps = thermal / ((gamma - 1.0) * s)
            eps = eps + cold_eos.eps_cold(c, s, ye, xn) + 0.5 * (v[0] * v[0] + v[1] * v[1])
            np.add.at(sig, idx, s)
            np.add.at(mx, idx, s * v[0])
            np.add.at(my, idx, s * v[1])
            np.add.at(en, idx, s * eps)
            for k, x in enumerate(comp):
                if x:
                    np.add.at(spm[k], idx, s * x)
        floor_sigma = c["sigma_floor"]
        floor_e = floor_sigma * (c["t_floor"] / ((gamma - 1.0) / inv_mu([1.0] + [0.0] * 9)) +
                                 float(cold_eos.eps_cold(c, floor_sigma, 1.0, 0.0)))
        if self.cold_on():
            self._refuse_held(c, sig, mx, my, en, spm)
        out = np.zeros((14, n), np.float64)
        empty = sig <= 0.0
        out[0] = np.where(empty, floor_sigma, sig)
        out[1] = np.where(empty, 0.0, mx)
        out[2] = np.where(empty, 0.0, my)
        out[3] = np.where(empty, floor_e, en)
        safe = np.where(empty, 1.0, sig)
        for k in range(len(ELEMENTS)):
            out[4 + k] = np.where(empty, 1.0 if k == 0 else 0.0, spm[k] / safe)
        return out.reshape(14, h, w).astype(np.float32)

    def _refuse_held(self, c, sig, mx, my, en, spm):
        """Refuses a start with a cell the hold would take (floors.rs:200-219, Floors::holds for a cell not yet held):
        Sigma >= sigma_hold, eps_th <= eta_hold eps_cold, and X_Fe < 1e-3 or X_n >= 0.5; the disk-mass gate is read
        as passing, so a cell it would free is refused too. The hold is not carried by this path."""
        eta = float(self.overrides.get("eta_hold", 0.1))
        for i in np.flatnonzero(sig >= self.sigma_hold):
            x = np.maximum(spm[:, i] / sig[i], 0.0)
            x = x / x.sum()
            cold = float(cold_eos.eps_cold(c, sig[i], float(cold_eos.y_e(x)), x[-1]))
            eps_th = en[i] / sig[i] - 0.5 * (mx[i] ** 2 + my[i] ** 2) / sig[i] ** 2 - cold
            if eps_th <= eta * cold and (x[8] < 1e-3 or x[9] >= 0.5):
                h, w = divmod(int(i), self.width)
                raise SceneError(f"the cell ({w}, {h}) starts held (Sigma {sig[i]:g} >= sigma_hold "
                                 f"{self.sigma_hold:g}, eps_th {eps_th:g} <= eta_hold eps_cold {eta * cold:g}): the "
                                 "hold (floors.rs:200-260) is not carried by this path")

    def physics_kwargs(self, gamma: float = 2.0, dx: float = 1.0) -> dict:
        """Keyword arguments of sandbox_reactions_amd.physics() for this scene (the reference opacity)."""
        kw = dict(self.constants(), gamma=gamma, dx=dx, edge=self.edge,
                  pin_temperature=int(self.test["pin_temperature"]),
                  constant_diffusivity=self.test["constant_diffusivity"] or 0.0,
                  flux_limiter=int(self.test["flux_limiter"]))
        return kw


def load(path: str) -> Scene:
    with open(path) as f:
        return Scene(json.load(f))


def parse(text: str) -> Scene:
    return Scene(json.loads(text))


def world(scene: Scene, heat_settings=None, planes=None, box_mode=None, **physics_over):
    """A World for the scene (its world size, switches, P6's records when reactions are on, and constants; gamma 2
    and unit cells unless overridden),
    with its planes uploaded (`planes`, (14, h, w), in place of the scene's own: a dump's cells) and, given
    `box_mode`, its active box in that mode (BOX_FIT, BOX_WHOLE_WORLD). Refuses a scene whose passes this path does
    not carry."""
    import sandbox_reactions_amd as sr
    scene.check_path()
    kw = scene.physics_kwargs()
    kw.update(physics_over)
    w = sr.World(scene.width, scene.height, physics=sr.physics(**kw), gravity=scene.switches["gravity"],
                 hydro=scene.switches["hydro"], heat=scene.switches["heat"], heat_settings=heat_settings)
    if box_mode is not None:
        w.set_box_mode(box_mode)
    rp = scene.react_params()
    if rp is not None:
        w.set_reactions(rp)
    sp = scene.sinks_params()
    if sp is not None:
        w.set_sinks(*sp)
    w.upload(scene.planes(kw["gamma"]) if planes is None else planes)
    return w
