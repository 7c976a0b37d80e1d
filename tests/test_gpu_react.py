"""P6 burning on the GPU (sandbox_reactions_amd/csrc/react.hip) through the C-ABI (sr_set_reactions, sr_react_apply,
sr_read_react): the reference's GPU burn cases (crates/sr-engine/tests/gpu/burn.rs:626-770, f32 planes) with their
tolerances, the three laws to their closed forms within 1e-4, the heat booked and the energy gained within 1e-5, the
shipped limits, cells burning alone bit for bit; the library's cells against the oracle's on mixed cells; N_Fe in closed
form; P6 in the world's step at the step's dt."""
import math

import numpy as np
import pytest

import oracle
from sandbox_reactions_amd import react

pytestmark = pytest.mark.gpu

NS = 10
Q_H = 8700.0
X_TOL = 1e-4
DX = 1e-4
LADDER = {"t_h": 100.0, "t_he": 140.0, "t_c": 196.0, "t_ne": 274.0, "t_o": 300.0, "t_si": 420.0}
T_H, T_HE, T_NE = 100.0, 140.0, 274.0
SP = react.SPECIES
Z = np.array([1, 2, 6, 8, 10, 12, 14, 16, 26, 0], float)
A = np.array([1, 4, 12, 16, 20, 24, 28, 32, 56, 1], float)
W = 8  # the smallest world: cells in row order from (0, 0), vacuum around


def x_of(parts):
    x = np.zeros(NS)
    for k, v in parts:
        x[SP.index(k)] = v
    return x


def inv_mu(x):
    return float(np.sum(x * (Z + 1.0) / A))


def planes_of(cells, w=W, h=W):
    p = np.zeros((14, h, w), np.float32)
    p[4] = 1.0  # vacuum: Sigma 0, pure hydrogen
    for i, (s, t, x) in enumerate(cells):
        yy, xx = divmod(i, w)
        p[0, yy, xx] = s
        p[3, yy, xx] = s * t * inv_mu(x)
        p[4:, yy, xx] = x
    return p


class Rig:
    """A world with the flow off and P6's records: each pass is P6 alone over its dt."""

    def __init__(self, sr, params, cells, w=W, h=W):
        self.w, self.h = w, h
        self.world = sr.World(w, h, physics=sr.physics(), gravity=False, hydro=False, heat=False)
        self.world.set_reactions(params)
        self.load(cells)

    def load(self, cells):
        self.world.upload(planes_of(cells, self.w, self.h))

    def run(self, dt):
        self.world.react_apply(dt)
        return self.world.read_react()

    def planes(self):
        return self.world.readback().astype(np.float64)


def cell(p, i, w=W):
    yy, xx = divmod(i, w)
    return p[:, yy, xx]


def frac(p, sp, i=0):
    return cell(p, i)[4 + SP.index(sp)]


def temp(p, i=0):
    c = cell(p, i)
    return c[3] / c[0] / inv_mu(c[4:])


def gamow(t, t_k, nu):
    return (t / t_k) ** (-2.0 / 3.0) * math.exp((3.0 * nu + 2.0) * (1.0 - (t_k / t) ** (1.0 / 3.0)))


def f_h(t, t_h):
    return 0.0 if t < 0.5 * t_h else gamow(t, t_h, 4.0) + gamow(t, 1.6 * t_h, 18.0)


K_H, K_HE, K_NE = 5e-4 * 100.0 * f_h(T_H, T_H), 1e-3 * 10.0 * 10.0, 1e-3 * 50.0
LAWS = [
    ("H_burn", "H", (100.0, T_H, x_of([("H", 1.0)])), [v / K_H for v in (0.5, 1.0, 2.0, 4.0, 6.0, 9.0)],
     lambda t: 1.0 / (1.0 + K_H * t)),
    ("He_burn", "He", (10.0, T_HE, x_of([("He", 1.0)])),
     [(m - 1.0) / (2.0 * K_HE) for m in (1.25, 2.0, 4.0, 9.0, 25.0, 100.0)],
     lambda t: 1.0 / math.sqrt(1.0 + 2.0 * K_HE * t)),
    ("Ne_burn", "Ne", (50.0, T_NE, x_of([("Ne", 0.7), ("Fe", 0.3)])),
     [v / K_NE for v in (0.2, 0.5, 1.0, 1.5, 2.0, math.log(10.0))], lambda t: 0.7 * math.exp(-K_NE * t)),
]


@pytest.mark.parametrize("rid,species,c0,at,closed", LAWS, ids=[l[0] for l in LAWS])
def test_the_three_laws_follow_their_closed_forms(gpu_lib, rid, species, c0, at, closed):
    sr = gpu_lib
    print(f"P6 {rid} closed form", flush=True)
    r = Rig(sr, react.params(only=[rid], overrides=dict(LADDER, burn_dx=DX), pin_temperature=True), [c0])
    t_prev = 0.0
    for t in at:
        rep = r.run(t - t_prev)
        t_prev = t
        p = r.planes()
        assert abs(frac(p, species) - closed(t)) <= X_TOL, (rid, t, frac(p, species), closed(t))
        assert abs(cell(p, 0)[4:].sum() - 1.0) <= 1e-6 and rep.unfinished == 0
    p = r.planes()
    assert abs(temp(p) - c0[1]) <= 1e-5 * c0[1]
    if rid == "Ne_burn":
        assert abs(frac(p, "Fe") - 0.3) <= 1e-6


@pytest.mark.parametrize("rid,group,q_ratio,parts,sigma,t,run", [
    ("H_burn", "H", 1.0, [("H", 1.0)], 100.0, T_H, 9.0 / K_H),
    ("Ne_burn", "Ne", 1.0 / 57.0, [("Ne", 0.7), ("Fe", 0.3)], 50.0, T_NE, math.log(10.0) / K_NE),
])
def test_the_heat_is_booked_per_substep(gpu_lib, rid, group, q_ratio, parts, sigma, t, run):
    sr = gpu_lib
    r = Rig(sr, react.params(only=[rid], overrides=dict(LADDER, burn_dx=DX), pin_temperature=True),
            [(sigma, t, x_of(parts))])
    before = frac(r.planes(), parts[0][0])
    rep = r.run(run)
    dm = sigma * (before - frac(r.planes(), parts[0][0]))
    want = q_ratio * Q_H * dm
    g = react.GROUPS.index(group)
    nuc, _, _ = r.world.read_react_planes()
    print(f"P6 {rid}: booked {rep.nuclear[g]:.6f} for q dm {want:.6f}, {rep.substeps} sub-steps", flush=True)
    assert rep.unfinished == 0 and abs(rep.nuclear[g] - want) <= 1e-5 * want
    assert all(rep.nuclear[o] == 0.0 for o in range(7) if o != g) and nuc[g, 0, 0] == rep.nuclear[g]


def test_the_heat_enters_the_energy(gpu_lib):
    sr = gpu_lib
    sigma, t0 = 1.0, 90.0
    r = Rig(sr, react.params(only=["H_burn"], overrides=dict(burn_dx=1e-3, burn_dt=0.01)),
            [(sigma, t0, x_of([("H", 0.3), ("He", 0.7)]))])
    start = r.planes()
    rep = r.run(400.0)
    p = r.planes()
    dm = sigma * (frac(start, "H") - frac(p, "H"))
    de, booked = cell(p, 0)[3] - cell(start, 0)[3], rep.nuclear[0]
    assert dm > 0.05 * sigma and temp(p) > 1.1 * t0
    assert abs(de - booked) <= 1e-5 * booked and abs(booked - Q_H * dm) <= 1e-5 * Q_H * dm


MIX = x_of([("H", 0.2), ("He", 0.2), ("C", 0.15), ("O", 0.15), ("Ne", 0.1), ("Mg", 0.05), ("Si", 0.05), ("S", 0.04),
            ("Fe", 0.06)])
SUN = x_of([("H", 0.7346), ("He", 0.2485), ("O", 0.0077), ("C", 0.0092)])


def test_the_fractions_stay_physical_at_the_shipped_limits(gpu_lib):
    sr = gpu_lib
    cells = [(s, t, MIX) for s, t in ((1.0, 150.0), (10.0, 300.0), (100.0, 450.0), (400.0, 520.0), (0.5, 60.0),
                                      (30.0, 250.0))]
    cells += [(1.0, 56.0, SUN), (1.0, 56.0, SUN)]
    r = Rig(sr, react.params(), cells)
    a = r.run(5.0)
    b = r.run(45.0)
    p = r.planes()
    print(f"P6 shipped limits: {a.substeps} + {b.substeps} sub-steps, most {a.max_substeps} / {b.max_substeps}, "
          f"runaway T {temp(p, 6):.1f}", flush=True)
    assert a.unfinished + b.unfinished == 0
    assert b.max_substeps > 20 and temp(p, 6) > 1000.0
    for i in range(8):
        x = cell(p, i)[4:]
        assert abs(x.sum() - 1.0) <= 1e-6 and np.all(x >= 0.0), i


CELLS = [(5.0, 120.0, x_of([("H", 1.0)])), (20.0, 200.0, x_of([("He", 0.625), ("C", 0.375)])),
         (1.0e-8, 400.0, x_of([("H", 1.0)])), (5.0, 20.0, x_of([("H", 0.5), ("He", 0.5)])),
         (50.0, 330.0, x_of([("Ne", 0.5), ("O", 0.25), ("Mg", 0.25)])), (50.0, 480.0, x_of([("Si", 0.625), ("S", 0.375)]))]


def test_a_cell_burns_alone(gpu_lib):
    """Every channel and every group's slot of each cell of a mixed world equals the same cell stepped alone, bit for
    bit; the cold cell is left bit for bit and the vacuum and cold cells book nothing."""
    sr = gpu_lib
    params = react.params()
    r = Rig(sr, params, CELLS)
    before = r.planes()
    r.run(2.0)
    p = r.planes()
    nuc, _, _ = r.world.read_react_planes()
    solo = Rig(sr, params, [CELLS[0]])
    for i, c in enumerate(CELLS):
        if i == 2:
            continue
        solo.load([c])
        solo.run(2.0)
        q = solo.planes()
        nq, _, _ = solo.world.read_react_planes()
        assert np.array_equal(cell(p, i).view(np.uint64), cell(q, 0).view(np.uint64)), i
        yy, xx = divmod(i, W)
        assert np.array_equal(nuc[:, yy, xx], nq[:, 0, 0]), i
    assert np.array_equal(cell(p, 3).view(np.uint64), cell(before, 3).view(np.uint64))
    assert not nuc[:, 0, 2].any() and not nuc[:, 0, 3].any() and nuc[:, 0, 0].any()


def random_cells(n, seed):
    rng = np.random.default_rng(seed)
    cells = []
    for _ in range(n):
        x = rng.dirichlet(np.ones(NS) * 0.5)
        x[9] = 0.0
        x /= x.sum()
        cells.append((float(10 ** rng.uniform(-0.3, 2.6)), float(rng.uniform(30.0, 500.0)), x))
    return cells


def test_the_cells_are_the_oracles(gpu_lib):
    """64 mixed cells (random fractions, Sigma 0.5-400, T 30-500), every record, the cooling on, one pass of 0.5: the
    library's fractions against the oracle's on the same f32 inputs within 1e-5, its energy within 1e-5 relative, the
    bookings per group within 1e-5 of the pass's nuclear energy; the report's counts the oracle's."""
    sr = gpu_lib
    params = react.params(neutrinos=True)
    cells = random_cells(64, 7)
    r = Rig(sr, params, cells)
    inp = r.planes()
    rep = r.run(0.5)
    p = r.planes()
    nuc, lost, _ = r.world.read_react_planes()
    o = np.ascontiguousarray(inp)
    orep, onuc, olost, _ = oracle.react_step(oracle.reference_physics(), params, o, 0.5)
    print(f"P6 64 cells: {rep.substeps} sub-steps (oracle {orep.substeps}), retried {rep.retried} ({orep.retried})",
          flush=True)
    assert rep.unfinished == orep.unfinished == 0
    g = np.abs(p[4:] - o[4:]).max()
    e = np.abs(p[3] - o[3]) / np.maximum(np.abs(o[3]), 1e-30)
    assert g <= 1e-5 and e.max() <= 1e-5, (g, e.max())
    total = np.abs(onuc).sum()
    assert np.abs(nuc - onuc).sum() <= 1e-5 * total and abs(lost.sum() - olost.sum()) <= 1e-5 * max(olost.sum(), 1e-30)
    assert abs(rep.substeps - orep.substeps) <= 2 and rep.latch == orep.latch


def test_n_fe_steps_in_closed_form(gpu_lib):
    """The iron block (pure Fe, Sigma 400, T 1) over 20 passes of 5.87e-3 with every record: no sub-step, the latch
    set, X_n the oracle's within 1e-6 (f32 planes), Fe to n mass for mass, the heat in N alone."""
    sr = gpu_lib
    params = react.params()
    c0 = (400.0, 1.0, x_of([("Fe", 1.0)]))
    r = Rig(sr, params, [c0])
    o = np.ascontiguousarray(r.planes())
    for step in range(20):
        rep = r.run(5.87e-3)
        orep, *_ = oracle.react_step(oracle.reference_physics(), params, o, 5.87e-3)
        assert rep.substeps == 0 and rep.latch, step
        p = r.planes()
        assert abs(frac(p, "n") - o[4 + 9, 0, 0]) <= 1e-6, step
        assert abs(frac(p, "n") + frac(p, "Fe") - 1.0) <= 1e-6
        n = react.GROUPS.index("N")
        assert all(rep.nuclear[k] == 0.0 for k in range(7) if k != n) and rep.nuclear[n] < 0.0


def test_p6_runs_in_the_step_at_the_steps_dt(gpu_lib):
    """A world with the flow off and P6 on: one step (P1, P6, P8, P9) is P6 alone at that step's dt followed by a step
    with the records off (P8's floors and renormalisation once each), the same bits in every cell; with the records off
    the fractions move by no more than P8's renormalisation."""
    sr = gpu_lib
    params = react.params()
    cells = random_cells(40, 3)
    a = Rig(sr, params, cells)
    a.world.run(1)
    dt = a.world.read_dt().dt
    b = Rig(sr, params, cells)
    b.run(dt)
    b.world.set_reactions(None)
    b.world.run(1)
    pa, pb = a.planes(), b.planes()
    gas = pa[0] >= 1e-6
    assert gas.sum() == 40
    assert np.array_equal(pa.view(np.uint64), pb.view(np.uint64))
    assert not np.array_equal(pa[4:, gas], planes_of(cells)[4:, gas].astype(np.float64))
    c = Rig(sr, params, cells)
    c.world.set_reactions(None)
    c.world.run(1)
    assert np.abs(c.planes()[4:, gas] - planes_of(cells)[4:, gas]).max() <= 1e-6
    with pytest.raises(sr.SrError):
        bad = react.params()
        bad.rec[0].group = 9
        a.world.set_reactions(bad)


@pytest.mark.parametrize("dt", [5e-3, 5e-2])
def test_the_burning_field_is_the_oracles(gpu_lib, dt):
    """The benchmark's burning field (the reference's ref_burn.json made uneven: Sigma 30 and T 100 lognormal, fresh
    and burnt fractions mixed, the cold pressure on, burn_k 10) at 128^2, one pass at the step's dt (one sub-step per
    cell) and at ten times it (sub-cycled, with retries): the fractions within 1e-5 of the oracle's from the same f32
    planes, E within 1e-5 relative, the bookings within 1e-5 of the pass's total, the same sub-steps within 0.1 %."""
    from sandbox_reactions_amd import scenes
    sr = gpu_lib
    n = 128
    print(f"P6 burning field {n}^2 dt {dt}", flush=True)
    planes, kw = scenes.burn_scene(n)
    params = react.params(overrides={"burn_k": 10.0})
    w = sr.World(n, n, physics=sr.physics(**kw), gravity=False, hydro=False, heat=False)
    w.set_reactions(params)
    w.upload(planes)
    w.react_apply(dt)
    rep = w.read_react()
    p = w.readback().astype(np.float64)
    nuc, lost, _ = w.read_react_planes()
    o = planes.astype(np.float64)
    orep, onuc, olost, _ = oracle.react_step(oracle.reference_physics(**kw), params, o, dt)
    print(f"P6 field: {rep.substeps} sub-steps (oracle {orep.substeps}), retried {rep.retried} ({orep.retried}), "
          f"most {rep.max_substeps}", flush=True)
    assert rep.unfinished == orep.unfinished == 0 and orep.substeps >= n * n * 0.9
    assert np.abs(p[4:] - o[4:]).max() <= 1e-5
    assert (np.abs(p[3] - o[3]) / np.abs(o[3])).max() <= 1e-5
    assert np.abs(nuc - onuc).sum() <= 1e-5 * np.abs(onuc).sum()
    assert abs(rep.substeps - orep.substeps) <= 1e-3 * orep.substeps
