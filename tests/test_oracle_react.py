"""The oracle's P6 burning (oracle/react.c, restating crates/sr-physics/src/reference/react.rs and rates.rs) pinned by
the reference's own G-BURN and G-ORDER cases (crates/sr-engine/tests/gpu/burn.rs:239-1133), whose oracles are
closed forms written from the contract, not the twin: H_burn (order 2) X = X0/(1 + k X0 t), He_burn (order 3)
X = X0/sqrt(1 + 2 k X0^2 t), Ne_burn (order 1) X = X0 e^{-kt}, each to X0/10 within 1e-4; the heat booked q dm within
1e-5 of the mass the planes say moved; E gaining exactly that; the burn clock's q_max dt; the sub-cycle limits; N_Fe
against the ODE marched by RK4; the ignition order; each record booking into its group. Tolerances are the
reference's."""
import math

import numpy as np
import pytest

import oracle
from sandbox_reactions_amd import react

NS = 10
Q_H = 8700.0
X_TOL = 1e-4
DX = 1e-4
LADDER = {"t_h": 100.0, "t_he": 140.0, "t_c": 196.0, "t_ne": 274.0, "t_o": 300.0, "t_si": 420.0}
T_H, T_HE, T_NE = 100.0, 140.0, 274.0
SP = react.SPECIES
Z = np.array([1, 2, 6, 8, 10, 12, 14, 16, 26, 0], float)
A = np.array([1, 4, 12, 16, 20, 24, 28, 32, 56, 1], float)
PH = oracle.reference_physics()  # the shipped constants, the cold pressure off


def x_of(parts):
    x = np.zeros(NS)
    for k, v in parts:
        x[SP.index(k)] = v
    return x


def inv_mu(x):
    return float(np.sum(x * (Z + 1.0) / A))


def world(cells):
    """(14, 1, n) planes of cells (Sigma, T, X) at rest: E = Sigma T / mu (the cold pressure off)."""
    n = len(cells)
    p = np.zeros((14, 1, n))
    for i, (s, t, x) in enumerate(cells):
        p[0, 0, i] = s
        p[3, 0, i] = s * t * inv_mu(x)
        p[4:, 0, i] = x
    return p


def temp(p, i=0):
    return p[3, 0, i] / p[0, 0, i] / inv_mu(p[4:, 0, i])


def frac(p, sp, i=0):
    return p[4 + SP.index(sp), 0, i]


def gamow(t, t_k, nu):
    return (t / t_k) ** (-2.0 / 3.0) * math.exp((3.0 * nu + 2.0) * (1.0 - (t_k / t) ** (1.0 / 3.0)))


def f_h(t, t_h):
    return 0.0 if t < 0.5 * t_h else gamow(t, t_h, 4.0) + gamow(t, 1.6 * t_h, 18.0)


def follow(p, params, species, checkpoints, closed):
    t_prev, rep = 0.0, None
    for t in checkpoints:
        rep, *_ = oracle.react_step(PH, params, p, t - t_prev)
        t_prev = t
        got, want = frac(p, species), closed(t)
        assert abs(got - want) <= X_TOL, (species, t, got, want)
        assert abs(p[4:, 0, 0].sum() - 1.0) <= 1e-6
        assert rep.unfinished == 0
    return rep


def test_h_burn_second_order():
    params = react.params(only=["H_burn"], overrides=dict(LADDER, burn_dx=DX), pin_temperature=True)
    sigma, t, x0 = 100.0, T_H, 1.0
    k = 5e-4 * sigma * f_h(t, T_H)
    p = world([(sigma, t, x_of([("H", x0)]))])
    at = [v / (k * x0) for v in (0.5, 1.0, 2.0, 4.0, 6.0, 9.0)]
    follow(p, params, "H", at, lambda tt: x0 / (1.0 + k * x0 * tt))
    assert abs(temp(p) - t) < 1e-9 * t
    assert abs(frac(p, "He") - (1.0 - frac(p, "H"))) < 1e-9


def test_he_burn_third_order():
    params = react.params(only=["He_burn"], overrides=dict(LADDER, burn_dx=DX), pin_temperature=True)
    sigma, t, x0 = 10.0, T_HE, 1.0
    k = 1e-3 * sigma * sigma
    p = world([(sigma, t, x_of([("He", x0)]))])
    at = [(m - 1.0) / (2.0 * k * x0 * x0) for m in (1.25, 2.0, 4.0, 9.0, 25.0, 100.0)]
    follow(p, params, "He", at, lambda tt: x0 / math.sqrt(1.0 + 2.0 * k * x0 * x0 * tt))
    assert abs(frac(p, "C") - (1.0 - frac(p, "He"))) < 1e-9


def test_ne_burn_first_order():
    params = react.params(only=["Ne_burn"], overrides=dict(LADDER, burn_dx=DX), pin_temperature=True)
    sigma, t, x0 = 50.0, T_NE, 0.7
    k = 1e-3 * sigma
    p = world([(sigma, t, x_of([("Ne", x0), ("Fe", 0.3)]))])
    at = [v / k for v in (0.2, 0.5, 1.0, 1.5, 2.0, math.log(10.0))]
    follow(p, params, "Ne", at, lambda tt: x0 * math.exp(-k * tt))
    assert abs(frac(p, "Fe") - 0.3) < 1e-12
    lost = x0 - frac(p, "Ne")
    assert abs(frac(p, "O") - 0.4 * lost) < 1e-9 and abs(frac(p, "Mg") - 0.6 * lost) < 1e-9


@pytest.mark.parametrize("rid,group,q_ratio,parts,sigma,t,run", [
    ("H_burn", "H", 1.0, [("H", 1.0)], 100.0, T_H, 9.0 / (5e-4 * 100.0 * f_h(T_H, T_H))),
    ("Ne_burn", "Ne", 1.0 / 57.0, [("Ne", 0.7), ("Fe", 0.3)], 50.0, T_NE, math.log(10.0) / (1e-3 * 50.0)),
])
def test_heat_booked_per_substep(rid, group, q_ratio, parts, sigma, t, run):
    params = react.params(only=[rid], overrides=dict(LADDER, burn_dx=DX), pin_temperature=True)
    p = world([(sigma, t, x_of(parts))])
    before = frac(p, parts[0][0])
    rep, *_ = oracle.react_step(PH, params, p, run)
    dm = sigma * (before - frac(p, parts[0][0]))
    want = q_ratio * Q_H * dm
    g = react.GROUPS.index(group)
    assert abs(rep.nuclear[g] - want) <= 1e-5 * want
    assert all(rep.nuclear[o] == 0.0 for o in range(7) if o != g)


def test_heat_enters_the_energy():
    params = react.params(only=["H_burn"], overrides=dict(burn_dx=1e-3, burn_dt=0.01))
    sigma, t0 = 1.0, 90.0
    p = world([(sigma, t0, x_of([("H", 0.3), ("He", 0.7)]))])
    e0, h0 = p[3, 0, 0], frac(p, "H")
    rep, *_ = oracle.react_step(PH, params, p, 400.0)
    dm = sigma * (h0 - frac(p, "H"))
    de, booked = p[3, 0, 0] - e0, rep.nuclear[0]
    assert dm > 0.05 * sigma and temp(p) > 1.1 * t0
    assert abs(de - booked) <= 1e-9 * booked
    assert abs(booked - Q_H * dm) <= 1e-5 * Q_H * dm


GIANT = (2.4458, 128.54, x_of([("H", 0.7149), ("He", 0.2635), ("C", 0.0136), ("O", 0.0080)]))
SUN = (9.67, 34.1, x_of([("H", 0.7200), ("He", 0.2631), ("O", 0.0077), ("C", 0.0092)]))


def q_max(t, x, k_burn):
    eps_th = t * inv_mu(x)
    return eps_th * math.sqrt(2.0 * eps_th) / k_burn


def test_the_burn_clock_caps_the_heating():
    on, off = react.params(), react.params(overrides={"burn_k": 0.0})
    assert on.burn_k == 15.0
    sigma, t0, x = GIANT
    a, b = world([GIANT]), world([GIANT])
    (pa, na, *_), (pb, nb, *_) = oracle.react_step(PH, on, a, 1e-4), oracle.react_step(PH, off, b, 1e-4)
    got, unclocked = na.sum() / sigma, nb.sum() / sigma
    want = q_max(t0, x, 15.0) * 1e-4
    assert abs(got - want) <= 1e-3 * want and unclocked > 5.0 * got
    a, b = world([GIANT]), world([GIANT])
    (pa, *_), (pb, *_) = oracle.react_step(PH, on, a, 0.00907), oracle.react_step(PH, off, b, 0.00907)
    assert frac(b, "Fe") > 0.99
    assert frac(a, "H") > 0.7 and frac(a, "Fe") < 1e-6
    assert pa.unfinished == pb.unfinished == 0
    a, b = world([SUN]), world([SUN])
    (pa, na, *_), (pb, nb, *_) = oracle.react_step(PH, on, a, 0.0117), oracle.react_step(PH, off, b, 0.0117)
    assert pa.substeps > 0
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64)) and np.array_equal(na, nb)


def test_the_shipped_limits_hold():
    params = react.params()
    assert (params.burn_dx, params.burn_dt) == (0.1, 0.05)
    p = world([(1.0, 100.0, x_of([("H", 0.7346), ("He", 0.2485), ("O", 0.0077), ("C", 0.0092)]))])
    rep, *_ = oracle.react_step(PH, params, p, 50.0)
    assert rep.substeps > 20 and rep.unfinished == 0
    assert rep.max_dx <= 0.1 * (1.0 + 1e-12) and 0.02 < rep.max_dt_rel <= 0.05 * (1.0 + 1e-12)


def test_the_shipped_burn_dx_overestimates_within_0_05():
    params = react.params(only=["H_burn"], overrides=LADDER, pin_temperature=True)
    sigma, t, x0 = 100.0, T_H, 1.0
    k = 5e-4 * sigma * f_h(t, T_H)
    run = 9.0 / (k * x0)
    p = world([(sigma, t, x_of([("H", x0)]))])
    oracle.react_step(PH, params, p, run)
    got, want = frac(p, "H"), x0 / (1.0 + k * x0 * run)
    assert want <= got < want + 0.05


MIX = x_of([("H", 0.2), ("He", 0.2), ("C", 0.15), ("O", 0.15), ("Ne", 0.1), ("Mg", 0.05), ("Si", 0.05), ("S", 0.04),
            ("Fe", 0.06)])


def test_fractions_stay_physical():
    cells = [(s, t, MIX) for s, t in ((1.0, 150.0), (10.0, 300.0), (100.0, 450.0), (400.0, 520.0), (0.5, 60.0),
                                      (30.0, 250.0))]
    p = world(cells)
    rep, *_ = oracle.react_step(PH, react.params(), p, 5.0)
    assert rep.unfinished == 0 and rep.substeps > 6
    x = p[4:, 0, :]
    assert np.all(np.abs(x.sum(axis=0) - 1.0) <= 1e-6) and np.all(x >= 0.0)


CELLS = [(5.0, 120.0, x_of([("H", 1.0)])), (20.0, 200.0, x_of([("He", 0.6), ("C", 0.4)])),
         (1.0e-8, 400.0, x_of([("H", 1.0)])), (5.0, 20.0, x_of([("H", 0.5), ("He", 0.5)])),
         (50.0, 330.0, x_of([("Ne", 0.5), ("O", 0.3), ("Mg", 0.2)])), (50.0, 480.0, x_of([("Si", 0.6), ("S", 0.4)]))]


def test_cells_burn_alone():
    """Each cell of a world steps as the same cell alone (bit for bit); a vacuum and a cold cell are left as they
    were, nothing booked."""
    params = react.params()
    p = world(CELLS)
    before = p.copy()
    rep, nuc, *_ = oracle.react_step(PH, params, p, 2.0)
    for i, c in enumerate(CELLS):
        q = world([c])
        _, nq, *_ = oracle.react_step(PH, params, q, 2.0)
        assert np.array_equal(p[:, 0, i].view(np.uint64), q[:, 0, 0].view(np.uint64)), i
        assert np.array_equal(nuc[:, 0, i], nq[:, 0, 0]), i
    for i in (2, 3):
        assert np.array_equal(p[:, 0, i].view(np.uint64), before[:, 0, i].view(np.uint64)) and not nuc[:, 0, i].any()
    assert nuc[:, 0, 0].any()


def n_fe_march(a, sigma, sigma_n, a_y, y_fe, x_fe, dt):
    """dX_Fe/dt = -A X_Fe max(0, (a_Y + y_Fe X_Fe) Sigma / Sigma_N - 1), classical RK4 at 4000 steps (burn.rs:521-536)."""
    rate = lambda x: -a * x * max((a_y + y_fe * x) * sigma / sigma_n - 1.0, 0.0)
    h = dt / 4000.0
    x = x_fe
    for _ in range(4000):
        k1 = rate(x)
        k2 = rate(x + 0.5 * h * k1)
        k3 = rate(x + 0.5 * h * k2)
        k4 = rate(x + h * k3)
        x += h / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    return x


@pytest.mark.parametrize("label,sigma,parts", [
    ("iron block", 400.0, [("Fe", 1.0)]),
    ("a_Y = b", 400.0, [("He", 0.8), ("Fe", 0.2)]),
    ("gate closes", 390.0, [("He", 0.8), ("Fe", 0.2)]),
    ("iron runs out", 410.0, [("He", 0.8), ("Fe", 0.2)]),
])
def test_n_fe_steps_in_closed_form(label, sigma, parts):
    params = react.params()
    a, sigma_n, q, y_fe, dt = 100.0, 160.0, -1.14 * Q_H, 26.0 / 56.0, 5.87e-3
    n_group = react.GROUPS.index("N")
    x0 = x_of(parts)
    others = x0.copy()
    others[SP.index("Fe")] = 0.0
    a_y = float(np.sum(others * Z / A))
    p = world([(sigma, 1.0, x0)])
    want_fe = frac(p, "Fe")
    for step in range(1, 21):
        e0, fe0, n0 = p[3, 0, 0], frac(p, "Fe"), frac(p, "n")
        rep, *_ = oracle.react_step(PH, params, p, dt)
        assert rep.substeps == 0 and rep.latch, (label, step)
        want_fe = n_fe_march(a, sigma, sigma_n, a_y, y_fe, want_fe, dt)
        fe, xn = frac(p, "Fe"), frac(p, "n")
        assert abs(xn - (x0[SP.index("Fe")] - want_fe)) <= 1e-9, (label, step)
        dn = xn - n0
        assert abs((fe0 - fe) - dn) <= 1e-15
        assert abs(p[4:, 0, 0].sum() - 1.0) <= 1e-12 and np.all(p[4:, 0, 0] >= 0.0)
        booked, want = rep.nuclear[n_group], q * dn * sigma
        assert abs(booked - want) <= 1e-12 * abs(q * sigma)
        assert all(rep.nuclear[g] == 0.0 for g in range(7) if g != n_group)
        assert abs((p[3, 0, 0] - e0) - booked) <= 1e-12 * abs(q * sigma)


@pytest.mark.parametrize("label,sigma,parts", [("gate", 400.0, [("Fe", 1.0)]), ("iron", 410.0, [("He", 0.8), ("Fe", 0.2)])])
def test_n_fe_lands_on_its_limits(label, sigma, parts):
    params = react.params()
    p = world([(sigma, 1.0, x_of(parts))])
    rep, *_ = oracle.react_step(PH, params, p, 1e6)
    assert rep.latch and rep.substeps == 0 and rep.unfinished == 0
    y_e_sigma = float(np.sum(p[4:, 0, 0] * Z / A)) * sigma
    if label == "gate":
        assert y_e_sigma <= 160.0
        assert abs(frac(p, "n") - (1.0 - 160.0 / (26.0 / 56.0 * sigma))) <= 1e-12
    else:
        assert frac(p, "Fe") == 0.0
    assert abs(p[4:, 0, 0].sum() - 1.0) <= 1e-12
    before = p.copy()
    again, *_ = oracle.react_step(PH, params, p, 1.0)
    assert not again.latch and np.array_equal(p, before)


def test_the_ignition_order():
    """G-ORDER (burn.rs:1040-1105): a cell of equal parts H, He, C, O, Ne, Si, T ramped from 40 by 0.5 per sample, one
    pass of 1e-9 each (T pinned): each group's specific power first passes 1e-3 Q_H in the order H, He, C, {Ne, O}, Si;
    H's power at its crossing is q A Sigma X^2 f(T) by hand; N never."""
    params = react.params(pin_temperature=True)
    eps_star = 1e-3 * Q_H
    eq = x_of([(k, 1.0 / 6.0) for k in ("H", "He", "C", "O", "Ne", "Si")])
    first = [None] * 7

    def powers(t):
        p = world([(1.0, t, eq)])
        rep, *_ = oracle.react_step(PH, params, p, 1e-9)
        return [rep.nuclear[g] / 1e-9 for g in range(7)]

    for k in range(int((600.0 - 40.0) / (100.0 * 0.005)) + 1):
        t = 40.0 + 100.0 * k * 0.005
        for g, pw in enumerate(powers(t)):
            if first[g] is None and pw >= eps_star:
                first[g] = t
    at = {g: first[react.GROUPS.index(g)] for g in ("H", "He", "C", "Ne", "O", "Si")}
    assert at["H"] < at["He"] < at["C"] < min(at["Ne"], at["O"]) and max(at["Ne"], at["O"]) < at["Si"]
    got, want = powers(at["H"])[0], Q_H * 5e-4 * (1.0 / 36.0) * f_h(at["H"], 56.0)
    assert abs(got - want) <= 1e-6 * want
    assert at["H"] >= 28.0 and at["He"] >= 50.0 and first[react.GROUPS.index("N")] is None


@pytest.mark.parametrize("rid", react.IDS)
def test_each_record_books_into_its_group(rid):
    rec = next(r for r in react.RECORDS if r[0] == rid)
    _, group, ins, _, q_ratio, _, _, terms, _, gate, _ = rec
    params = react.params(only=[rid], pin_temperature=True)
    sigma = 600.0 if gate is not None else 2.0
    t = 1.2 * params.rec[0].t_k[0] if terms else 1.0
    p = world([(sigma, t, x_of(ins))])
    before = p[4:, 0, 0].copy()
    i0 = SP.index(ins[0][0])
    rep, *_ = oracle.react_step(PH, params, p, 1e-3)
    moved = sigma * (before[i0] - p[4 + i0, 0, 0]) / ins[0][1]
    g = react.GROUPS.index(group)
    want = q_ratio * Q_H * moved
    assert moved > 0.0 and abs(rep.nuclear[g] - want) <= 1e-5 * abs(want)
    assert all(rep.nuclear[o] == 0.0 for o in range(7) if o != g)


@pytest.mark.parametrize("rid", ["H_burn", "He_burn", "C_burn", "Si_burn"])
def test_the_rate_table_holds_its_budget(rid):
    """f through the 256-node log-log table against the direct form within 1e-4 relative (rates.rs's budget), at the
    nodes' midpoints over the span and past it; 0 below T_thr."""
    params = react.params(only=[rid])
    rec = params.rec[0]
    x = np.ones(NS)
    t_thr = rec.t_thr
    for i in range(255):
        t = t_thr * 2.5 ** ((i + 0.5) / 255.0)
        a, b = oracle.react_omega(rec, 1.0, x, t, 0.5), oracle.react_omega(rec, 1.0, x, t, 0.5, table=False)
        assert abs(a - b) <= 1e-4 * b, (rid, t)
    for t in (3.0 * t_thr, 10.0 * t_thr):
        assert oracle.react_omega(rec, 1.0, x, t, 0.5) == oracle.react_omega(rec, 1.0, x, t, 0.5, table=False)
    assert oracle.react_omega(rec, 1.0, x, 0.999 * t_thr, 0.5) == 0.0


def test_the_twin_runs_p6_between_p5_and_p8():
    """The twin's step with P6 (mod.rs:263-297) on a burning field with gravity, hydro and heat off: P1's dt, P6 over
    it, P8's floors, the next dt -- the same f64 values as the sequence done by hand; with the records off the step is
    P8 alone."""
    rng = np.random.default_rng(5)
    w = h = 8
    planes = np.zeros((14, h, w))
    planes[0] = rng.uniform(5.0, 60.0, (h, w))
    x = rng.dirichlet(np.ones(NS), size=(h, w)).transpose(2, 0, 1)
    x[9] = 0.0
    x /= x.sum(axis=0)
    planes[4:] = x
    inv = np.einsum("s,shw->hw", (Z + 1.0) / A, x)
    planes[3] = planes[0] * rng.uniform(100.0, 400.0, (h, w)) * inv
    params = react.params()
    tw = oracle.Twin(PH, planes, gravity=False, hydro=False, heat=False, reactions=params)
    dt = tw.times()[1]
    tw.step()
    by_hand = planes.copy()
    assert dt == oracle.hydro_dt(PH, by_hand)
    rep, *_ = oracle.react_step(PH, params, by_hand, dt)
    by_hand, _, _ = oracle.floors_apply(PH, by_hand)
    assert np.array_equal(tw.state(), by_hand)
    assert tw.times()[1] == oracle.hydro_dt(PH, by_hand)
    got = tw.last_react()
    assert got.substeps == rep.substeps > 0 and list(got.nuclear) == list(rep.nuclear)
    off = oracle.Twin(PH, planes, gravity=False, hydro=False, heat=False)
    off.step()
    assert np.array_equal(off.state(), oracle.floors_apply(PH, planes)[0])
