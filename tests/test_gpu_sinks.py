"""P3 sinks on the GPU (sandbox_reactions_amd/csrc/sinks.hip) through the C-ABI (sr_set_sinks, sr_put_sinks,
sr_read_sinks, sr_sinks_apply), graded as the reference grades its shader pass (crates/sr-engine/tests/gpu/sink.rs:
760-1402): against the oracle's P3 from the same f32 start, and on the conservation itself. Accretion keeps grid +
sink mass and momentum and touches nothing beyond r_acc; boundness in the sink's frame; two sinks over one cell take it
once; formation where all four criteria hold and each refusal alone; the densest first, ties to the lowest y then x,
one a step, none past 16; a sink from mass; the kick-drift-kick and merging against the oracle in the gas; no net force
between the sinks and the gas; the sinks' part of P2's field; the cloud falling into a sink inside the world's step
with grid + sink mass kept every step; the step with P3 reproducible and, with P3 off, the step without it."""
import math

import numpy as np
import pytest

import oracle
from test_oracle_sinks import M_TOV, SHIPPED_COLD, SIGMA_FLOOR, blob, cloud, core_world, vacuum, well_world

pytestmark = pytest.mark.gpu


def rig(sr, planes, sinks=(), m_tov=float("inf"), gravity=False, hydro=False, edge=None, **phys):
    _, h, w = planes.shape
    kw = dict(phys)
    if edge is not None:
        kw["edge"] = edge
    world = sr.World(w, h, physics=sr.physics(**kw), gravity=gravity, hydro=hydro, heat=False)
    world.set_sinks(6000.0, m_tov)
    world.upload(planes.astype(np.float32))
    if sinks:
        world.put_sinks(sinks)
    return world


def twin(planes, sinks=(), m_tov=float("inf"), **phys):
    k = oracle.Sinks(oracle.reference_physics(**phys), m_tov=m_tov)
    for s in sinks:
        k.push(*s)
    return k, np.ascontiguousarray(planes.astype(np.float32).astype(np.float64))


def pass_on_twin(k, st, phi, dt):
    """The oracle's P3 with phi as given (no sinks' term: the rig uploads phi by hand)."""
    formed = k.form(st, phi)
    acc = k.accrete(st)
    k.advance(st, dt)
    merged = k.merge()
    return formed, acc, merged


def same_sinks(gpu, k, tol):
    assert len(gpu) == len(k.sinks)
    for a, b in zip(gpu, k.sinks):
        for u, v in zip(a, b):
            assert abs(u - v) <= tol * max(1.0, abs(v)), (a, b)


def test_accretion_matches_the_oracle_and_touches_nothing_beyond_r_acc(gpu_lib):
    sr = gpu_lib
    print("P3 accretion on the cloud", flush=True)
    st = cloud()
    s0 = [(32.5, 32.5, 0.0, 0.0, 2000.0)]
    w = rig(sr, st, s0)
    m0 = w.readback_sigma().sum()
    p0 = w.readback()[1:3].astype(np.float64).sum(axis=(1, 2))
    # dt 0: accretion alone (the kick-drift-kick moves the sinks by the gas's pull, whose reaction is P4's)
    w.sinks_apply(0.0, np.zeros((64, 64), np.float32))
    sinks, rep = w.read_sinks()
    out = w.readback().astype(np.float64)
    k, t = twin(st, s0)
    _, acc, _ = pass_on_twin(k, t, np.zeros((64, 64)), 0.0)
    print(f"P3 accreted {rep.accreted:.6f} (oracle {acc:.6f}), swallowed {rep.swallowed:.6f}", flush=True)
    assert rep.accreted > 10.0 and abs(rep.accreted - acc) <= 1e-9 * acc and rep.formed_x == -1
    assert abs(rep.swallowed - k.swallowed) <= 1e-6 * k.swallowed
    same_sinks(sinks, k, 1e-9)
    m1 = w.readback_sigma().sum()
    p1 = out[1:3].sum(axis=(1, 2))
    sm = sum(s[4] for s in sinks)
    sp = [sum(s[4] * s[2] for s in sinks), sum(s[4] * s[3] for s in sinks)]
    assert abs(m1 + sm - (m0 + 2000.0)) <= 1e-12 * (m0 + 2000.0)
    assert abs(p1[0] + sp[0] - p0[0]) <= 1e-6 * abs(p0[0]) and abs(p1[1] + sp[1] - p0[1]) <= 1e-6 * abs(p0[1])
    ra = 2.0
    yy, xx = np.mgrid[0:64, 0:64]
    far = np.hypot(xx + 0.5 - 32.5, yy + 0.5 - 32.5) > ra + 1e-9
    assert np.array_equal(out[:, far].astype(np.float32), st.astype(np.float32)[:, far])
    assert np.abs(out[:4] - t[:4]).max() <= 1e-6 * np.abs(t[:4]).max()


def test_accretion_is_bound_gas_only_in_the_sinks_frame(gpu_lib):
    sr = gpu_lib
    hot = cloud()
    hot[3] *= 1e5
    w = rig(sr, hot, [(32.5, 32.5, 0.0, 0.0, 2000.0)])
    before = w.readback()
    w.sinks_apply(0.0, np.zeros((64, 64), np.float32))
    sinks, rep = w.read_sinks()
    assert rep.accreted == 0.0 and rep.swallowed == 0.0 and sinks[0][4] == 2000.0
    assert np.array_equal(w.readback(), before)
    fast = cloud()
    fast[1] *= 300.0
    fast[2] *= 300.0
    fast[3] += 0.5 * (fast[1] ** 2 + fast[2] ** 2) / fast[0]
    w = rig(sr, fast, [(32.5, 32.5, 0.0, 0.0, 2000.0)])
    w.sinks_apply(0.0, np.zeros((64, 64), np.float32))
    assert w.read_sinks()[1].accreted == 0.0
    w = rig(sr, fast, [(32.5, 32.5, 90.0, -60.0, 2000.0)])
    w.sinks_apply(0.0, np.zeros((64, 64), np.float32))
    assert w.read_sinks()[1].accreted > 10.0


def test_two_sinks_over_one_cell_take_it_once(gpu_lib):
    sr = gpu_lib
    st = cloud()
    s0 = [(32.5, 32.5, 0.0, 0.0, 2000.0), (33.5, 32.5, 0.0, 0.0, 2000.0)]
    w = rig(sr, st, s0)
    m0 = w.readback_sigma().sum()
    w.sinks_apply(0.0, np.zeros((64, 64), np.float32))
    sinks, rep = w.read_sinks()
    k, t = twin(st, s0)
    _, acc, merged = pass_on_twin(k, t, np.zeros((64, 64)), 0.0)
    # 1 cell apart: the pass merges them; the mass each took is the oracle's, and the grid's loss is what they gained
    assert rep.merged == merged == 1 and abs(rep.accreted - acc) <= 1e-9 * acc
    same_sinks(sinks, k, 1e-9)
    assert abs(w.readback_sigma().sum() + sinks[0][4] - (m0 + 4000.0)) <= 1e-12 * m0


def test_formation_matches_the_oracle_and_refuses_each_criterion(gpu_lib):
    sr = gpu_lib
    print("P3 formation", flush=True)
    cases = [((8, 8, 7000.0), -1.0, 0.0, (8, 8)), ((8, 8, 6000.0 * (1 - 1e-6)), -1.0, 0.0, None),
             ((8, 8, 6000.0), -1.0, 0.0, (8, 8)), ((8, 8, 7000.0), 1.0, 0.0, None),
             ((8, 8, 7000.0), -1.0, 200.0, None), ((8, 8, 7000.0), -1.0, 100.0 - 0.1 + 0.01, None),
             ((8, 8, 7000.0), -1.0, 100.0 - 0.1 - 0.01, (8, 8))]
    for well, sign, off, want in cases:
        st, phi = well_world(16, 16, [well], sign, off)
        w = rig(sr, st)
        before = w.readback()
        w.sinks_apply(0.0, phi.astype(np.float32))
        sinks, rep = w.read_sinks()
        got = None if rep.formed_x < 0 else (rep.formed_x, rep.formed_y)
        assert got == want, (well, sign, off, got)
        if want is None:
            assert not sinks and rep.swallowed == 0.0 and np.array_equal(w.readback(), before)
        else:  # the new sink accretes in the same pass, as the oracle's
            k, t = twin(st)
            assert pass_on_twin(k, t, phi, 0.0)[0] == want
            same_sinks(sinks, k, 1e-9)
            assert abs(rep.swallowed - k.swallowed) <= 1e-9 * k.swallowed
    st, phi = well_world(16, 16, [(8, 8, 7000.0)], -1.0, 0.0)
    phi[7, 9] = phi[8, 8] - 1.0
    w = rig(sr, st)
    w.sinks_apply(0.0, phi.astype(np.float32))
    assert w.read_sinks()[1].formed_x == -1


def test_the_densest_forms_first_ties_to_the_lowest_y_then_x_one_a_step_none_past_16(gpu_lib):
    sr = gpu_lib
    for wells, want in (([(18, 4, 7000.0), (5, 11, 7000.0)], [(18, 4), (5, 11), None]),
                        ([(18, 4, 7000.0), (5, 4, 7000.0)], [(5, 4)]),
                        ([(5, 4, 7000.0), (18, 11, 8000.0)], [(18, 11)])):
        st, phi = well_world(24, 16, wells, -1.0, 0.0)
        w = rig(sr, st)
        for expect in want:
            w.sinks_apply(0.0, phi.astype(np.float32))
            rep = w.read_sinks()[1]
            assert (None if rep.formed_x < 0 else (rep.formed_x, rep.formed_y)) == expect, wells
    st, phi = well_world(16, 16, [(8, 8, 7000.0)], -1.0, 0.0)
    w = rig(sr, st, [(100.0 + 10.0 * q, 1.0, 0.0, 0.0, 1.0) for q in range(16)])
    w.sinks_apply(0.0, phi.astype(np.float32))
    assert w.read_sinks()[1].formed_x == -1
    w = rig(sr, st, [(100.0 + 10.0 * q, 1.0, 0.0, 0.0, 1.0) for q in range(15)])
    w.sinks_apply(0.0, phi.astype(np.float32))
    sinks, rep = w.read_sinks()
    assert (rep.formed_x, rep.formed_y) == (8, 8) and rep.count == 16


def test_a_sink_from_mass_as_the_oracle(gpu_lib):
    sr = gpu_lib
    print("P3 sink from mass", flush=True)
    for mass, species, sign, y_e, forms in ((1.05, "n", -1.0, None, True), (0.95, "n", -1.0, None, False),
                                            (1.05, "Fe", -1.0, None, False), (1.05, "n", 1.0, None, False),
                                            (1.05, "Fe", -1.0, 0.38, True), (1.05, "Fe", -1.0, 0.42, False)):
        st, phi = core_world(24, 24, (12, 12), mass * M_TOV, species, sign, y_e)
        w = rig(sr, st, m_tov=M_TOV, **SHIPPED_COLD)
        w.sinks_apply(0.0, phi.astype(np.float32))
        sinks, rep = w.read_sinks()
        assert (rep.formed_x == 12) == forms, (mass, species, sign, y_e)
        if forms:
            k, t = twin(st, m_tov=M_TOV, **SHIPPED_COLD)
            assert pass_on_twin(k, t, phi, 0.0)[0] == (12, 12)
            same_sinks(sinks, k, 1e-9)
            out = w.readback().astype(np.float64)
            assert np.abs(out[:4] - t[:4]).max() <= 1e-6 * np.abs(t[:4]).max()


def test_kick_drift_kick_and_merging_match_the_oracle(gpu_lib):
    sr = gpu_lib
    print("P3 motion", flush=True)
    st = blob()
    s0 = [(20.5, 38.5, 0.3, 0.1, 50.0), (40.5, 30.5, 0.0, -0.1, 400.0), (44.0, 30.5, 0.0, 0.0, 300.0)]
    w = rig(sr, st, s0)
    k, t = twin(st, s0)
    for step in range(6):
        w.sinks_apply(0.5, np.zeros((64, 64), np.float32))
        pass_on_twin(k, t, np.zeros((64, 64)), 0.5)
        sinks, rep = w.read_sinks()
        same_sinks(sinks, k, 1e-10)
        assert rep.count == len(k.sinks)


def test_no_net_force_between_the_sinks_and_the_gas(gpu_lib):
    """The sinks' pull on the cells (P2's part of the field, read back) against the gas's pull on the sinks (the
    velocity the closing half-kick leaves over a tiny dt with the drift negligible): equal and opposite within the
    reference's 1e-5 of the gross."""
    sr = gpu_lib
    st = blob()
    s0 = [(17.5, 41.5, 0.0, 0.0, 100.0), (24.5, 30.5, 0.0, 0.0, 400.0)]
    w = rig(sr, st, s0, gravity=True)
    dt = 1e-7
    w.sinks_apply(dt, np.zeros((64, 64), np.float32))
    sinks, _ = w.read_sinks()
    # the gas's pull on each sink from its velocity change (pair pulls removed: written here)
    net, gross = [0.0, 0.0], 0.0
    for i, s in enumerate(sinks):
        ax, ay = s[2] / dt, s[3] / dt
        for j, o in enumerate(s0):
            if j == i:
                continue
            dx, dy = o[0] - s0[i][0], o[1] - s0[i][1]
            soft = dx * dx + dy * dy + 4.0
            ax -= o[4] * dx / soft ** 1.5
            ay -= o[4] * dy / soft ** 1.5
        net[0] += s[4] * ax
        net[1] += s[4] * ay
        gross += s[4] * math.hypot(ax, ay)
    k = oracle.Sinks(oracle.reference_physics())
    for s in s0:
        k.push(*s)
    gx, gy = np.zeros((64, 64)), np.zeros((64, 64))
    k.add_field(None, gx, gy)
    m = np.where(st[0] >= 1e-6, st[0], 0.0)
    net[0] += float((m * gx).sum())
    net[1] += float((m * gy).sum())
    gross += float((m * np.hypot(gx, gy)).sum())
    print(f"P3 no net force: {math.hypot(*net) / gross:.3e}", flush=True)
    assert math.hypot(*net) <= 1e-5 * gross


def test_p2_carries_the_sinks_part_of_the_field(gpu_lib):
    """One step with gravity and P3 (hydro off): phi and g read back are the gas's field plus the sinks' softened
    potential and exact pull at the step's start (the oracle's field and add_field), within f32 and the FFT's error."""
    sr = gpu_lib
    st = blob()
    s0 = [(30.5, 20.5, 0.0, 0.0, 500.0)]
    w = rig(sr, st, s0, gravity=True)
    w.run(1)
    phi, gx = w.read_field(sr.FIELD_PHI), w.read_field(sr.FIELD_GX)
    ph = oracle.reference_physics()
    ophi, ogx, ogy = oracle.gravity_field(ph, st[0])
    k = oracle.Sinks(ph)
    for s in s0:
        k.push(*s)
    k.add_field(ophi, ogx, ogy)
    assert np.abs(phi - ophi).max() <= 1e-5 * np.abs(ophi).max()
    assert np.abs(gx - ogx).max() <= 1e-4 * np.abs(ogx).max()


def test_the_cloud_falls_into_a_sink_in_the_step_with_mass_kept(gpu_lib):
    sr = gpu_lib
    print("P3 cloud fall in the step", flush=True)
    st = cloud()
    w = rig(sr, st, [(32.5, 32.5, 0.0, 0.0, 2000.0)], gravity=True, hydro=True, edge=sr.EDGE_BOUNCE)
    total = w.readback_sigma().sum() + 2000.0
    last = 2000.0
    for step in range(40):
        w.run(1)
        sinks, rep = w.read_sinks()
        assert rep.formed_x == -1 and rep.merged == 0 and len(sinks) == 1, step
        now = w.readback_sigma().sum() + sinks[0][4]
        # P8 refills a taken cell to the floor (Sigma_floor of 1e-8 per cell): within that, mass is kept
        assert abs(now - total) <= 1e-6 * total, (step, now, total)
        assert sinks[0][4] >= last
        last = sinks[0][4]
    assert last > 2000.0
    assert not w.read_dt().non_finite


def test_the_step_with_p3_is_reproducible_and_p3_off_is_the_plain_step(gpu_lib):
    sr = gpu_lib
    st = cloud()
    outs = []
    for _ in range(2):
        w = rig(sr, st, [(32.5, 32.5, 0.0, 0.0, 2000.0)], gravity=True, hydro=True)
        w.run(3)
        outs.append((w.readback(), w.read_sinks()[0]))
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32)) and outs[0][1] == outs[1][1]
    a = sr.World(64, 64, physics=sr.physics(), gravity=True, hydro=True, heat=False)
    a.upload(st.astype(np.float32))
    a.run(3)
    b = rig(sr, st, gravity=True, hydro=True)
    b.set_sinks(on=False)
    b.upload(st.astype(np.float32))
    b.run(3)
    assert np.array_equal(a.readback().view(np.uint32), b.readback().view(np.uint32))
    with pytest.raises(sr.SrError):
        sr.World(16, 16, physics=sr.physics(dx=0.5), gravity=True).set_sinks()
