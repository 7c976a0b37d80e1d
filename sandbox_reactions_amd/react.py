"""P6 burning's registry: the shipped reaction records (assets/reactions.json) and the constants the pass reads
(assets/physics.json), restated as data, resolved into the sr_react_params the library takes. The keys a record names
(its temperatures T_k, its A, its nu, the gate's Sigma_g) resolve against the physics values, overridable as a scene's
`overrides` or the reference's tests override them (crates/sr-physics/src/registry.rs:485-535, :680-753; the rate law
crates/sr-physics/src/rates.rs:1-60)."""
from __future__ import annotations

from ._lib import REACT_GROUPS, REACT_RECORDS, ReactParams

SPECIES = ("H", "He", "C", "O", "Ne", "Mg", "Si", "S", "Fe", "n")  # assets/elements.json order
Z = (1, 2, 6, 8, 10, 12, 14, 16, 26, 0)
GROUPS = ("H", "He", "C", "Ne", "O", "Si", "N")  # registry.rs:460
assert len(GROUPS) == REACT_GROUPS

# assets/reactions.json's records in file order: id, group, inputs [(species, share)], outputs, q_ratio, A's exponent a,
# orders, terms [(temperature key, factor, nu or its key, w)], t_thr_factor, gate (Sigma key, p) or None, A's key
RECORDS = (
    ("H_burn", "H", (("H", 1.0),), (("He", 1.0),), 1.0, 1, (2,), (("t_h", 1.0, 4, 1.0), ("t_h", 1.6, 18, 1.0)), 0.5,
     None, "a_h_burn"),
    ("He_burn", "He", (("He", 1.0),), (("C", 1.0),), 0.09259259259259259, 2, (3,), (("t_he", 1.0, 40, 1.0),), 0.5,
     None, "a_he_burn"),
    ("C_alpha", "He", (("C", 0.75), ("He", 0.25)), (("O", 1.0),), 0.0684931506849315, 1, (1, 1),
     (("t_he", 1.0, 40, 1.0),), 0.5, None, "a_c_alpha"),
    ("C_burn", "C", (("C", 1.0),), (("Ne", 0.8), ("Mg", 0.2)), 0.06369426751592357, 1, (2,),
     (("t_c", 1.0, "nu_c_burn", 1.0),), 0.5, None, "a_c_burn"),
    ("Ne_burn", "Ne", (("Ne", 1.0),), (("O", 0.4), ("Mg", 0.6)), 0.017543859649122806, 1, (1,),
     (("t_ne", 1.0, "nu_ne_burn", 1.0),), 0.5, None, "a_ne_burn"),
    ("O_burn", "O", (("O", 1.0),), (("Si", 0.7), ("S", 0.3)), 0.07936507936507936, 1, (2,),
     (("t_o", 1.0, "nu_o_burn", 1.0),), 0.5, None, "a_o_burn"),
    ("Mg_burn", "O", (("Mg", 1.0),), (("Si", 1.0),), 0.07936507936507936, 1, (1,),
     (("t_o", 1.0, "nu_mg_burn", 1.0),), 0.5, None, "a_mg_burn"),
    ("Si_burn", "Si", (("Si", 1.0),), (("Fe", 1.0),), 0.030303030303030304, 1, (1,),
     (("t_si", 1.0, "nu_si_burn", 1.0),), 0.5, None, "a_si_burn"),
    ("S_burn", "Si", (("S", 1.0),), (("Fe", 1.0),), 0.030303030303030304, 1, (1,),
     (("t_si", 1.0, "nu_s_burn", 1.0),), 0.5, None, "a_s_burn"),
    ("N_Fe", "N", (("Fe", 1.0),), (("n", 1.0),), -1.14, 0, (1,), (), None, ("sigma_n", 1), "a_n_fe"),
)
IDS = tuple(r[0] for r in RECORDS)

# the physics.json values the records and the pass read
CONSTANTS = {
    "t_h": 56.0, "t_he": 100.0, "t_c": 109.8, "t_ne": 118.0, "t_o": 122.0, "t_si": 135.0,
    "a_h_burn": 5e-4, "a_he_burn": 1e-3, "a_c_alpha": 1e-3, "a_c_burn": 1e-3, "a_ne_burn": 1e-3, "a_o_burn": 1e-3,
    "a_mg_burn": 1e-3, "a_si_burn": 1e-3, "a_s_burn": 1e-3, "a_n_fe": 100.0,
    "nu_c_burn": 30.0, "nu_ne_burn": 30.0, "nu_o_burn": 30.0, "nu_mg_burn": 30.0, "nu_si_burn": 30.0,
    "nu_s_burn": 30.0, "sigma_n": 160.0, "q_h": 8700.0,
    "burn_dx": 0.1, "burn_dt": 0.05, "burn_k": 15.0, "a_nu": 1e-3, "t_nu": 93.0, "m_nu": 8.0,
}


def params(only=None, overrides=None, pin_temperature=False, neutrinos=False) -> ReactParams:
    """The pass's parameters: the records of `only` (ids; all when None) enabled in file order, every key resolved
    against CONSTANTS updated by `overrides`; N_Fe stepped in closed form when it is enabled and of that shape (one
    input of order 1 into a species of Z = 0, no temperature law, the gate's p = 1; react.rs:197-219)."""
    c = dict(CONSTANTS)
    for k, v in (overrides or {}).items():
        if k in c:
            c[k] = float(v)
    val = lambda v: c[v] if isinstance(v, str) else float(v)
    p = ReactParams()
    recs = [r for r in RECORDS if only is None or r[0] in only]
    if only is not None and len(recs) != len(set(only)):
        raise KeyError(f"unknown records in {sorted(set(only) - set(IDS))}")
    assert len(recs) <= REACT_RECORDS
    p.n_records = len(recs)
    p.nfe = -1
    for k, (rid, group, ins, outs, q_ratio, a, orders, terms, t_thr_factor, gate, a_key) in enumerate(recs):
        r = p.rec[k]
        r.n_in, r.n_out = len(ins), len(outs)
        for j, (sp, share) in enumerate(ins):
            r.in_species[j], r.in_share[j], r.in_order[j] = SPECIES.index(sp), share, float(orders[j])
        for j, (sp, share) in enumerate(outs):
            r.out_species[j], r.out_share[j] = SPECIES.index(sp), share
        r.group = GROUPS.index(group)
        r.q = q_ratio * c["q_h"]
        r.a_coef, r.a_sigma = c[a_key], float(a)
        r.n_terms = len(terms)
        for j, (t_key, factor, nu, w) in enumerate(terms):
            r.t_k[j], r.nu[j], r.w[j] = factor * c[t_key], val(nu), float(w)
        r.t_thr = t_thr_factor * r.t_k[0] if terms else 0.0
        r.gated = int(gate is not None)
        if gate is not None:
            r.sigma_g, r.gate_p = c[gate[0]], float(gate[1])
        r.latch = int(rid == "N_Fe")
        shaped = (not terms and tuple(orders) == (1,) and len(ins) == 1 and len(outs) == 1 and
                  Z[SPECIES.index(outs[0][0])] == 0 and gate is not None and float(gate[1]) == 1.0)
        if rid == "N_Fe" and shaped:
            p.nfe = k
    for k in ("burn_dx", "burn_dt", "burn_k", "a_nu", "t_nu", "m_nu"):
        setattr(p, k, c[k])
    p.neutrinos, p.pin_temperature = int(neutrinos), int(pin_temperature)
    return p
