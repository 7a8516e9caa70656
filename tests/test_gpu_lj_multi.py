"""lj_multi_force (per-type-pair LJ parameters) through the C ABI against the oracle.

Two species with three different pair parameter sets (one pair at epsilon 0, the
cutoffs 5.9-8 A): forces within 1e-11 of rms|F| in FP64 and 1e-5 in the mixed form,
energies 1e-12, virials 1e-12, on the block list and on the global list; with equal
parameters for every pair the operator gives the single-parameter operator's
results bit for bit; the reference's multi-species LJ regression deck runs and
follows the oracle's NVE run (1e-9)."""
import numpy as np
import pytest

from conftest import GOLDEN
from test_gpu_parity import frel, make_oracle, make_sim

pytestmark = pytest.mark.gpu

MASSES = [65.38, 63.546]
EV = 9648.533212331


def _pairs():
    from exastamp_amd import units
    J, m = units.parse_quantity("1 J"), units.parse_quantity("1 m")
    return [{"type_a": "Zn", "type_b": "Zn", "rcut": "6.10 ang",
             "parameters": {"epsilon": 2.522e-20 * J, "sigma": 0.244e-9 * m}},
            {"type_a": "Cu", "type_b": "Zn", "rcut": "5.89 ang",
             "parameters": {"epsilon": 4.853e-20 * J, "sigma": 0.236e-9 * m}}]


def _tables(names, pairs, common=(0.0, 0.0), rc=8.0):
    from exastamp_amd.units import parse_quantity
    nt = len(names)
    n = nt * (nt + 1) // 2
    eps, sig, rcs = np.full(n, common[0]), np.full(n, common[1]), np.full(n, rc)
    for e in pairs:
        a, b = names.index(e["type_a"]), names.index(e["type_b"])
        pid = max(a, b) * (max(a, b) + 1) // 2 + min(a, b)
        eps[pid] = parse_quantity(e["parameters"]["epsilon"])
        sig[pid] = parse_quantity(e["parameters"]["sigma"])
        rcs[pid] = parse_quantity(e["rcut"])
    return eps, sig, rcs


def _config(seed=3):
    from exastamp_amd import inputs
    cfg = inputs.lattice_config((8, 8, 8), 5.0, MASSES, T=300.0, jitter=0.15, seed=seed)
    cfg.type[:] = inputs.rng(seed + 1).integers(0, 2, size=cfg.n)
    return cfg


@pytest.mark.parametrize("kind", ["block", "global"])
@pytest.mark.parametrize("mixed", [False, True])
def test_lj_multi_matches_oracle(monkeypatch, oracle_mod, kind, mixed):
    if kind == "global":
        monkeypatch.setenv("XSMD_LIST", "global")
    else:
        monkeypatch.delenv("XSMD_LIST", raising=False)
    cfg = _config()
    names = ["Zn", "Cu"]
    sim = make_sim(cfg, MASSES, names)
    sim.chunk_neighbors(rcut_max=8.0, rcut_inc=0.3)
    assert sim.list_info()["kind"] == kind
    sim.zero_force_energy()
    sim.lj_multi_force(rcut="8 ang", parameters=_pairs(), common_parameters={"epsilon": 0.0, "sigma": 0.0},
                       trigger_thermo_state=True, compute_virial=True, mixed=mixed)
    sim.force_to_accel()
    o = make_oracle(oracle_mod, cfg, MASSES)
    o.set_lj_multi(2, *_tables(names, _pairs()))
    o.build_neighbors(8.0, 0.3)
    o.compute_forces("lj_multi")
    ref = o.state()
    assert frel(sim.forces(), o.forces()) <= (1e-5 if mixed else 1e-11)
    assert abs(sim.ep().sum() - ref["ep"].sum()) <= (1e-6 if mixed else 1e-12) * abs(ref["ep"].sum())
    if not mixed:
        vir, vref = sim.virial().sum(0), ref["virial"].sum(0)
        assert np.abs(vir - vref).max() <= 1e-12 * np.abs(vref).max()


def test_lj_multi_with_equal_pairs_is_lj():
    cfg = _config(seed=9)
    sim = make_sim(cfg, MASSES, ["Zn", "Cu"])
    sim.chunk_neighbors(rcut_max=8.0, rcut_inc=0.3)
    p = {"epsilon": 0.0104 * EV, "sigma": 3.4}
    sim.zero_force_energy()
    with pytest.raises(ValueError, match="exactly 1 atom type"):
        sim.lj_compute_force(parameters=p, rcut=8.0)
    sim.lj_compute_force(parameters=p, rcut=8.0, trigger_thermo_state=True, compute_virial=True, type="Zn")
    a = (sim.accelerations(), sim.ep(), sim.virial())
    sim.zero_force_energy()
    sim.lj_multi_force(rcut=8.0, parameters=(), common_parameters=p, trigger_thermo_state=True,
                       compute_virial=True)
    b = (sim.accelerations(), sim.ep(), sim.virial())
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_lj_multi_species_deck_follows_oracle(oracle_mod):
    """data/regression_new/potentials/pair/lj/multi_species_nosym.msp (fixture): 2 048
    Cu/Zn atoms, two parameter sets and the Cu-Cu pair at epsilon 0, 10 NVE steps,
    against the oracle's run."""
    from exastamp_amd import deck, units
    cfg = deck.load_deck(GOLDEN / "decks" / "lj" / "multi_species_nosym.msp", [GOLDEN],
                         [(["global", "simulation_thermostate_screen_frequency"], "10")])
    r = deck.Runner(cfg, out=open("/dev/null", "w"))
    recs = r.run()
    assert [x.step for x in recs] == [0, 10]
    names = [n for n, _ in r.species]
    o = oracle_mod.Oracle()
    o.set_config(r.init_atoms, [m for _, m in r.species])
    pairs = cfg["lj_multi_force"]["parameters"]
    o.set_lj_multi(len(names), *_tables(names, pairs, rc=8.0))
    _, th = o.md_run("lj_multi", 10, 0.001, 8.0, 1.0)
    for rec, row in ((recs[0], th[0]), (recs[1], th[10])):
        pe = rec.potential_ev * units.EV * rec.particles
        ke = rec.kinetic_ev * units.EV * rec.particles
        assert abs(pe - row[24]) <= 1e-9 * abs(row[24])
        assert abs(ke - row[21:24].sum()) <= 1e-9 * row[21:24].sum()
