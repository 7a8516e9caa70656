"""Deck loading (host logic, no GPU): the reference's regression decks under
tests/golden/decks (copied from data/regression_new/potentials/{pair/lj,eam/eam_alloy})
resolve their includes, override the default graph by node name, and take
--set overrides; list appends ('+name') and include cycles behave as stated."""
import pytest

from conftest import GOLDEN

DECKS = GOLDEN / "decks"


def test_lj_deck_resolves():
    from exastamp_amd import deck
    cfg = deck.load_deck(DECKS / "lj" / "single_specy_nosym.msp")
    assert cfg["compute_force"] == "lj_compute_force"
    assert cfg["lj_compute_force"]["rcut"] == "8.0 ang"
    assert cfg["global"]["max_iteration"] == 10          # from the included baseline
    assert cfg["global"]["dt"] == "1.0e-3 ps"
    assert cfg["global"]["md_loop_continue"] is True     # default graph value kept
    names = [next(iter(s)) if isinstance(s, dict) else s for s in cfg["setup_system"]]
    assert names == ["particle_types", "particle_type_add_properties", "domain", "init_rcb_grid", "lattice",
                     "gaussian_noise_r", "gaussian_noise_v"]


def test_eam_split_deck_blocks_and_sets():
    from exastamp_amd import deck
    cfg = deck.load_deck(DECKS / "eam_alloy" / "single_specy_nosym_cs1.msp",
                         sets=[(["global", "max_iteration"], "3"), (["chunk_neighbors", "config", "chunk_size"], "4")])
    assert cfg["global"]["max_iteration"] == 3
    assert cfg["chunk_neighbors"]["config"] == {"chunk_size": 4, "build_particle_offset": True}
    assert cfg["compute_force"]["rebind"] == {"parameters": "eam_alloy_parameters"}
    assert cfg["compute_force"]["body"] == ["eam_rho", "eam_rho2emb", "ghost_update_opt", "eam_force"]
    (op, p), = cfg["eam_rho"][0].items()
    assert op == "eam_alloy_force" and p["eam_rho"] is True and p["eam_force"] is False


def test_set_parsing_and_appends(tmp_path):
    from exastamp_amd import deck
    sets, rest = deck.parse_sets(["d.msp", "--set-global-max_iteration", "20", "--mixed"])
    assert sets == [(["global", "max_iteration"], "20")] and rest == ["d.msp", "--mixed"]
    base = tmp_path / "base.yaml"
    base.write_text("setup_system:\n  - a\n  - b\nglobal: { dt: 2.0e-3 ps }\n")
    top = tmp_path / "top.msp"
    top.write_text("includes: [ base.yaml ]\n+setup_system:\n  - c\nglobal: { max_iteration: 5 }\n")
    cfg = deck.load_deck(top)
    assert cfg["setup_system"] == ["a", "b", "c"]
    assert cfg["global"]["dt"] == "2.0e-3 ps" and cfg["global"]["max_iteration"] == 5


def test_include_cycle_and_missing(tmp_path):
    from exastamp_amd import deck
    (tmp_path / "a.msp").write_text("includes: [ b.msp ]\n")
    (tmp_path / "b.msp").write_text("includes: [ a.msp ]\n")
    with pytest.raises(deck.DeckError, match="cycle"):
        deck.load_deck(tmp_path / "a.msp")
    (tmp_path / "c.msp").write_text("includes: [ nowhere.yaml ]\n")
    with pytest.raises(deck.DeckError, match="cannot find"):
        deck.load_deck(tmp_path / "c.msp")


def _record_from_row(row):
    from exastamp_amd import deck
    v = row.split()
    return deck.ThermoRecord(step=int(v[0]), time_ps=float(v[1]), particles=int(v[2]), total_ev=float(v[3]),
                             kinetic_ev=float(v[4]), potential_ev=float(v[5]), temperature_k=float(v[6]),
                             t_xyz=(0.0, 0.0, 0.0), pressure_pa=0.0, p_full=tuple(float(x) for x in v[7:13]),
                             p_vir=(0.0, 0.0, 0.0), vonmises_pa=0.0, lengths=tuple(float(x) for x in v[13:16]),
                             volume=float(v[19]), mass=0.0, density=float(v[20]))


def test_thermo_file_rows_match_recorded_file():
    """The reference's recorded thermodynamic_state.csv (AlCu deck, fixture): its header
    and every row, re-printed from the parsed values, come out byte for byte."""
    from exastamp_amd import deck
    lines = (GOLDEN / "alcu_b2_thermodynamic_state.csv").read_text().split("\n")
    assert lines[0] == deck.THERMO_HEADER
    rows = [l for l in lines[1:] if l.strip()]
    assert len(rows) >= 2
    for row in rows:
        assert _record_from_row(row).line() == row


def test_screen_log_modes_and_status_column():
    """Screen log (print_thermodynamic_state): item lists per mode, the default list for
    an unknown mode, the moves character of the status column."""
    from exastamp_amd import deck
    row = (GOLDEN / "alcu_b2_thermodynamic_state.csv").read_text().split("\n")[1]
    rec = _record_from_row(row)
    for mode, items in deck.LOG_MODES.items():
        assert len(deck.log_line(rec, mode).split()) == len(items) - 1   # blank status column
        assert deck.log_header(mode).startswith("      Step")
    assert deck.log_header("stp;pht;vol") == deck.log_header("no such mode")
    assert deck.log_line(rec, "mechanical").split()[:3] == ["0", "0.000000e+00", "16000"]
    for moves, ch in ((0, None), (1, "m"), (3, "3"), (12, "M")):
        cols = deck.log_line(rec, "mechanical", moves).split()
        assert (len(cols) == 11) if ch is None else (cols[3] == ch)
    assert "13.672" in deck.log_line(rec, "default")


class _StubSim:
    """Stands in for the operator mirror (no GPU): records the operator calls and
    answers particle_displ_over from a script."""
    over_script = ()

    def __init__(self, device=0):
        import numpy as np
        self.calls = []
        self.n = 0
        self.xform = np.eye(3)
        self._over = list(type(self).over_script)

    def _log(self, name):
        self.calls.append(name)

    def species(self, spec):
        self._log("species")

    def set_config(self, cfg):
        self.n = cfg.n
        self._log("set_config")

    def chunk_neighbors(self, **kw):
        self._log("chunk_neighbors")

    def particle_displ_over(self, th):
        self._log("particle_displ_over")
        return self._over.pop(0) if self._over else False

    def zero_force_energy(self):
        self._log("zero")

    def force_to_accel(self):
        self._log("accel")

    def lj_compute_force(self, **kw):
        self._log("lj")

    def johnson_force(self, rcut, parameters, compute_virial=False):
        from exastamp_amd.operators import JOHNSON_PARAMETERS
        from exastamp_amd.units import parse_quantity
        self.johnson_args = (rcut, [parse_quantity(parameters[k]) for k in JOHNSON_PARAMETERS])
        self._log("johnson")

    def push_f_v_r(self, dt, dt_scale=1.0, xform_mode=None):
        self._log("push_f_v_r")

    def push_f_v(self, dt, dt_scale=0.5, xform_mode=None):
        self._log("push_f_v")

    def simulation_thermodynamic_state(self):
        import numpy as np
        from exastamp_amd.operators import ThermodynamicState
        raw = np.zeros(27)
        raw[21:24] = 1.0
        raw[25] = 1.0
        raw[26] = max(self.n, 1)
        return ThermodynamicState(raw, 1000.0)


def _stub_run(monkeypatch, deck_path, steps, over):
    from exastamp_amd import deck, operators
    monkeypatch.setattr(operators, "Simulation", type("S", (_StubSim,), {"over_script": tuple(over)}))
    cfg = deck.load_deck(deck_path, [GOLDEN], [(["global", "max_iteration"], str(steps))])
    r = deck.Runner(cfg, out=open("/dev/null", "w"))
    r.run()
    return r, r.sim.calls


def _steps(calls):
    """the calls after the initial build and forces, split at each push_f_v_r"""
    out, cur = [], None
    for c in calls[calls.index("push_f_v_r"):]:
        if c == "push_f_v_r":
            cur = []
            out.append(cur)
        cur.append(c)
    return out


def test_deck_standard_scheme_call_order(monkeypatch):
    """verlet_nve with check_and_update_particles as in config_numerical_schemes.msp and
    config_move_particles.msp: the trigger before the forces, a rebuild only when it fires."""
    _, calls = _stub_run(monkeypatch, GOLDEN / "decks" / "lj" / "single_specy_nosym.msp", 3, [False, True, False])
    steps = _steps(calls)
    assert len(steps) == 3
    quiet = ["push_f_v_r", "push_f_v", "particle_displ_over", "zero", "lj", "accel", "push_f_v"]
    assert steps[0] == quiet and steps[2] == quiet
    assert steps[1] == ["push_f_v_r", "push_f_v", "particle_displ_over", "chunk_neighbors", "zero", "lj", "accel",
                        "push_f_v"]


def test_deck_speculative_scheme_call_order(monkeypatch):
    """config_speculative_verlet.msp: forces on the old list first, the async trigger read
    after them, and on a hit a rebuild and the forces again."""
    _, calls = _stub_run(monkeypatch, GOLDEN / "decks" / "lj" / "single_specy_nosym_speculative_verlet.msp", 3,
                         [False, True, False])
    steps = _steps(calls)
    assert len(steps) == 3
    quiet = ["push_f_v_r", "particle_displ_over", "push_f_v", "zero", "lj", "accel", "push_f_v"]
    assert steps[0] == quiet and steps[2] == quiet
    assert steps[1] == ["push_f_v_r", "particle_displ_over", "push_f_v", "zero", "lj", "accel", "chunk_neighbors",
                        "zero", "lj", "accel", "push_f_v"]


def test_deck_block_conditions():
    from exastamp_amd import deck
    r = deck.Runner.__new__(deck.Runner)
    r.slots, r.rebind = {"flag": True, "off": False}, []
    assert r.condition("flag") and not r.condition("not flag")
    assert not r.condition("off") and r.condition("not off")
    assert r.condition("not missing") and r.condition(True)


def test_johnson_deck_reaches_the_operator(monkeypatch):
    """The reference's Johnson deck: rcut 6.1 A published to rcut_max (6.1 + rcut_inc
    1 A = neighbour distance), johnson_force called once per force evaluation, its 19
    parameters converted with their units (fe, rhoe in eV/A; energies in eV)."""
    from exastamp_amd import units
    r, calls = _stub_run(monkeypatch, DECKS / "eam_johnson" / "single_specy.msp", 3, ())
    assert calls.count("johnson") == 4                     # initial forces + 3 steps
    rc, p = r.sim.johnson_args
    assert rc == 6.1
    assert p[0] == 2.860082 and p[1] == 3.08634 * units.EV and p[3] == 8.489528
    assert p[13] == -5.14 * units.EV and p[18] == 0.848843 * units.EV
    assert float(r.get("rcut_max")) == 6.1


def test_johnson_init_supplies_the_parameters_slot(monkeypatch, tmp_path):
    """johnson_init's parameters slot is an OUTPUT (eam_potential.cu:62-70) and the force
    operator's an INPUT_OUTPUT (:92): a deck that sets the parameters through johnson_init,
    rebound to a shared name as the eam_alloy decks do, and gives johnson_force only its
    rcut, runs with those parameters; with neither source it is a deck error."""
    import yaml
    from exastamp_amd import deck, units
    src = yaml.safe_load((DECKS / "eam_johnson" / "single_specy.msp").read_text())
    params = src["johnson_force"].pop("parameters")
    src["init_parameters"] = {"rebind": {"parameters": "johnson_parameters"},
                              "body": [{"johnson_init": {"parameters": params}}]}
    src["compute_force"] = {"rebind": {"parameters": "johnson_parameters"}, "body": ["johnson_force"]}
    path = tmp_path / "johnson_init.msp"
    path.write_text(yaml.safe_dump(src))
    r, calls = _stub_run(monkeypatch, path, 2, ())
    assert calls.count("johnson") == 3
    rc, p = r.sim.johnson_args
    assert rc == 6.1 and p[1] == 3.08634 * units.EV and p[18] == 0.848843 * units.EV
    del src["init_parameters"]
    path.write_text(yaml.safe_dump(src))
    with pytest.raises(deck.DeckError, match="no parameters"):
        _stub_run(monkeypatch, path, 2, ())
