"""The FP32 tables in s = r^2 of the mixed single-element EAM force pass (cubic
pieces uniform in r^2; pairs closer than r_lo or past the table's end read the FP32
knot rows), against the FP64 oracle and against the knot rows alone
(XSMD_EAM_S_TABLE=0).

Bars: the host's acceptance figures (FP32 errors within 8 ulps of the largest
|rho'/r| and |phi'/r| and 4 of the largest |phi| over the table's range);
mixed forces vs the FP64 oracle max_i |dF_i| / rms|F| <= 1e-5 and, per atom,
|dF_i| / |F_i| <= 1e-5 for atoms with |F_i| > 0.1 rms|F|; energies 1e-6 relative;
every form of the operator (block list, global list, Newton-symmetric, split passes)
on the same tables within 1e-6 of rms|F| of each other.
"""
import numpy as np
import pytest

from conftest import cu_tables
from test_gpu_parity import CU_MASS, cu_config, make_oracle

pytestmark = pytest.mark.gpu


def _sim(cfg, tab, monkeypatch=None, env=None, lists=None):
    from exastamp_amd.operators import Simulation
    if monkeypatch is not None:
        if env is None:
            monkeypatch.delenv("XSMD_EAM_S_TABLE", raising=False)
        else:
            monkeypatch.setenv("XSMD_EAM_S_TABLE", env)
        if lists is None:
            monkeypatch.delenv("XSMD_LIST", raising=False)
        else:
            monkeypatch.setenv("XSMD_LIST", lists)
    sim = Simulation(0)
    sim.species([("Cu", CU_MASS)])
    sim.set_config(cfg)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    return sim


def _mixed(sim, tab, **kw):
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, mixed=True, **kw)
    sim.force_to_accel()
    return sim.forces(), sim.ep().sum()


def _oracle(oracle_mod, cfg, tab):
    o = make_oracle(oracle_mod, cfg, [CU_MASS])
    o.set_eam(tab)
    o.build_neighbors(tab.rc, 0.3)
    o.compute_forces("eam")
    return o.forces(), o.state()["ep"].sum()


def _errors(f, ref):
    mag = np.sqrt((ref ** 2).sum(1))
    rms = np.sqrt((mag ** 2).mean())
    d = np.sqrt(((f - ref) ** 2).sum(1))
    sel = mag > 0.1 * rms
    return d.max() / rms, (d[sel] / mag[sel]).max()


def test_s_table_layout(monkeypatch):
    tab = cu_tables()
    sim = _sim(cu_config((6, 6, 6)), tab, monkeypatch)
    info = sim.eam_table_info()
    print(info)
    assert info["s_pieces"] >= 256
    assert info["r_lo"] == pytest.approx(0.3 * 4999 * tab.dr)
    u = info["ulps"]
    assert u["phi"] <= 4 and u["rho_slope_over_r"] <= 8 and u["phi_slope_over_r"] <= 8
    sim.close()
    sim = _sim(cu_config((6, 6, 6)), tab, monkeypatch, env="0")
    assert sim.eam_table_info()["s_pieces"] == 0
    sim.close()


def test_s_table_matches_oracle_600K(oracle_mod, monkeypatch):
    tab = cu_tables()
    cfg = cu_config((10, 10, 10), T=600.0, jitter=0.08, seed=17)
    ref, eref = _oracle(oracle_mod, cfg, tab)
    sim = _sim(cfg, tab, monkeypatch)
    f, e = _mixed(sim, tab)
    e_rms, e_atom = _errors(f, ref)
    print(f"s-table ({sim.eam_table_info()['s_pieces']} pieces): max|dF|/rms {e_rms:.2e}, per-atom {e_atom:.2e}, "
          f"energy {abs(e - eref) / abs(eref):.2e}")
    assert e_rms <= 1e-5
    assert e_atom <= 1e-5
    assert abs(e - eref) <= 1e-6 * abs(eref)
    # every operator form reads the same tables
    rms = np.sqrt((ref ** 2).sum(1).mean())
    for kw in ({"eam_symmetry": True},):
        f2, _ = _mixed(sim, tab, **kw)
        assert np.abs(f2 - f).max() <= 1e-6 * rms, kw
    sim.zero_force_energy()
    for r, em, fo in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        sim.eam_alloy_force(rcut=tab.rc, eam_rho=r, eam_rho2emb=em, eam_force=fo, mixed=True)
    sim.force_to_accel()
    assert np.abs(sim.forces() - f).max() <= 1e-6 * rms
    sim.close()
    sim = _sim(cfg, tab, monkeypatch, lists="global")
    assert sim.list_info()["kind"] == "global"
    f3, _ = _mixed(sim, tab)
    assert np.abs(f3 - f).max() <= 1e-6 * rms
    sim.close()


def test_s_table_vs_knot_rows(oracle_mod, monkeypatch):
    tab = cu_tables()
    cfg = cu_config((8, 8, 8), T=600.0, jitter=0.08, seed=19)
    ref, _ = _oracle(oracle_mod, cfg, tab)
    sim = _sim(cfg, tab, monkeypatch, env="0")
    fk, _ = _mixed(sim, tab)
    sim.close()
    sim = _sim(cfg, tab, monkeypatch)
    fs, _ = _mixed(sim, tab)
    sim.close()
    ek, ek_atom = _errors(fk, ref)
    es, es_atom = _errors(fs, ref)
    print(f"knot rows: {ek:.2e} / {ek_atom:.2e}; s-table: {es:.2e} / {es_atom:.2e}")
    assert es <= 1e-5 and ek <= 1e-5 and es_atom <= 1e-5 and ek_atom <= 1e-5


def test_close_pairs_read_knot_rows(oracle_mod, monkeypatch):
    """Atoms pushed to 1.3 A of a neighbour (below r_lo = 1.48 A) read the knot rows:
    forces still follow the oracle."""
    tab = cu_tables()
    from exastamp_amd import inputs
    cfg = cu_config((8, 8, 8), T=300.0, jitter=0.04, seed=29)
    r = cfg.r.copy()
    for a, b in ((0, 1), (100, 101), (777, 778)):
        d = r[b] - r[a]
        r[b] = r[a] + d / np.linalg.norm(d) * 1.3
    cfg = inputs.Config(r, cfg.v, cfg.type, cfg.lo, cfg.hi)
    ref, eref = _oracle(oracle_mod, cfg, tab)
    sim = _sim(cfg, tab, monkeypatch)
    assert sim.eam_table_info()["r_lo"] > 1.3
    f, e = _mixed(sim, tab)
    e_rms, _ = _errors(f, ref)
    print(f"close pairs: max|dF|/rms {e_rms:.2e}, energy {abs(e - eref) / abs(eref):.2e}")
    assert e_rms <= 1e-5
    assert abs(e - eref) <= 1e-6 * abs(eref)
    sim.close()


def test_gather_ceiling(oracle_mod, monkeypatch):
    """xs_gather_ceiling times both passes' gather kernels, leaves the simulation state
    alone, and refuses the lists and tables it does not model."""
    from exastamp_amd._lib import XsError
    tab = cu_tables()
    cfg = cu_config((10, 10, 10), T=600.0, jitter=0.08, seed=23)
    sim = _sim(cfg, tab, monkeypatch)
    f0, e0 = _mixed(sim, tab)
    for force in (False, True):
        for level in range(4):
            ms = sim.gather_ceiling(tab.rc, force, 3, level=level)
            print(f"gather ceiling ({'force' if force else 'density'} pass, level {level}): {ms:.4f} ms")
            assert 0.0 < ms < 1e3
    with pytest.raises(XsError):
        sim.gather_ceiling(tab.rc, False, 2, level=4)
    f1, e1 = _mixed(sim, tab)
    assert np.array_equal(f0, f1) and e0 == e1
    with pytest.raises(XsError):
        sim.gather_ceiling(tab.rc, False, 0)
    sim.close()
    sim = _sim(cfg, tab, monkeypatch, env="0")
    assert sim.gather_ceiling(tab.rc, False, 2) > 0.0
    with pytest.raises(XsError):
        sim.gather_ceiling(tab.rc, True, 2)
    sim.close()
    sim = _sim(cfg, tab, monkeypatch, lists="global")
    with pytest.raises(XsError):
        sim.gather_ceiling(tab.rc, False, 2)
    sim.close()
