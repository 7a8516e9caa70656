"""The types slot of eam_alloy_force (eam_potential_multimat.cu:90,154-166) through the
C ABI against the oracle.

A random Cu/Ni FCC crystal with the synthetic two-element tables and types: [Cu]
(or [Ni]): pairs that involve the unlisted species contribute nothing, its atoms get
rho_dEmb = 0, no energy and no force.  Checked for the block list (FP64 1e-11 of
rms|F|, mixed 1e-5), the global list, the Newton-symmetric form, the split passes, the
fused NVE loop (20 steps, 1e-9 in energy) and the adjoint (zero gradient for every
table only the unlisted species uses; FP64 central differences of the oracle on
the others, 1e-6).
"""
import numpy as np
import pytest

from conftest import cuni_tables
from test_gpu_parity import frel, make_oracle, make_sim

pytestmark = pytest.mark.gpu
MASSES = [63.546, 58.693]
NAMES = ["Cu", "Ni"]


def _cuni(cells=(5, 5, 5), seed=41, T=300.0):
    from exastamp_amd import inputs
    cfg = inputs.lattice_config(cells, 3.57, MASSES, T=T, jitter=0.06, seed=seed)
    t = (inputs.rng(seed + 1).random(cfg.n) < 0.5).astype(np.uint8)
    return inputs.Config(cfg.r, cfg.v, t, cfg.lo, cfg.hi)


def _oracle(oracle_mod, cfg, tab, mask):
    o = make_oracle(oracle_mod, cfg, MASSES)
    o.set_eam(tab)
    o.set_eam_types(mask)
    o.build_neighbors(tab.rc, 0.3)
    o.compute_forces("eam")
    return o


@pytest.mark.parametrize("listed", [["Cu"], ["Ni"]])
def test_types_subset_matches_oracle(oracle_mod, listed):
    tab = cuni_tables(2000)
    cfg = _cuni()
    mask = sum(1 << NAMES.index(t) for t in listed)
    off = np.array([(mask >> int(t)) & 1 == 0 for t in cfg.type])
    o = _oracle(oracle_mod, cfg, tab, mask)
    ref, fref = o.state(), o.forces()
    sim = make_sim(cfg, MASSES, NAMES)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    assert sim.list_info()["kind"] == "block"
    for mixed, sym, tol in ((False, False, 1e-11), (True, False, 1e-5), (False, True, 1e-11)):
        sim.zero_force_energy()
        sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, compute_virial=True, mixed=mixed,
                            eam_symmetry=sym, types=listed)
        fp = sim.rho_dEmb()
        sim.force_to_accel()
        f = sim.forces()
        err = frel(f, fref)
        print(f"types {listed} mixed={mixed} symmetric={sym}: force err {err:.2e} of rms|F|")
        assert err <= tol
        assert np.all(f[off] == 0.0) and np.all(sim.ep()[off] == 0.0) and np.all(fp[off] == 0.0)
        assert abs(sim.ep().sum() - ref["ep"].sum()) <= max(tol, 1e-12) * abs(ref["ep"].sum())
        vir, vref = sim.virial().sum(0), ref["virial"].sum(0)
        assert np.abs(vir - vref).max() <= max(tol, 1e-12) * np.abs(vref).max()
    # the split passes (eam_rho, then eam_rho2emb, then eam_force) give the same forces
    sim.zero_force_energy()
    for r, e, fo in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        sim.eam_alloy_force(rcut=tab.rc, eam_rho=r, eam_rho2emb=e, eam_force=fo, types=listed)
    sim.force_to_accel()
    assert frel(sim.forces(), fref) <= 1e-11
    # the empty list is every species
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, types=[])
    sim.force_to_accel()
    o.set_eam_types(0xFFFFFFFF)
    o.compute_forces("eam")
    assert frel(sim.forces(), o.forces()) <= 1e-11
    with pytest.raises(ValueError, match="unknown types"):
        sim.eam_alloy_force(rcut=tab.rc, types=["Al"])
    sim.close()


def test_types_subset_global_list(oracle_mod, monkeypatch):
    monkeypatch.setenv("XSMD_LIST", "global")
    tab = cuni_tables(2000)
    cfg = _cuni(seed=43)
    o = _oracle(oracle_mod, cfg, tab, 0b01)
    sim = make_sim(cfg, MASSES, NAMES)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    assert sim.list_info()["kind"] == "global"
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, types=["Cu"])
    sim.force_to_accel()
    assert frel(sim.forces(), o.forces()) <= 1e-11
    assert abs(sim.ep().sum() - o.state()["ep"].sum()) <= 1e-12 * abs(o.state()["ep"].sum())
    sim.close()


def test_types_subset_nve_loop(oracle_mod):
    tab = cuni_tables(2000)
    cfg = _cuni(seed=45, T=600.0)
    sim = make_sim(cfg, MASSES, NAMES)
    sim.eam_alloy_init(tab)
    sim.ctx.xs_eam_alloy_types(0b01)
    st, th = sim.md_run("eam", 20, 0.001, tab.rc, 0.3, thermo_every=10)
    o = make_oracle(oracle_mod, cfg, MASSES)
    o.set_eam(tab)
    o.set_eam_types(0b01)
    _, th_ref = o.md_run("eam", 20, 0.001, tab.rc, 0.3)
    for k, step in enumerate((0, 10, 20)):
        e, e_ref = th[k].total_energy, th_ref[step, 21:25].sum()
        print(f"types [Cu] NVE step {step}: total energy {e:.12e} vs {e_ref:.12e}")
        assert abs(e - e_ref) <= 1e-9 * abs(e_ref)
    sim.close()


def test_types_subset_adjoint(oracle_mod):
    """With types [Cu] the loss does not depend on any Ni table (F_Ni, rho_Ni, the
    Cu-Ni and Ni-Ni r*phi): those gradients are exactly 0; the Cu tables' gradients
    match FP64 central differences of the oracle's masked loss."""
    from exastamp_amd import inputs
    from exastamp_amd.fitting import fit_loss_np
    from exastamp_amd.units import EV
    tab = cuni_tables(2000)
    cfg = _cuni(cells=(4, 4, 4), seed=47)
    theta0 = tab.theta()
    tab_t = tab.with_theta(theta0 * (1.0 + inputs.rng(48).uniform(-0.02, 0.02, theta0.size)))

    def oracle_fe(t):
        o = _oracle(oracle_mod, cfg, t, 0b01)
        return o.forces() / EV, np.array([o.state()["ep"].sum() / EV])

    Ft, Et = oracle_fe(tab_t)
    sim = make_sim(cfg, MASSES, NAMES)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, types=["Cu"])
    sim.force_to_accel()
    F, E = sim.forces() / EV, sim.replica_energies() / EV
    nper = np.array([cfg.n])
    gF = 2.0 * (F - Ft) / F.size / EV
    gE = 0.1 * 2.0 * (E - Et) / nper ** 2 / E.size / EV
    g = sim.eam_adjoint(gF, gE)
    nrho, nr = tab.nrho, tab.nr
    # setfl order: F_Cu, F_Ni | rho_Cu, rho_Ni | z2r 11 (Cu-Cu), 21 (Ni-Cu), 22 (Ni-Ni)
    ni_only = np.r_[nrho:2 * nrho, 2 * nrho + nr:2 * nrho + 2 * nr, 2 * nrho + 3 * nr:2 * nrho + 5 * nr]
    assert np.all(g[ni_only] == 0.0)
    cu = np.setdiff1d(np.arange(g.size), ni_only)
    active = cu[np.abs(g[cu]) > 1e-3 * np.abs(g).max()]
    pick = np.random.default_rng(4).choice(active, size=12, replace=False)
    h = 1e-6
    for k in pick:
        v = []
        for s in (+1, -1):
            th = theta0.copy()
            th[k] += s * h
            Fk, Ek = oracle_fe(tab.with_theta(th))
            v.append(fit_loss_np(Fk, Ek, Ft, Et, nper))
        fd = (v[0] - v[1]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-6), k
    sim.close()
