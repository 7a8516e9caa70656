"""EAM adjoint (dLoss/dtheta for the tabulated knots) against FP64 central finite
differences (h = 1e-6) of the oracle's forces and energies.

C4 at its full size: 64 independently strained (+-3 %), thermally displaced
(300-1500 K) Cu boxes of 6^3 FCC cells = 864 atoms each, 3 x 500-knot synthetic
tables, targets from a seeded +-2 % perturbation of every knot, loss = force MSE +
0.1 * per-atom energy MSE (eV units), 32 random knots among those the boxes sample.
Tolerances: FP64 adjoint <= 1e-6 relative to the difference quotient, the
FP32-table adjoint (distances, weights and accumulation FP64) <= 1e-5.

C5's two-species case: a Cu/Ni slab (random 50/50 substitution, 1 % vacancies, free
surface along z) with 7 x 2000-knot tables, two knots from each of the seven tables
(F_Cu, F_Ni, rho_Cu, rho_Ni, r*phi Cu-Cu, Ni-Cu, Ni-Ni), same tolerances; and the
full 64^3-cell slab through size-independent properties.
"""
import concurrent.futures as cf

import numpy as np
import pytest

from conftest import cu_tables, cuni_tables

pytestmark = pytest.mark.gpu
CU = 63.546
NI = 58.693


class OracleBoxes:
    """One FP64 oracle per periodic box (lists built once): forces (eV/A) and per-box
    energies (eV) for any table set, boxes evaluated in parallel (one thread each)."""

    def __init__(self, boxes, masses, rc, threads=1):
        import pyoracle
        self.boxes = boxes
        self.rc = rc
        self.os = []
        for b in boxes:
            o = pyoracle.Oracle(threads=threads)
            o.set_config(b, masses)
            self.os.append(o)
        self.built = False
        self.pool = cf.ThreadPoolExecutor(max_workers=min(16, len(boxes)))

    def fe(self, tab, mask=None):
        from exastamp_amd import units

        def one(o):
            o.set_eam(tab)
            if mask is not None:
                o.set_eam_types(mask)
            if not self.built:
                o.build_neighbors(self.rc, 0.3)
            o.compute_forces("eam")
            return o.forces() / units.EV, o.state()["ep"].sum() / units.EV

        res = list(self.pool.map(one, self.os))
        self.built = True
        return np.concatenate([r[0] for r in res]), np.array([r[1] for r in res])

    def close(self):
        self.pool.shutdown()
        for o in self.os:
            o.close()


def _c4_batch(nconf=64, cells=(6, 6, 6), seed=41):
    from exastamp_amd import inputs
    return inputs.strained_batch(nconf, cells, 3.615, [CU], seed=seed)


def _targets(tab, seed=5):
    from exastamp_amd import inputs
    theta0 = tab.theta()
    return tab.with_theta(theta0 * (1.0 + inputs.rng(seed).uniform(-0.02, 0.02, theta0.size)))


def _sim(cfg, tab, masses, names):
    from exastamp_amd.operators import Simulation
    sim = Simulation(0)
    sim.species(list(zip(names, masses)))
    sim.set_config(cfg)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    return sim


def _grads(sim, tab, Ft, Et, nper):
    """dL/dtheta through the torch harness for each (forward, adjoint) precision."""
    import torch
    from exastamp_amd.fitting import EamForcesEnergies, fit_loss
    out = {}
    for fwd_mixed, adj_mixed in ((False, False), (False, True), (True, True)):
        theta = torch.tensor(tab.theta(), requires_grad=True)
        F, E = EamForcesEnergies.apply(theta, sim, fwd_mixed, adj_mixed)
        fit_loss(F, E, torch.from_numpy(Ft), torch.from_numpy(Et), nper).backward()
        out[(fwd_mixed, adj_mixed)] = theta.grad.numpy().copy()
    # the default is the FP64 forward with the FP32-table adjoint
    theta = torch.tensor(tab.theta(), requires_grad=True)
    F, E = EamForcesEnergies.apply(theta, sim)
    fit_loss(F, E, torch.from_numpy(Ft), torch.from_numpy(Et), nper).backward()
    np.testing.assert_allclose(theta.grad.numpy(), out[(False, True)], rtol=1e-9,
                               atol=1e-12 * np.abs(out[(False, True)]).max())
    return out


def _central_differences(fe, tab, pick, Ft, Et, nper, h=1e-6):
    from exastamp_amd.fitting import fit_loss_np
    theta0 = tab.theta()
    fd = []
    for k in pick:
        vals = []
        for s in (+1, -1):
            th = theta0.copy()
            th[k] += s * h
            F, E = fe(tab.with_theta(th))
            vals.append(fit_loss_np(F, E, Ft, Et, nper))
        fd.append((vals[0] - vals[1]) / (2 * h))
    return np.array(fd)


def _report(name, grads, pick, fd):
    err = {k: float((np.abs(v[pick] - fd) / np.abs(fd)).max()) for k, v in grads.items()}
    print(f"{name} adjoint vs FD ({len(pick)} knots): FP64 fwd+adj {err[(False, False)]:.3e}; "
          f"FP64 fwd + FP32-table adj {err[(False, True)]:.3e}; "
          f"mixed fwd + FP32-table adj {err[(True, True)]:.3e}")
    return err


def test_adjoint_c4_full_batch():
    tab = cu_tables(500)
    from exastamp_amd import inputs
    batch = _c4_batch()
    assert batch.nrep == 64 and batch.n == 64 * 864
    boxes = inputs.split_replicas(batch)
    ob = OracleBoxes(boxes, [CU], tab.rc)
    Ft, Et = ob.fe(_targets(tab))
    nper = np.bincount(batch.rep)
    sim = _sim(batch, tab, [CU], ["Cu"])
    grads = _grads(sim, tab, Ft, Et, nper)
    g64 = grads[(False, False)]
    active = np.flatnonzero(np.abs(g64) > 1e-3 * np.abs(g64).max())
    pick = np.random.default_rng(3).choice(active, size=32, replace=False)
    fd = _central_differences(ob.fe, tab, pick, Ft, Et, nper)
    err = _report("C4 64 x 864 atoms", grads, pick, fd)
    ob.close()
    sim.close()
    assert err[(False, False)] <= 1e-6
    assert err[(False, True)] <= 1e-5
    # an FP32 forward bounds the gradient by (force error) / (residual size)
    assert err[(True, True)] <= 1e-4


def _c5_slab(cells, seed=42):
    from exastamp_amd import inputs
    return inputs.alloy_slab_config(cells, 3.57, [CU, NI], vacancy_frac=0.01, vacuum=10.0, T=300.0, seed=seed)


def test_adjoint_c5_two_species():
    tab = cuni_tables(2000)
    cfg = _c5_slab((8, 8, 8))
    ob = OracleBoxes([cfg], [CU, NI], tab.rc, threads=8)
    Ft, Et = ob.fe(_targets(tab, seed=6))
    nper = np.array([cfg.n])
    sim = _sim(cfg, tab, [CU, NI], ["Cu", "Ni"])
    grads = _grads(sim, tab, Ft, Et, nper)
    g64 = grads[(False, False)]
    # two sampled knots of each of the 7 tables (setfl order: F_Cu, F_Ni, rho_Cu,
    # rho_Ni, z2r 11, 21, 22)
    bounds = np.cumsum([0, tab.nrho, tab.nrho, tab.nr, tab.nr, tab.nr, tab.nr, tab.nr])
    rng = np.random.default_rng(9)
    pick = []
    for t in range(7):
        seg = np.arange(bounds[t], bounds[t + 1])
        act = seg[np.abs(g64[seg]) > 1e-3 * np.abs(g64[seg]).max()]
        pick += list(rng.choice(act, size=2, replace=False))
    pick = np.array(pick)
    fd = _central_differences(ob.fe, tab, pick, Ft, Et, nper)
    err = _report(f"C5 slab {cfg.n} atoms", grads, pick, fd)
    ob.close()
    sim.close()
    assert err[(False, False)] <= 1e-6
    assert err[(False, True)] <= 1e-5
    assert err[(True, True)] <= 1e-3


def test_c5_full_slab_properties():
    """The 64^3-cell Cu/Ni slab (1.04 M atoms): mixed forces within 1e-5 of rms|F| of
    FP64, zero net force, and the adjoint's dE/dtheta (dL/dF = 0, dL/dE = 1) against
    FP64 central differences of the GPU's own total energy on one knot per table."""
    tab = cuni_tables(2000)
    cfg = _c5_slab((64, 64, 64))
    sim = _sim(cfg, tab, [CU, NI], ["Cu", "Ni"])
    forces = {}
    for mixed in (False, True):
        sim.zero_force_energy()
        sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, mixed=mixed)
        sim.force_to_accel()
        forces[mixed] = sim.forces()
    f64, f32 = forces[False], forces[True]
    rms = np.sqrt((f64 ** 2).sum(1).mean())
    err = np.sqrt(((f32 - f64) ** 2).sum(1)).max() / rms
    net = np.abs(f64.sum(0)).max() / (rms * np.sqrt(cfg.n))
    print(f"C5 full slab {cfg.n} atoms: mixed force err {err:.2e} of rms|F|, net force {net:.2e}")
    assert err <= 1e-5
    assert net <= 1e-10
    # energy gradient: the FP64 forward above left the checkpoint of the mixed call; redo FP64
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True)
    g = sim.eam_adjoint(np.zeros((sim.n, 3)), np.ones(1))
    theta0 = tab.theta()
    bounds = np.cumsum([0, tab.nrho, tab.nrho, tab.nr, tab.nr, tab.nr, tab.nr, tab.nr])
    h = 1e-6
    for t in range(7):
        k = bounds[t] + int(np.argmax(np.abs(g[bounds[t]:bounds[t + 1]])))
        es = []
        for s in (+1, -1):
            th = theta0.copy()
            th[k] += s * h
            sim.eam_alloy_init(tab.with_theta(th))
            sim.zero_force_energy()
            sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True)
            es.append(sim.ep().sum())
        fd = (es[0] - es[1]) / (2 * h)
        print(f"  table {t} knot {k}: dE/dtheta {g[k]:.9e} vs FD {fd:.9e}")
        assert g[k] == pytest.approx(fd, rel=1e-6)
    sim.close()


def test_adjoint_requires_checkpoint():
    from exastamp_amd.operators import XsError
    from exastamp_amd import inputs
    tab = cu_tables(500)
    batch = _c4_batch(6, (4, 4, 4), seed=21)
    sim = _sim(batch, tab, [CU], ["Cu"])
    with pytest.raises(XsError, match="checkpoint"):
        sim.eam_adjoint(np.zeros((sim.n, 3)), np.zeros(batch.nrep))
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True)
    sim.push_f_v_r(0.001)
    with pytest.raises(XsError, match="checkpoint"):
        sim.eam_adjoint(np.zeros((sim.n, 3)), np.zeros(batch.nrep))
    assert len(inputs.split_replicas(batch)) == 6
    sim.close()


def test_energy_only_gradient_is_linear_in_dE():
    """With dL/dF = 0 and dL/dE = 1 per replica, the gradient is dE_total/dtheta:
    check it against a central difference of the GPU energy itself (FP64)."""
    tab = cu_tables(500)
    batch = _c4_batch(6, (4, 4, 4), seed=21)
    sim = _sim(batch, tab, [CU], ["Cu"])
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True)
    g = sim.eam_adjoint(np.zeros((sim.n, 3)), np.ones(batch.nrep))
    theta0 = tab.theta()
    k = int(np.argmax(np.abs(g)))
    h = 1e-6
    es = []
    for s in (+1, -1):
        th = theta0.copy()
        th[k] += s * h
        sim.eam_alloy_init(tab.with_theta(th))
        sim.zero_force_energy()
        sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True)
        es.append(sim.ep().sum())
    fd = (es[0] - es[1]) / (2 * h)
    assert g[k] == pytest.approx(fd, rel=1e-6)
    sim.close()
