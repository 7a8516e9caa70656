"""The device-resident fitting path (xs_ctx_set_stream, xs_eam_theta_dev,
xs_forces_energies_dev, xs_eam_adjoint_dev) against the host path it replaces.

Bars: spline rows built on the device from theta give FP64 forces equal bit for bit
to those of the host build (xs_eam_alloy_tables with the same knots) and replica
energies within 1e-13 (summation order); the device adjoint equals the host-output
adjoint within 1e-12 of its largest component (the same kernels; their LDS
accumulation adds in scheduling order, so two host calls differ as much); the torch harness with theta on the GPU gives the
CPU-theta harness's gradients within 1e-10 (the loss itself is summed by torch on
another device).  Cases: a batch of 8 strained Cu boxes (4^3 FCC cells each,
500-knot tables) and a Cu/Ni slab (two species, 2000-knot tables).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import cu_tables, cuni_tables

pytestmark = pytest.mark.gpu
CU = 63.546
NI = 58.693


def _case(name):
    from exastamp_amd import inputs
    if name == "cu_batch":
        cfg = inputs.strained_batch(8, (4, 4, 4), 3.615, [CU], seed=7)
        return cfg, [CU], ["Cu"], cu_tables(500)
    cfg = inputs.alloy_slab_config((6, 6, 6), 3.57, [CU, NI], vacancy_frac=0.01, vacuum=10.0, T=300.0, seed=43)
    return cfg, [CU, NI], ["Cu", "Ni"], cuni_tables(2000)


def _sim(cfg, masses, names, tab):
    from exastamp_amd.operators import Simulation
    sim = Simulation(0)
    sim.species(list(zip(names, masses)))
    sim.set_config(cfg)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    return sim


def _perturbed(tab, seed=11):
    from exastamp_amd import inputs
    th = tab.theta()
    return th * (1.0 + inputs.rng(seed).uniform(-0.02, 0.02, th.size))


def _forward(sim, tab):
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True)
    sim.force_to_accel()


@pytest.mark.parametrize("name", ["cu_batch", "cuni_slab"])
def test_theta_dev_forces_energies(name):
    import torch
    cfg, masses, names, tab = _case(name)
    th = _perturbed(tab)
    sim = _sim(cfg, masses, names, tab)
    sim.eam_alloy_init(tab.with_theta(th))
    _forward(sim, tab)
    F_host, E_host = sim.forces(), sim.replica_energies()
    assert sim.eam_table_info()["s_pieces"] >= 0
    theta = torch.tensor(th, device="cuda")
    sim.ctx.xs_ctx_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
    sim.ctx.xs_eam_theta_dev(C.c_void_p(theta.data_ptr()))
    assert sim.eam_table_info()["s_pieces"] == 0
    _forward(sim, tab)
    F = torch.empty((cfg.n, 3), dtype=torch.float64, device="cuda")
    E = torch.empty(cfg.nrep, dtype=torch.float64, device="cuda")
    sim.ctx.xs_forces_energies_dev(C.c_void_p(F.data_ptr()), C.c_void_p(E.data_ptr()), 1.0)
    F, E = F.cpu().numpy(), E.cpu().numpy()
    print(f"{name}: max |dF| {np.abs(F - F_host).max():.3e}, max |dE|/|E| {np.abs((E - E_host) / E_host).max():.3e}")
    assert np.array_equal(F, F_host)
    assert np.abs(E - E_host).max() <= 1e-13 * np.abs(E_host).max()
    # the host view of the same state agrees
    assert np.array_equal(sim.forces(), F)
    sim.ctx.xs_ctx_own_stream()
    sim.close()


@pytest.mark.parametrize("name", ["cu_batch", "cuni_slab"])
@pytest.mark.parametrize("mixed", [False, True])
def test_adjoint_dev_equals_host(name, mixed):
    import torch
    from exastamp_amd import _lib
    cfg, masses, names, tab = _case(name)
    sim = _sim(cfg, masses, names, tab)
    _forward(sim, tab)
    g = np.random.default_rng(5).normal(size=(cfg.n, 3))
    ge = np.random.default_rng(6).normal(size=cfg.nrep)
    host = sim.eam_adjoint(g, ge, mixed=mixed)
    gd = torch.tensor(g, device="cuda")
    ged = torch.tensor(ge, device="cuda")
    out = torch.empty(host.size, dtype=torch.float64, device="cuda")
    sim.ctx.xs_ctx_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
    # dL/dF and dL/dE given at twice their value, scaled back by the call
    gd2, ged2 = 2.0 * gd, 2.0 * ged
    sim.ctx.xs_eam_adjoint_dev(float(tab.rc), C.c_void_p(gd2.data_ptr()), C.c_void_p(ged2.data_ptr()), 0.5,
                               C.c_void_p(out.data_ptr()), _lib.XS_MIXED if mixed else 0)
    dev = out.cpu().numpy()
    sim.ctx.xs_ctx_own_stream()
    again = sim.eam_adjoint(g, ge, mixed=mixed)
    scale = np.abs(host).max()
    print(f"{name} mixed={mixed}: device vs host {np.abs(dev - host).max() / scale:.3e}, "
          f"host vs host {np.abs(again - host).max() / scale:.3e}")
    # the adjoint's LDS accumulation adds in scheduling order: last bits vary per call
    assert np.abs(dev - host).max() <= 1e-12 * scale
    sim.close()


@pytest.mark.parametrize("name", ["cu_batch", "cuni_slab"])
def test_torch_harness_on_device(name):
    import torch
    from exastamp_amd.fitting import EamForcesEnergies, fit_loss
    cfg, masses, names, tab = _case(name)
    sim = _sim(cfg, masses, names, tab)
    th = _perturbed(tab, seed=13)
    # targets: the FP64 forward at the initial tables
    from exastamp_amd.units import EV
    _forward(sim, tab)
    Ft, Et = sim.forces() / EV, sim.replica_energies() / EV
    nper = np.bincount(cfg.rep) if cfg.rep is not None else np.array([cfg.n])
    res = {}
    for dev in ("cpu", "cuda"):
        theta = torch.tensor(th, device=dev, requires_grad=True)
        F, E = EamForcesEnergies.apply(theta, sim, False, True)
        loss = fit_loss(F, E, torch.tensor(Ft, device=dev), torch.tensor(Et, device=dev), nper)
        loss.backward()
        res[dev] = (F.detach().cpu().numpy(), E.detach().cpu().numpy(), theta.grad.cpu().numpy(), float(loss.detach()))
    (Fc, Ec, gc, lc), (Fg, Eg, gg, lg) = res["cpu"], res["cuda"]
    print(f"{name}: loss {lc:.6e} vs {lg:.6e}; grad rel diff {np.abs(gg - gc).max() / np.abs(gc).max():.3e}")
    # F / EV: torch divides by a Python scalar through its reciprocal on the GPU
    assert np.abs(Fg - Fc).max() <= 1e-15 * np.abs(Fc).max()
    assert np.abs(Eg - Ec).max() <= 1e-13 * np.abs(Ec).max()
    assert abs(lg - lc) <= 1e-12 * abs(lc)
    assert np.abs(gg - gc).max() <= 1e-10 * np.abs(gc).max()
    sim.ctx.xs_ctx_own_stream()
    sim.close()


def test_device_entry_points_refuse_out_of_order_calls():
    """Each device entry point names what it needs: tables before xs_eam_theta_dev, a
    neighbour list before xs_forces_energies_dev, the forward's checkpoint before
    xs_eam_adjoint_dev; null pointers are argument errors."""
    import torch
    from exastamp_amd._lib import XsError
    from exastamp_amd.operators import Simulation
    cfg, masses, names, tab = _case("cu_batch")
    buf = torch.zeros(max(3 * cfg.n, tab.theta().size), dtype=torch.float64, device="cuda")
    p = C.c_void_p(buf.data_ptr())
    sim = Simulation(0)
    sim.species(list(zip(names, masses)))
    sim.set_config(cfg)
    with pytest.raises(XsError, match="eam tables"):
        sim.ctx.xs_eam_theta_dev(p)
    with pytest.raises(XsError, match="neighbour list"):
        sim.ctx.xs_forces_energies_dev(p, p, 1.0)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    with pytest.raises(XsError, match="checkpoint"):
        sim.ctx.xs_eam_adjoint_dev(float(tab.rc), p, p, 1.0, p, 0)
    with pytest.raises(XsError):
        sim.ctx.xs_eam_theta_dev(None)
    with pytest.raises(XsError):
        sim.ctx.xs_forces_energies_dev(None, p, 1.0)
    sim.close()


def test_harness_on_a_side_stream():
    """The device harness on a non-default torch stream (default priority) gives the
    default stream's results bit for bit, and the context returns to its own stream for
    host-path calls afterwards."""
    import torch
    from exastamp_amd.fitting import EamForcesEnergies, fit_loss
    cfg, masses, names, tab = _case("cu_batch")
    sim = _sim(cfg, masses, names, tab)
    from exastamp_amd.units import EV
    _forward(sim, tab)
    Ft = torch.tensor(sim.forces() / EV, device="cuda")
    Et = torch.tensor(sim.replica_energies() / EV, device="cuda")
    nper = torch.tensor(np.bincount(cfg.rep), dtype=torch.float64, device="cuda")
    th = _perturbed(tab, seed=17)

    def step():
        theta = torch.tensor(th, device="cuda", requires_grad=True)
        F, E = EamForcesEnergies.apply(theta, sim)
        fit_loss(F, E, Ft, Et, nper).backward()
        return F.detach().cpu().numpy(), theta.grad.cpu().numpy()

    F0, g0 = step()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        F1, g1 = step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert np.array_equal(F0, F1)
    scale = np.abs(g0).max()
    assert np.abs(g1 - g0).max() <= 1e-12 * scale   # LDS accumulation order
    sim.ctx.xs_ctx_own_stream()
    # the host path on the context's own stream still works and agrees
    sim.eam_alloy_init(tab.with_theta(th))
    _forward(sim, tab)
    assert np.abs(sim.forces() / EV - F0).max() <= 1e-12 * np.abs(F0).max()
    sim.close()
