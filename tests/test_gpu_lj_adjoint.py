"""LJ adjoint dLoss/d(epsilon, sigma) (xs_lj_adjoint, the torch LjForcesEnergies) against
FP64 central differences of the oracle's lj_compute_force.

Workload: the LJ argon configuration (FCC a = 5.26 A, 10^3 cells = 4 000 atoms,
epsilon = 0.0104 eV, sigma = 3.4 A, rc = 8.5 A) with 0.1 A thermal jitter; targets are
the oracle's forces and energy at epsilon and sigma moved by a seeded +-2 %;
loss = force MSE (eV/A)^2 + 0.1 per-atom energy MSE (eV^2).  The FP64 adjoint matches
the oracle's central differences (relative step 1e-6) to 1e-6; a replica batch gives
the sum of its boxes' gradients.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
AR = 39.948
RC = 8.5


def _config(cells=(10, 10, 10), seed=17):
    from exastamp_amd import inputs
    return inputs.lattice_config(cells, 5.26, [AR], T=100.0, jitter=0.1, seed=seed)


def _oracle_fe(oracle_mod, cfg, eps, sigma):
    from exastamp_amd.units import EV
    o = oracle_mod.Oracle(threads=8)
    o.set_config(cfg, [AR])
    o.set_lj(eps, sigma, RC)
    o.build_neighbors(RC, 0.3)
    o.compute_forces("lj")
    F, E = o.forces() / EV, np.array([o.state()["ep"].sum() / EV])
    o.close()
    return F, E


def test_lj_adjoint_matches_central_differences(oracle_mod):
    import torch
    from exastamp_amd import inputs
    from exastamp_amd.fitting import LjForcesEnergies, fit_loss, fit_loss_np
    from exastamp_amd.operators import Simulation
    from exastamp_amd.units import EV
    cfg = _config()
    eps0, sig0 = 0.0104, 3.4
    pert = 1.0 + inputs.rng(19).uniform(-0.02, 0.02, 2)
    Ft, Et = _oracle_fe(oracle_mod, cfg, eps0 * pert[0] * EV, sig0 * pert[1])
    nper = np.array([cfg.n])
    sim = Simulation(0)
    sim.species([("Ar", AR)])
    sim.set_config(cfg)
    sim.chunk_neighbors(rcut_max=RC, rcut_inc=0.3)
    theta = torch.tensor([eps0, sig0], dtype=torch.float64, requires_grad=True)
    F, E = LjForcesEnergies.apply(theta, sim, RC)
    assert np.abs(F.detach().numpy() - _oracle_fe(oracle_mod, cfg, eps0 * EV, sig0)[0]).max() <= \
        1e-11 * np.abs(Ft).max()
    loss = fit_loss(F, E, torch.from_numpy(Ft), torch.from_numpy(Et), nper)
    loss.backward()
    g = theta.grad.numpy()
    fd = []
    for k, x0 in enumerate((eps0, sig0)):
        h = 1e-6 * x0
        v = []
        for s in (+1, -1):
            th = [eps0, sig0]
            th[k] = x0 + s * h
            Fk, Ek = _oracle_fe(oracle_mod, cfg, th[0] * EV, th[1])
            v.append(fit_loss_np(Fk, Ek, Ft, Et, nper))
        fd.append((v[0] - v[1]) / (2 * h))
    fd = np.array(fd)
    err = np.abs(g - fd) / np.abs(fd)
    print(f"LJ adjoint: grad {g}, central differences {fd}, rel err {err}")
    assert err.max() <= 1e-6
    sim.close()


def test_lj_adjoint_energy_shift_and_batch(oracle_mod):
    """dL/dE = 1 alone gives dE/d(epsilon, sigma), shift term included: against central
    differences of the oracle's energy on one box, and for a batch of two boxes (as
    replicas) the sum of the boxes' gradients."""
    from exastamp_amd import inputs
    from exastamp_amd.operators import Simulation
    from exastamp_amd.units import EV
    boxes = [_config((6, 6, 6), seed=s) for s in (23, 24)]
    eps, sig = 0.0104 * EV, 3.4
    sims, grads = [], []
    for b in boxes:
        sim = Simulation(0)
        sim.species([("Ar", AR)])
        sim.set_config(b)
        sim.chunk_neighbors(rcut_max=RC, rcut_inc=0.3)
        sim.lj_parameters(parameters={"epsilon": eps, "sigma": sig}, rcut=RC)
        grads.append(sim.lj_adjoint(np.zeros((b.n, 3)), np.ones(1)))
        sims.append(sim)
    for b, g in zip(boxes, grads):
        fd = []
        for k, x0 in enumerate((eps, sig)):
            h = 1e-6 * x0
            es = []
            for s in (+1, -1):
                th = [eps, sig]
                th[k] = x0 + s * h
                es.append(_oracle_fe(oracle_mod, b, th[0], th[1])[1][0] * EV)
            fd.append((es[0] - es[1]) / (2 * h))
        assert np.allclose(g, fd, rtol=1e-7)
    batch = inputs.Config(np.concatenate([b.r for b in boxes]), np.concatenate([b.v for b in boxes]),
                          np.concatenate([b.type for b in boxes]), np.concatenate([b.lo for b in boxes]),
                          np.concatenate([b.hi for b in boxes]), (1, 1, 1),
                          np.repeat(np.arange(2), [b.n for b in boxes]).astype(np.uint16))
    sim = Simulation(0)
    sim.species([("Ar", AR)])
    sim.set_config(batch)
    sim.chunk_neighbors(rcut_max=RC, rcut_inc=0.3)
    sim.lj_parameters(parameters={"epsilon": eps, "sigma": sig}, rcut=RC)
    gb = sim.lj_adjoint(np.zeros((batch.n, 3)), np.ones(2))
    assert np.allclose(gb, grads[0] + grads[1], rtol=1e-12)
    for s in sims + [sim]:
        s.close()
