"""PyTorch fitting harness for EAM tables, bound over libxsmd's C ABI.

``EamForcesEnergies`` is a torch.autograd.Function: forward loads the knot vector
theta (setfl order: F tables, rho tables, r*phi tables) into the context, runs
eam_alloy_force on the GPU and returns forces (n, 3) and per-replica energies in
eV units; backward calls the hand-written adjoint kernels (xs_eam_adjoint) with
the incoming gradients and returns dL/dtheta.  The forward's rho_i and F'(rho_i)
are the checkpoint the adjoint reads, so the system must not move between the
two calls.

    sim = Simulation(); sim.species(...); sim.set_config(batch); sim.eam_alloy_init(tables)
    sim.chunk_neighbors(rcut_max=tables.rc, rcut_inc=0.3)
    theta = torch.tensor(tables.theta(), requires_grad=True)
    F, E = EamForcesEnergies.apply(theta, sim)      # FP64 forward, FP32-table adjoint
    loss = fit_loss(F, E, F_target, E_target, atoms_per_replica)
    loss.backward()

With theta on the GPU the step stays on the device: the context runs on torch's
current stream (xs_ctx_set_stream), the spline coefficients are built from theta by a
kernel (xs_eam_theta_dev), forces and replica energies come back as device tensors
(xs_forces_energies_dev) and the adjoint writes dL/dtheta into a device tensor
(xs_eam_adjoint_dev); nothing is copied through the host.  On that path the mixed
forward reads the FP32 knot rows (the s = r^2 tables need the host's accuracy check).

forward_mixed selects the forward precision (FP32 pair/spline math), adjoint_mixed
the FP32 table copies in the adjoint sweeps.  A force-matching gradient is only as
accurate as the residual F - F_target it is built from: with an FP32 forward the
force error (~1e-6 of |F|) divided by the residual size bounds the gradient error
(6e-4 - 9e-4 relative to FP64 central differences for targets 2 % away).  The
default is therefore an FP64 forward with the FP32-table adjoint, whose gradients
agree with FP64 central differences to a few 1e-7; the forward is the small part
of a fitting step (the adjoint sweeps dominate).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import XS_MIXED
from .units import EV


def _on_torch_stream(sim, device):
    sim.ctx.xs_ctx_set_stream(C.c_void_p(torch.cuda.current_stream(device).cuda_stream))


class EamForcesEnergies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, sim, forward_mixed=False, adjoint_mixed=True):
        ctx.sim = sim
        ctx.adjoint_mixed = adjoint_mixed
        ctx.device = theta.is_cuda
        if theta.is_cuda:
            _on_torch_stream(sim, theta.device)
            th = theta.detach().to(torch.float64).contiguous()
            if th.numel() != sim.eam_knots:
                raise ValueError("theta does not match the loaded tables' knot count")
            sim.ctx.xs_eam_theta_dev(C.c_void_p(th.data_ptr()))
            sim.zero_force_energy()
            sim.eam_alloy_force(rcut=sim.eam.rc, trigger_thermo_state=True, mixed=forward_mixed)
            sim.force_to_accel()
            F = torch.empty((sim.n, 3), dtype=torch.float64, device=theta.device)
            E = torch.empty(sim.cfg.nrep, dtype=torch.float64, device=theta.device)
            sim.ctx.xs_forces_energies_dev(C.c_void_p(F.data_ptr()), C.c_void_p(E.data_ptr()), 1.0 / EV)
            return F, E
        t = sim.eam.with_theta(theta.detach().to("cpu", torch.float64).numpy())
        sim.eam_alloy_init(t)
        sim.zero_force_energy()
        sim.eam_alloy_force(rcut=t.rc, trigger_thermo_state=True, mixed=forward_mixed)
        sim.force_to_accel()
        F = torch.from_numpy(sim.forces() / EV).to(theta.device)
        E = torch.from_numpy(sim.replica_energies() / EV).to(theta.device)
        return F, E

    @staticmethod
    def backward(ctx, gF, gE):
        # outputs are in eV units: dL/dF_internal = dL/dF_eV / EV, likewise for E
        sim = ctx.sim
        if ctx.device:
            _on_torch_stream(sim, gF.device)
            gf = gF.detach().to(torch.float64).contiguous()
            ge = gE.detach().to(torch.float64).contiguous()
            grad = torch.empty(sim.eam_knots, dtype=torch.float64, device=gF.device)
            sim.ctx.xs_eam_adjoint_dev(float(sim.eam.rc), C.c_void_p(gf.data_ptr()), C.c_void_p(ge.data_ptr()),
                                       1.0 / EV, C.c_void_p(grad.data_ptr()), XS_MIXED if ctx.adjoint_mixed else 0)
            return grad, None, None, None
        gf = gF.detach().to("cpu", torch.float64).numpy() / EV
        ge = gE.detach().to("cpu", torch.float64).numpy() / EV
        g = ctx.sim.eam_adjoint(gf, ge, mixed=ctx.adjoint_mixed)
        return torch.from_numpy(g).to(gF.device), None, None, None


class LjForcesEnergies(torch.autograd.Function):
    """theta = (epsilon in eV, sigma in A) -> forces (n, 3) in eV/A and per-replica
    energies in eV of lj_compute_force at rcut; backward calls the hand-written LJ
    adjoint (xs_lj_adjoint: one FP64 sweep of the list, the energy shift's dependence
    on epsilon and sigma included).  The system must not move between the calls."""

    @staticmethod
    def forward(ctx, theta, sim, rcut, mixed=False):
        eps_ev, sigma = (float(x) for x in theta.detach().to("cpu", torch.float64))
        params = {"epsilon": eps_ev * EV, "sigma": sigma}
        sim.zero_force_energy()
        sim.lj_compute_force(parameters=params, rcut=rcut, compute_virial=False, trigger_thermo_state=True,
                             mixed=mixed)
        sim.force_to_accel()
        F = torch.from_numpy(sim.forces() / EV).to(theta.device)
        E = torch.from_numpy(sim.replica_energies() / EV).to(theta.device)
        ctx.sim, ctx.params, ctx.rcut = sim, params, rcut
        return F, E

    @staticmethod
    def backward(ctx, gF, gE):
        gf = gF.detach().to("cpu", torch.float64).numpy() / EV
        ge = gE.detach().to("cpu", torch.float64).numpy() / EV
        g = ctx.sim.lj_adjoint(gf, ge, parameters=ctx.params, rcut=ctx.rcut)
        g[0] *= EV   # d/d(epsilon in eV) = EV * d/d(epsilon in internal units)
        return torch.from_numpy(g).to(gF.device), None, None, None


class _FitLoss(torch.autograd.Function):
    """fit_loss as one autograd node: L = mean(rF^2) + w mean(rE^2) with rF = F - F_t and
    rE = (E - E_t) / n; dL/dF = 2 rF / size(rF), dL/dE = 2 w rE / (n size(rE)).  One node
    instead of the dozen of the elementwise form keeps a device fitting step from being
    bound by the autograd engine's per-node work on the host."""

    @staticmethod
    def forward(ctx, F, E, F_target, E_target, n, energy_weight):
        rF = F - F_target
        rE = (E - E_target) / n
        ctx.save_for_backward(rF, rE, n)
        ctx.energy_weight = energy_weight
        return (rF * rF).mean() + energy_weight * (rE * rE).mean()

    @staticmethod
    def backward(ctx, g):
        rF, rE, n = ctx.saved_tensors
        gF = rF * (g * (2.0 / rF.numel()))
        gE = rE / n * (g * (2.0 * ctx.energy_weight / rE.numel()))
        return gF, gE, None, None, None, None


def fit_loss(F, E, F_target, E_target, atoms_per_replica, energy_weight=0.1):
    """Force MSE (eV/A)^2 over all components plus energy_weight * MSE of the
    per-atom energy (eV) over replicas.  In a device loop give atoms_per_replica as a
    tensor on the device: a host array is copied (synchronously) at every call.
    Gradients flow to F and E (not to the targets)."""
    n = torch.as_tensor(atoms_per_replica, dtype=F.dtype, device=F.device)
    return _FitLoss.apply(F, E, F_target.detach(), E_target.detach(), n, energy_weight)


def fit_loss_np(F, E, F_target, E_target, atoms_per_replica, energy_weight=0.1):
    n = np.asarray(atoms_per_replica, dtype=np.float64)
    return float(((F - F_target) ** 2).mean() + energy_weight * (((E - E_target) / n) ** 2).mean())
