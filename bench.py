#!/usr/bin/env python3
"""EAM Cu NVE throughput on one MI355X: mixed precision vs plain FP64.

Workload (the headline configuration): FCC copper, a = 3.615 A, 128^3 cubic cells
= 8 388 608 atoms, periodic; synthetic Johnson-type setfl tables with 5000 knots
(rc = 4.95 A); neighbour skin 0.3 A; Maxwell-Boltzmann velocities at 600 K from a
seed; velocity-Verlet NVE with dt = 1 fs.  A step is one NVE step of the whole
system (drift/kick, displacement check and list rebuild when triggered, EAM
density pass + embedding, EAM force pass + force->acceleration + kick).

Prints ONE JSON line.  value = atom-steps/s of the mixed-precision form over the
K timed steps (inputs resident in HBM, list rebuilds included, setup excluded).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "atom-steps/s EAM Cu 8.4M atoms, mixed vs FP64 GPU, at ≤1e-5 rel force error"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
# Algorithmic per-atom figures of the EAM force pass as fused here (force pass +
# force_to_accel + half kick): bytes of the compulsory FP64 state stream (read r,
# F', v; write a, v) and FLOPs per interacting pair / per atom.  In the MD loop the
# pass also applies the next step's drift and kick (positions written, reference
# positions of the displacement test read): 48 B more, 18 FLOP more.
FORCE_BYTES_PER_ATOM = 104.0
DRIFT_BYTES_PER_ATOM = 48.0
FORCE_FLOPS_PER_PAIR = 49.0
FORCE_FLOPS_PER_ATOM = 3.0 + 6.0
DRIFT_FLOPS_PER_ATOM = 12.0 + 6.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=128, help="FCC cells per edge (128 -> 8.39M atoms)")
    ap.add_argument("--fp64-steps", type=int, default=None, help="timed steps of the FP64 form (default: --steps)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-1t-steps", type=int, default=1, help="steps of the single-thread CPU sample (0: skip)")
    ap.add_argument("--drift-steps", type=int, default=1000,
                    help="NVE steps of the energy-drift leg of each form (thermo every 10 steps, linear fit; 0: skip)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-symmetric", action="store_true", help="skip the full-list vs Newton-symmetric operator timing")
    ap.add_argument("--no-lj", action="store_true", help="skip the LJ argon leg")
    ap.add_argument("--no-c2", action="store_true",
                    help="skip the 500 000-atom FP64 GPU-vs-CPU leg (EAM Cu 50^3 cells, 100 steps)")
    ap.add_argument("--lj-cells", type=int, default=64, help="FCC cells per edge of the LJ leg (64 -> 1.05M atoms)")
    ap.add_argument("--stream-copy-gib", type=float, default=0.0,
                    help="also time a plain HBM copy of this many GiB (the counter calibration kernel)")
    ap.add_argument("--seed", type=int, default=2024)
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def setup(sim, cfg, tab):
    sim.species([("Cu", 63.546)])
    sim.set_config(cfg)
    sim.eam_alloy_init(tab)


def energy(sim, rc, mixed):
    """Total energy at the current state (recomputes forces, outside timed regions)."""
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=rc, trigger_thermo_state=True, compute_virial=True, mixed=mixed)
    sim.force_to_accel()
    return sim.simulation_thermodynamic_state()


def run_form(sim, cfg, tab, mixed, steps, warmup, dt, rc, skin):
    """Warm-up, then `steps` timed NVE steps (no profiling inside the timed region),
    then the same number of steps again with per-kernel HIP-event timers on."""
    from exastamp_amd import _lib
    import ctypes as C
    sim.set_config(cfg)
    st_w, _ = sim.md_run("eam", warmup, dt, rc, skin, mixed=mixed)
    th0 = energy(sim, rc, mixed)
    flags = (_lib.XS_MIXED if mixed else 0) | _lib.XS_CONTINUE
    sim.ctx.xs_synchronize()
    t0 = time.perf_counter()
    st = _lib.RunStats()
    sim.ctx.xs_md_run(_lib.XS_POT_EAM, steps, dt, rc, skin, flags, 0, None, C.byref(st))
    sim.ctx.xs_synchronize()
    t1 = time.perf_counter()
    sim.ctx.xs_profile_enable(1)
    st_p = _lib.RunStats()
    sim.ctx.xs_md_run(_lib.XS_POT_EAM, steps, dt, rc, skin, flags, 0, None, C.byref(st_p))
    sim.ctx.xs_synchronize()
    ms = np.zeros(len(_lib.PROF_CLASSES))
    cnt = np.zeros(len(_lib.PROF_CLASSES), dtype=np.int64)
    sim.ctx.xs_profile_get(_lib.dptr(ms), _lib.ptr(cnt, C.c_int64))
    sim.ctx.xs_profile_enable(0)
    thK = energy(sim, rc, mixed)
    n = cfg.n
    sec = t1 - t0
    kern = {name: {"ms_total": float(ms[k]), "launches": int(cnt[k]),
                   "ms_avg": float(ms[k] / cnt[k]) if cnt[k] else None}
            for k, name in enumerate(_lib.PROF_CLASSES)}
    return {
        "value": n * steps / sec, "seconds": sec, "ms_per_step": 1e3 * sec / steps,
        "rebuilds": int(st.rebuilds), "kernels": kern, "kernels_run_rebuilds": int(st_p.rebuilds),
        "T0": th0.temperature, "T_end": thK.temperature,
        "warmup_rebuilds": int(st_w.rebuilds),
    }


C2_CELLS, C2_STEPS = 50, 100


def c2_config(seed):
    """EAM Cu FCC 50^3 cells = 500 000 atoms, periodic, seeded 0.05 A Gaussian jitter,
    zero initial velocities (the plain-FP64 GPU-vs-CPU configuration)."""
    from exastamp_amd import inputs
    return inputs.lattice_config((C2_CELLS,) * 3, 3.615, [63.546], T=0.0, jitter=0.05, seed=seed)


def c2_leg(tab, dt, rc, skin, seed):
    """The 500 000-atom FP64 configuration on the GPU: list build and first forces
    untimed (as the CPU sample's setup), then 100 NVE steps of 1 fs timed."""
    from exastamp_amd import _lib
    from exastamp_amd.operators import Simulation
    import ctypes as C
    cfg = c2_config(seed)
    sim = Simulation(0)
    setup(sim, cfg, tab)
    sim.md_run("eam", 5, dt, rc, skin)  # warm-up (kernels, buffers)
    sim.set_config(cfg)
    sim.md_run("eam", 0, dt, rc, skin)  # first_iteration: list and forces at t = 0
    sim.ctx.xs_synchronize()
    st = _lib.RunStats()
    t0 = time.perf_counter()
    sim.ctx.xs_md_run(_lib.XS_POT_EAM, C2_STEPS, dt, rc, skin, _lib.XS_CONTINUE, 0, None, C.byref(st))
    sim.ctx.xs_synchronize()
    sec = time.perf_counter() - t0
    sim.close()
    return {"workload": f"EAM Cu FCC {C2_CELLS}^3 cells = {cfg.n} atoms, 0.05 A jitter, 0 K, FP64, "
                        f"{C2_STEPS} NVE steps of 1 fs",
            "gpu_fp64": {"value": cfg.n * C2_STEPS / sec, "unit": "atom-steps/s", "seconds": sec,
                         "rebuilds": int(st.rebuilds)}}


def drift_leg(sim, cfg, tab, mixed, steps, dt, rc, skin):
    """NVE energy drift of one form: `steps` steps from the initial state with the
    total energy every 10 steps, slope of a least-squares line (eV/atom/ps)."""
    from exastamp_amd.units import EV
    sim.set_config(cfg)
    _, th = sim.md_run("eam", steps, dt, rc, skin, mixed=mixed, thermo_every=10)
    t = np.arange(len(th)) * 10 * dt
    e = np.array([x.total_energy for x in th]) / EV / cfg.n
    slope = float(np.polyfit(t, e, 1)[0])
    return {"ev_per_atom_per_ps": slope, "steps": steps, "samples": len(th),
            "rms_fluctuation_ev_per_atom": float(np.std(e - np.polyval(np.polyfit(t, e, 1), t)))}


def gather_ceiling(sim, rc, kernels):
    """Gather-ceiling kernels of both passes (their data movement without the spline
    arithmetic) against the passes' own average times in the MD loop.  The force pass's
    ceiling models its s = r^2 table rows, so it is left out when the tables are off."""
    out = {}
    for name, force in (("density", False), ("force", True)):
        if force and sim.eam_table_info()["s_pieces"] == 0:
            out[name] = None
            continue
        ms = sim.gather_ceiling(rc, force, 10)
        k = kernels["eam_force" if force else "eam_rho"]["ms_avg"]
        out[name] = {"ceiling_ms": ms, "pass_ms": k, "frac": ms / k if k else None,
                     # the same kernel with parts taken away: what each part costs
                     "without_table_rows_ms": sim.gather_ceiling(rc, force, 10, level=1),
                     "list_words_only_ms": sim.gather_ceiling(rc, force, 10, level=2),
                     "staging_only_ms": sim.gather_ceiling(rc, force, 10, level=3)}
    return out


def force_kernel_name(list_info):
    if list_info["kind"] != "block":
        return "k_eam_force<float, false, false, false, true>"
    return "k_block_eam_force<float, false, false, false, true, true>"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed counter
    summary (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            parts = line.strip().strip("`").split("\t")
            if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == kernel:
                return float(parts[2])
    return None


def force_error(sim, rc):
    """Mixed vs FP64 forces on the current positions."""
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=rc, mixed=False)
    sim.force_to_accel()
    f64 = sim.accelerations()
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=rc, mixed=True)
    sim.force_to_accel()
    f32 = sim.accelerations()
    mag = np.sqrt((f64 ** 2).sum(1))
    rms = float(np.sqrt((mag ** 2).mean()))
    d = np.sqrt(((f32 - f64) ** 2).sum(1))
    sel = mag > 0.1 * rms
    return {
        "max_abs_over_rms": float(d.max() / rms),
        "rms_abs_over_rms": float(np.sqrt((d ** 2).mean()) / rms),
        "max_rel_atoms_above_0.1rms": float((d[sel] / mag[sel]).max()),
        "metric": "max_i |F_mixed,i - F_fp64,i| / rms_j |F_fp64,j|",
    }


def _fp64_forces_energies(sim, tab):
    """FP64 forces (eV/A, caller's atom order) and per-replica energies (eV) of the GPU
    form with tables `tab` (the FP64 path equals the oracle to 1e-11, tests/)."""
    from exastamp_amd.units import EV
    sim.eam_alloy_init(tab)
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, mixed=False)
    sim.force_to_accel()
    return sim.forces() / EV, sim.replica_energies() / EV


def adjoint_leg(name, cfg, masses, names, tab, nfd, repeats):
    """Forward + adjoint dLoss/dtheta on one batch: GPU time per forward+backward,
    device memory, and gradient error vs FP64 central differences of the FP64 forward
    (targets: the FP64 forward with tables perturbed by a seeded +-2 %).
    Loss = force MSE (eV/A)^2 + 0.1 per-atom energy MSE (eV^2), the gradients of which
    are formed here with numpy (the torch harness in exastamp_amd.fitting does the same)."""
    import ctypes as C
    from exastamp_amd import _lib, inputs
    from exastamp_amd.operators import Simulation
    from exastamp_amd.units import EV

    def fit_loss_np(F, E, F_target, E_target, atoms_per_replica, energy_weight=0.1):
        n = np.asarray(atoms_per_replica, dtype=np.float64)
        return float(((F - F_target) ** 2).mean() + energy_weight * (((E - E_target) / n) ** 2).mean())

    theta0 = tab.theta()
    tab_t = tab.with_theta(theta0 * (1.0 + inputs.rng(5).uniform(-0.02, 0.02, theta0.size)))
    nper = np.bincount(cfg.rep) if cfg.rep is not None else np.array([cfg.n])
    sim = Simulation(0)
    sim.species(list(zip(names, masses)))
    sim.set_config(cfg)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    Ft, Et = _fp64_forces_energies(sim, tab_t)

    def step(fwd_mixed, adj_mixed):
        sim.eam_alloy_init(tab.with_theta(theta0))
        sim.zero_force_energy()
        sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, mixed=fwd_mixed)
        sim.force_to_accel()
        F = sim.forces() / EV
        E = sim.replica_energies() / EV
        gF = 2.0 * (F - Ft) / F.size / EV
        gE = 0.1 * 2.0 * (E - Et) / nper ** 2 / E.size / EV
        return sim.eam_adjoint(gF, gE, mixed=adj_mixed)

    grads = {("mixed", "mixed"): step(True, True), ("fp64", "fp64"): step(False, False),
             ("fp64", "mixed"): step(False, True)}
    # the fitting step as a training loop runs it: theta, targets, forces, loss and
    # gradient on the device (torch harness, device path), the library on torch's stream
    import torch
    from exastamp_amd.fitting import EamForcesEnergies, fit_loss
    dev = torch.device("cuda", 0)
    theta_d = torch.tensor(theta0, device=dev, requires_grad=True)
    Ft_d, Et_d = torch.tensor(Ft, device=dev), torch.tensor(Et, device=dev)
    nper_d = torch.tensor(nper, dtype=torch.float64, device=dev)

    def dstep(fwd_mixed):
        theta_d.grad = None
        F, E = EamForcesEnergies.apply(theta_d, sim, fwd_mixed, True)
        fit_loss(F, E, Ft_d, Et_d, nper_d).backward()

    wall, gpu_ms = {}, {}
    for label, fwd_mixed in (("fp64_forward", False), ("mixed_forward", True)):
        for _ in range(20):  # clocks, torch's allocator and the library's buffers settle
            dstep(fwd_mixed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()  # wall time: no per-launch event timers in this loop
        for _ in range(repeats):
            dstep(fwd_mixed)
        torch.cuda.synchronize()
        wall[label] = 1e3 * (time.perf_counter() - t0) / repeats
        sim.ctx.xs_profile_enable(1)  # GPU time per kernel class: the same steps again
        for _ in range(repeats):
            dstep(fwd_mixed)
        torch.cuda.synchronize()
        ms = np.zeros(len(_lib.PROF_CLASSES))
        cnt = np.zeros(len(_lib.PROF_CLASSES), dtype=np.int64)
        sim.ctx.xs_profile_get(_lib.dptr(ms), _lib.ptr(cnt, C.c_int64))
        sim.ctx.xs_profile_enable(0)
        gpu_ms[label] = {k: float(ms[i] / repeats) for i, k in enumerate(_lib.PROF_CLASSES) if cnt[i]}
    dstep(False)
    g_dev = theta_d.grad.cpu().numpy()
    g_host = grads[("fp64", "mixed")]
    dev_vs_host = float(np.abs(g_dev - g_host).max() / np.abs(g_host).max())
    sim.ctx.xs_ctx_own_stream()
    # without the rho/F' checkpoint the backward would first redo the density pass and
    # the embedding (F' is the force pass's input, so only rho's 8 bytes per atom are
    # the checkpoint's own): time that recomputation
    sim.ctx.xs_profile_enable(1)
    for _ in range(repeats):
        sim.eam_alloy_force(rcut=tab.rc, eam_rho=True, eam_rho2emb=True, eam_force=False, mixed=True)
    sim.ctx.xs_synchronize()
    sim.ctx.xs_profile_get(_lib.dptr(ms), _lib.ptr(cnt, C.c_int64))
    sim.ctx.xs_profile_enable(0)
    recompute_ms = float(ms[_lib.PROF_CLASSES.index("eam_rho")] / repeats)
    g64 = grads[("fp64", "fp64")]
    active = np.flatnonzero(np.abs(g64) > 1e-3 * np.abs(g64).max())
    pick = np.random.default_rng(3).choice(active, size=min(nfd, active.size), replace=False)
    h = 1e-6
    fd = []
    for k in pick:
        v = []
        for sgn in (+1, -1):
            th = theta0.copy()
            th[k] += sgn * h
            F, E = _fp64_forces_energies(sim, tab.with_theta(th))
            v.append(fit_loss_np(F, E, Ft, Et, nper))
        fd.append((v[0] - v[1]) / (2 * h))
    fd = np.array(fd)
    err = {f"fwd_{a}_adj_{b}": float((np.abs(g[pick] - fd) / np.abs(fd)).max()) for (a, b), g in grads.items()}
    free = C.c_size_t(0)
    total = C.c_size_t(0)
    sim.ctx.xs_device_memory(C.byref(free), C.byref(total))
    sim.close()
    return {
        "workload": name, "atoms": int(cfg.n), "replicas": int(cfg.nrep), "knots": int(theta0.size),
        "ms_forward_backward_wall": wall, "gpu_ms_per_forward_backward": gpu_ms,
        "harness": "torch autograd, theta/targets/forces/loss/gradient on the device (xs_eam_theta_dev, "
                   "xs_forces_energies_dev, xs_eam_adjoint_dev on torch's stream); FP32-table adjoint",
        "device_vs_host_grad_max_rel": dev_vs_host,
        "grad_rel_err_vs_fd_max": err, "fd_knots": int(len(pick)), "fd_h": h,
        "device_used_bytes": int(total.value - free.value),
        "checkpoint": {
            "with": {"peak_device_bytes": int(total.value - free.value), "checkpoint_bytes": int(16 * cfg.n),
                     "note": "rho (8 B/atom) and F'(rho) (8 B/atom, the w of the position record) kept from the forward"},
            "without": {"peak_device_bytes": int(total.value - free.value - 8 * cfg.n),
                        "extra_gpu_ms_per_backward": recompute_ms,
                        "note": "rho recomputed by a density pass before the backward; F' must exist between the "
                                "forward's two passes either way, so 8 B/atom are saved"},
        },
    }


def lj_leg(cells, steps, warmup):
    """LJ argon (the C1 parameters: a = 5.26 A, eps = 0.0104 eV, sigma = 3.4 A, rc = 8.5 A,
    skin 0.3 A, 100 K, dt = 2 fs) at cells^3 FCC cells: NVE atom-steps/s of both forms."""
    from exastamp_amd import _lib, inputs
    from exastamp_amd.operators import Simulation
    from exastamp_amd.units import EV
    import ctypes as C
    eps, sig, rc, skin, dt = 0.0104 * EV, 3.4, 8.5, 0.3, 0.002
    cfg = inputs.lattice_config((cells, cells, cells), 5.26, [39.948], T=100.0, seed=7)
    sim = Simulation(0)
    sim.species([("Ar", 39.948)])
    out = {"workload": f"LJ Ar FCC {cells}^3 cells = {cfg.n} atoms, rc 8.5 A, skin 0.3 A, NVE dt 2 fs",
           "atoms": cfg.n, "steps": steps}
    for name, mixed in (("mixed", True), ("fp64", False)):
        sim.set_config(cfg)
        sim.lj_parameters(parameters={"epsilon": eps, "sigma": sig}, rcut=rc)
        sim.md_run("lj", warmup, dt, rc, skin, mixed=mixed)
        sim.ctx.xs_profile_enable(1)
        sim.ctx.xs_synchronize()
        t0 = time.perf_counter()
        st = _lib.RunStats()
        flags = (_lib.XS_MIXED if mixed else 0) | _lib.XS_CONTINUE
        sim.ctx.xs_md_run(_lib.XS_POT_LJ, steps, dt, rc, skin, flags, 0, None, C.byref(st))
        sim.ctx.xs_synchronize()
        sec = time.perf_counter() - t0
        ms = np.zeros(len(_lib.PROF_CLASSES))
        cnt = np.zeros(len(_lib.PROF_CLASSES), dtype=np.int64)
        sim.ctx.xs_profile_get(_lib.dptr(ms), _lib.ptr(cnt, C.c_int64))
        sim.ctx.xs_profile_enable(0)
        k = {c: i for i, c in enumerate(_lib.PROF_CLASSES)}
        out[name] = {"value": cfg.n * steps / sec, "rebuilds": int(st.rebuilds),
                     "lj_ms_avg": float(ms[k["lj"]] / max(cnt[k["lj"]], 1))}
    out["list_kind"] = sim.list_info()["kind"]
    sim.close()
    return out


def cpu_baseline(cfg, tab, steps, threads, dt, rc, skin):
    """The FP64 oracle (OpenMP, full list, the reference's CPU structure) on the same
    system for a bounded number of steps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    if threads <= 0:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    o = pyoracle.Oracle(threads=threads)
    o.set_config(cfg, [63.546])
    o.set_eam(tab)
    t0 = time.perf_counter()
    o.build_neighbors(rc, skin)
    o.compute_forces("eam", 0)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    nreb = o.md_steps("eam", steps, dt, rc, skin)
    sec = time.perf_counter() - t0
    o.close()
    return {
        "value": cfg.n * steps / sec, "unit": "atom-steps/s", "cores": threads, "kind": "port",
        "sample": f"FP64 oracle, same {cfg.n}-atom system, {steps} NVE steps after setup "
                  f"({t_setup:.1f} s setup excluded), {nreb} rebuilds in the sample",
        "seconds": sec,
    }


def operator_forms(sim, rc, reps=5):
    """The eam_alloy_force operator (density, embedding and force passes, mixed) in
    its full-list form and in its Newton-symmetric form (half list, FP64 atomics):
    GPU time per call of each pass class, and the symmetric form's force deviation."""
    from exastamp_amd import _lib
    import ctypes as C
    out = {}
    forces = {}
    for name, sym in (("full_list", False), ("symmetric", True)):
        sim.zero_force_energy()
        sim.eam_alloy_force(rcut=rc, mixed=True, eam_symmetry=sym)   # untimed: builds the half list
        forces[name] = sim.accelerations()
        sim.ctx.xs_synchronize()
        sim.ctx.xs_profile_enable(1)
        for _ in range(reps):
            sim.zero_force_energy()
            sim.eam_alloy_force(rcut=rc, mixed=True, eam_symmetry=sym)
        sim.ctx.xs_synchronize()
        ms = np.zeros(len(_lib.PROF_CLASSES))
        cnt = np.zeros(len(_lib.PROF_CLASSES), dtype=np.int64)
        sim.ctx.xs_profile_get(_lib.dptr(ms), _lib.ptr(cnt, C.c_int64))
        sim.ctx.xs_profile_enable(0)
        k = {c: i for i, c in enumerate(_lib.PROF_CLASSES)}
        rho, force = ms[k["eam_rho"]] / reps, ms[k["eam_force"]] / reps
        out[name] = {"density_ms": float(rho), "force_ms": float(force), "total_ms": float(rho + force)}
    f0, f1 = forces["full_list"], forces["symmetric"]
    rms = float(np.sqrt((f0 ** 2).sum(1).mean()))
    out["symmetric_max_dev_over_rms"] = float(np.sqrt(((f1 - f0) ** 2).sum(1)).max() / rms)
    out["note"] = ("operator calls outside the timed MD loop; symmetric = each pair once from a half list, "
                   "partner share by FP64 atomics on the global list")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this library runs the force "
              "path on one GPU (several GPUs would each own a spatial domain of bricks with ghost shells).",
              file=sys.stderr)
        sys.exit(2)
    from exastamp_amd import inputs, potentials
    from exastamp_amd.operators import Simulation

    dt, rc, skin = 0.001, 4.95, 0.3
    nc = args.cells
    log(f"building {nc}^3-cell Cu lattice + 5000-knot tables")
    tab = potentials.johnson_tables(nr=5000, nrho=5000, rc=rc)
    cfg = inputs.lattice_config((nc, nc, nc), 3.615, [63.546], T=600.0, jitter=0.0, seed=args.seed)
    sim = Simulation(0)
    setup(sim, cfg, tab)
    s_table = sim.eam_table_info()

    log(f"mixed precision: {args.warmup} warmup + {args.steps} timed steps, {cfg.n} atoms")
    mixed = run_form(sim, cfg, tab, True, args.steps, args.warmup, dt, rc, skin)
    ferr = force_error(sim, rc)
    list_info = sim.list_info()
    entries, within = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
    import ctypes as C
    sim.ctx.xs_pair_count(rc, entries.ctypes.data_as(C.POINTER(C.c_int64)),
                          within.ctypes.data_as(C.POINTER(C.c_int64)))
    pairs_per_atom = float(within[0]) / cfg.n
    log(f"mixed: {mixed['value']:.4g} atom-steps/s, force err {ferr['max_abs_over_rms']:.2e}")
    gather = gather_ceiling(sim, rc, mixed["kernels"]) if list_info["kind"] == "block" else None
    if gather:
        for name, g in gather.items():
            if g:
                log("gather ceiling, %s pass: %.3f of %.3f ms" % (name, g["ceiling_ms"], g["pass_ms"]))

    sym = None
    if not args.no_symmetric:
        sym = operator_forms(sim, rc)
        log(f"operator eam_alloy_force: full list {sym['full_list']['total_ms']:.3f} ms, "
            f"symmetric {sym['symmetric']['total_ms']:.3f} ms")

    fp64 = None
    if not args.no_fp64:
        k64 = args.fp64_steps or args.steps
        log(f"FP64: {args.warmup} warmup + {k64} timed steps")
        fp64 = run_form(sim, cfg, tab, False, k64, args.warmup, dt, rc, skin)
        log(f"fp64: {fp64['value']:.4g} atom-steps/s")

    drift = None
    if args.drift_steps > 0:
        log(f"NVE drift legs: {args.drift_steps} steps per form, thermo every 10 steps")
        drift = {"mixed": drift_leg(sim, cfg, tab, True, args.drift_steps, dt, rc, skin)}
        if not args.no_fp64:
            drift["fp64"] = drift_leg(sim, cfg, tab, False, args.drift_steps, dt, rc, skin)
        drift["ratio_mixed_over_fp64"] = (drift["mixed"]["ev_per_atom_per_ps"] / drift["fp64"]["ev_per_atom_per_ps"]
                                          if "fp64" in drift and drift["fp64"]["ev_per_atom_per_ps"] else None)
        log("drift: " + ", ".join(f"{k} {v['ev_per_atom_per_ps']:.3e}" for k, v in drift.items() if isinstance(v, dict)))

    # roofline of the dominant kernel (EAM force pass, mixed)
    kf = mixed["kernels"]["eam_force"]
    kname = force_kernel_name(list_info)
    drift_fused = kname.endswith(", true>")
    bytes_atom = FORCE_BYTES_PER_ATOM + (DRIFT_BYTES_PER_ATOM if drift_fused else 0.0)
    flops_atom = FORCE_FLOPS_PER_ATOM + (DRIFT_FLOPS_PER_ATOM if drift_fused else 0.0)
    avg_s = kf["ms_avg"] * 1e-3
    gbs = bytes_atom * cfg.n / avg_s / 1e9
    tflops = (FORCE_FLOPS_PER_PAIR * pairs_per_atom + flops_atom) * cfg.n / avg_s / 1e12
    roof = {
        "kernel": kname + " (EAM force pass + force_to_accel + half kick"
                  + (" + next-step drift/kick)" if drift_fused else ")"),
        "bound": "gather",
        "bound_note": "the passes are bound by the L1/texture path of their table and neighbour gathers (TD busy "
                      "~95 %); achieved/peak/frac are the compulsory HBM stream against 8 TB/s, gather.frac the "
                      "pass's time explained by its data movement alone (gather-ceiling kernel)",
        "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
        "traffic": measured_traffic(kname),
        "bytes_per_atom": bytes_atom, "avg_launch_ms": kf["ms_avg"],
        "valu": {"achieved_tflops": tflops, "peak_tflops": FP32_PEAK_TFLOPS, "frac": tflops / FP32_PEAK_TFLOPS,
                 "flops_per_pair": FORCE_FLOPS_PER_PAIR, "pairs_per_atom": pairs_per_atom},
        "gather": gather,
        "gather_frac": gather["force"]["frac"] if gather and gather["force"] else None,
    }

    stream = None
    if args.stream_copy_gib > 0:
        g = C.c_double(0.0)
        sim.ctx.xs_stream_copy(int(args.stream_copy_gib * (1 << 30)), 3, C.byref(g))
        stream = {"gib": args.stream_copy_gib, "gbs": g.value}
        log(f"stream copy: {g.value:.1f} GB/s")

    c2 = None
    if not args.no_c2:
        log(f"C2 leg: EAM Cu {C2_CELLS}^3 cells, FP64, {C2_STEPS} steps on the GPU")
        c2 = c2_leg(tab, dt, rc, skin, args.seed)
        log(f"C2: GPU FP64 {c2['gpu_fp64']['value']:.4g} atom-steps/s")

    lj = None
    if not args.no_lj:
        log(f"LJ argon leg: {args.lj_cells}^3 FCC cells, 200 steps per form")
        lj = lj_leg(args.lj_cells, 200, 10)
        log(f"LJ: mixed {lj['mixed']['value']:.4g}, FP64 {lj['fp64']['value']:.4g} atom-steps/s")

    adjoint = None
    if not args.no_adjoint:
        from exastamp_amd import inputs as inp
        sim.close()
        log("adjoint C4: 64 strained Cu boxes x 864 atoms, 3 x 500-knot tables, FD on 32 knots")
        c4 = inp.strained_batch(64, (6, 6, 6), 3.615, [63.546], seed=41)
        a4 = adjoint_leg("C4 batch", c4, [63.546], ["Cu"], potentials.johnson_tables(nr=500, nrho=500), 32, 400)
        log(f"C4: {a4['ms_forward_backward_wall']} ms/forward+backward, FD err {a4['grad_rel_err_vs_fd_max']}")
        log("adjoint C5: Cu/Ni 64^3 FCC slab, 1% vacancies, 7 x 2000-knot tables, FD on 16 knots")
        c5 = inp.alloy_slab_config((64, 64, 64), 3.57, [63.546, 58.693], vacancy_frac=0.01, vacuum=10.0,
                                   T=300.0, seed=42)
        t5 = potentials.johnson_tables((potentials.CU, potentials.NI), nr=2000, nrho=2000)
        a5 = adjoint_leg("C5 Cu/Ni slab", c5, [63.546, 58.693], ["Cu", "Ni"], t5, 16, 30)
        log(f"C5: {a5['ms_forward_backward_wall']} ms/forward+backward, FD err {a5['grad_rel_err_vs_fd_max']}")
        adjoint = {"C4": a4, "C5": a5}

    # The CPU samples run last: a multi-threaded burn of the host cores ahead of the
    # fitting legs would slow the host side of their launch-bound steps.
    cpu = None
    if not args.no_cpu:
        log(f"CPU oracle baseline: {args.cpu_steps} steps")
        cpu = cpu_baseline(cfg, tab, args.cpu_steps, args.cpu_threads, dt, rc, skin)
        log(f"cpu: {cpu['value']:.4g} atom-steps/s on {cpu['cores']} threads")
        if args.cpu_1t_steps > 0:
            log(f"CPU oracle baseline, one thread: {args.cpu_1t_steps} steps")
            one = cpu_baseline(cfg, tab, args.cpu_1t_steps, 1, dt, rc, skin)
            cpu["single_thread"] = {k: one[k] for k in ("value", "cores", "sample", "seconds")}
            log(f"cpu 1 thread: {one['value']:.4g} atom-steps/s")
        if c2 is not None:
            log(f"C2: CPU oracle, {C2_STEPS} steps")
            c2cpu = cpu_baseline(c2_config(args.seed), tab, C2_STEPS, args.cpu_threads, dt, rc, skin)
            cpu["c2"] = {k: c2cpu[k] for k in ("value", "cores", "sample", "seconds")}
            c2["gpu_over_cpu"] = c2["gpu_fp64"]["value"] / c2cpu["value"]
            log(f"C2: CPU {c2cpu['value']:.4g} atom-steps/s on {c2cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": mixed["value"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64 mixed (FP32 pair+spline math; FP64 positions, displacements, accumulation, reductions)",
        "data": "synthetic (seeded FCC lattice + Maxwell-Boltzmann 600 K; synthetic Johnson-type 5000-knot tables)",
        "config": {
            "workload": f"EAM Cu FCC {nc}^3 cells = {cfg.n} atoms, periodic, rc 4.95 A, skin 0.3 A, "
                        f"NVE dt 1 fs, T0 600 K",
            "atoms": cfg.n, "dt_ps": dt, "rcut": rc, "skin": skin, "table_knots": 5000,
            "parallelism": "single GPU",
        },
        "fp64": None if fp64 is None else {k: fp64[k] for k in ("value", "ms_per_step", "rebuilds")},
        "mixed_over_fp64": None if fp64 is None else mixed["value"] / fp64["value"],
        "force_error": ferr,
        "nve_drift_ev_per_atom_per_ps": drift,
        "rebuilds": mixed["rebuilds"],
        "pairs_per_atom_within_rc": pairs_per_atom,
        "list_entries_per_atom": float(entries[0]) / cfg.n,
        "neighbor_list": list_info,
        "fp32_s_table": s_table,
        "kernels_mixed": mixed["kernels"],
        "kernels_fp64": None if fp64 is None else fp64["kernels"],
        "newton_symmetric": sym,
        "lj": lj,
        "c2_fp64_500k": c2,
        "roofline": roof,
        "stream_copy": stream,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
    }
    if adjoint is None:
        sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
