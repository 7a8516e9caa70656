"""Newton-symmetric operators (XS_SYMMETRIC: eam_symmetry: true,
lj_compute_force_symetric) against the full-list operators.

The symmetric form evaluates each pair once, from half of the neighbour list, and
adds the partner's share with FP64 atomics, so per-atom sums differ from the
full-list form by FP64 summation order only (and vary in their last bits from run
to run).  Tolerances: per-atom outputs within 1e-10 of their rms in FP64.  In the
mixed form a single species keeps the pair arithmetic symmetric in i and j
((F'_i + F'_j) rho'(r)), so only the FP32 rounding of rho before the embedding
lookup can move (1e-6); for alloys the FP32 sum F'_i rho'_j + F'_j rho'_i rounds
differently from the two ends of a pair, so the symmetric form is held to the
mixed-precision bar itself (1e-5 of the rms).  Cases include a
box of one cell per axis (each atom sees periodic images of itself and of its
neighbours several times), a two-species slab and a batch of replicas.
"""
import numpy as np
import pytest

from test_gpu_block import CASES, _case, _close
from test_gpu_parity import AR_MASS, frel, lj_c1_config, lj_params, make_oracle, make_sim

pytestmark = pytest.mark.gpu


def _eam_sim(cfg, masses, names, tab):
    sim = make_sim(cfg, masses, names)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    return sim


def _eam_outputs(sim, rc, mixed, symmetric):
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=rc, trigger_thermo_state=True, compute_virial=True, mixed=mixed,
                        eam_symmetry=symmetric)
    return sim.accelerations(), sim.ep(), sim.virial(), sim.rho_dEmb()


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("mixed", [False, True])
def test_eam_symmetric_matches_full(name, mixed):
    cfg, masses, names, tab = _case(name)
    sim = _eam_sim(cfg, masses, names, tab)
    full = _eam_outputs(sim, tab.rc, mixed, False)
    sym = _eam_outputs(sim, tab.rc, mixed, True)
    tol = (1e-5 if len(masses) > 1 else 1e-6) if mixed else 1e-10
    for x, y in zip(sym, full):
        assert _close(x, y, tol)
    # each pair's forces cancel: the net force vanishes to rounding
    f = sym[0]
    assert np.abs(f.sum(0)).max() <= 1e-9 * np.sqrt((f ** 2).sum(1).mean()) * len(f)


@pytest.mark.parametrize("name", ["cu_one_cell", "cuni_slab"])
def test_eam_symmetric_vs_oracle(oracle_mod, name):
    cfg, masses, names, tab = _case(name)
    sim = _eam_sim(cfg, masses, names, tab)
    _eam_outputs(sim, tab.rc, False, True)
    sim.force_to_accel()
    o = make_oracle(oracle_mod, cfg, masses)
    o.set_eam(tab)
    o.build_neighbors(tab.rc, 0.3)
    o.compute_forces("eam")
    assert frel(sim.forces(), o.forces()) <= 1e-11
    ref = o.state()["ep"].sum()
    assert abs(sim.ep().sum() - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("mixed", [False, True])
def test_eam_symmetric_split_passes(mixed):
    """eam_rho, eam_rho2emb and eam_force as separate symmetric calls."""
    cfg, masses, names, tab = _case("cuni_slab")
    sim = _eam_sim(cfg, masses, names, tab)
    full = _eam_outputs(sim, tab.rc, mixed, False)
    sim.zero_force_energy()
    kw = dict(rcut=tab.rc, trigger_thermo_state=True, compute_virial=True, mixed=mixed, eam_symmetry=True)
    sim.eam_alloy_force(eam_rho=True, eam_rho2emb=False, eam_force=False, **kw)
    sim.eam_alloy_force(eam_rho=False, eam_rho2emb=True, eam_force=False, **kw)
    sim.eam_alloy_force(eam_rho=False, eam_rho2emb=False, eam_force=True, **kw)
    for x, y in zip((sim.accelerations(), sim.ep(), sim.virial(), sim.rho_dEmb()), full):
        assert _close(x, y, 1e-5 if mixed else 1e-10)   # two species: see the module notes


@pytest.mark.parametrize("mixed", [False, True])
def test_lj_symmetric_matches_full(mixed):
    eps, sig, rc = lj_params()
    from exastamp_amd import inputs
    cfg = lj_c1_config(seed=31, cells=(7, 6, 5))
    cfg = inputs.Config(cfg.r + inputs.rng(8).normal(scale=0.08, size=cfg.r.shape), cfg.v, cfg.type, cfg.lo, cfg.hi)
    sim = make_sim(cfg, [AR_MASS])
    sim.chunk_neighbors(rcut_max=rc, rcut_inc=0.3)
    out = []
    for sym in (False, True):
        sim.zero_force_energy()
        p = {"epsilon": eps, "sigma": sig}
        if sym:
            sim.lj_compute_force_symetric(parameters=p, rcut=rc, trigger_thermo_state=True, compute_virial=True,
                                          mixed=mixed)
        else:
            sim.lj_compute_force(parameters=p, rcut=rc, trigger_thermo_state=True, compute_virial=True, mixed=mixed)
        out.append((sim.accelerations(), sim.ep(), sim.virial()))
    for x, y in zip(out[1], out[0]):
        assert _close(x, y, 1e-6 if mixed else 1e-10)


def test_md_run_rejects_symmetric_flag():
    from exastamp_amd import _lib
    from conftest import cu_tables
    from test_gpu_parity import CU_MASS, cu_config
    tab = cu_tables()
    sim = make_sim(cu_config((4, 4, 4), seed=2), [CU_MASS])
    sim.eam_alloy_init(tab)
    import ctypes
    st = _lib.RunStats()
    with pytest.raises(_lib.XsError, match="XS_SYMMETRIC"):
        sim.ctx.xs_md_run(_lib.XS_POT_EAM, 2, 0.001, tab.rc, 0.3, _lib.XS_SYMMETRIC, 0, None, ctypes.byref(st))
