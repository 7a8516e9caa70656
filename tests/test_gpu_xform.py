"""A deformed domain (a non-identity xform) on the GPU against the oracle.

Positions are grid coordinates; the xform m maps grid displacements to physical ones
(exaNBody's LinearXForm traversal, pair_potential_impl.hxx:388-413), forces and
velocities are physical, and push_f_v_r's INV_XFORM mode maps the physical step back by
m^-1 (push_vec3_2nd_order_xform_minimization.cpp:34).  The oracle's xform path is pinned
by physics in tests/test_oracle_xform.py.

Tolerances: neighbour lists bit-exact (both builds map each candidate displacement in
the oracle's operation order); FP64 forces within 1e-11 of the rms (the cell-block
passes difference staged physical positions, m r_j - m r_i, which rounds differently
from m (r_j - r_i)); mixed forces 1e-5; NVE energies 1e-9 relative and positions 1e-8 A.
"""
import numpy as np
import pytest

from conftest import cu_tables
from test_gpu_parity import AR_MASS, CU_MASS, cu_config, frel, lj_c1_config, lj_params, make_oracle, make_sim

pytestmark = pytest.mark.gpu

SHEAR = np.array([[1.02, 0.17, -0.06], [0.0, 0.97, 0.11], [0.0, 0.0, 1.01]])


def _rot(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * k @ k


def _env(monkeypatch, kind):
    if kind == "global":
        monkeypatch.setenv("XSMD_LIST", "global")
    else:
        monkeypatch.delenv("XSMD_LIST", raising=False)


def _eam_sim(cfg, tab, xf):
    sim = make_sim(cfg, [CU_MASS])
    sim.set_xform(xf)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    return sim


def _eam_forces(sim, tab, mixed=False):
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, compute_virial=True, mixed=mixed)
    sim.force_to_accel()


def _open_z(cfg):
    """cfg open along z, with two atoms moved below the bottom face and two above the top."""
    from exastamp_amd import inputs
    r = cfg.r.copy()
    by_z = np.argsort(r[:, 2])
    r[by_z[:2], 2] -= np.array([0.3, 0.6])
    r[by_z[-2:], 2] += 1.8 + np.array([0.2, 0.5])  # the top plane is at hi - a/2
    return inputs.Config(r, cfg.v, cfg.type, cfg.lo, cfg.hi, periodic=(1, 1, 0))


@pytest.mark.parametrize("open_z", [False, True])
@pytest.mark.parametrize("kind", ["block", "global"])
def test_sheared_eam_matches_oracle(oracle_mod, monkeypatch, kind, open_z):
    tab = cu_tables()
    cfg = cu_config((9, 8, 7), T=600.0, jitter=0.1, seed=21)
    if open_z:
        cfg = _open_z(cfg)
    _env(monkeypatch, kind)
    sim = _eam_sim(cfg, tab, SHEAR)
    assert sim.list_info()["kind"] == kind
    o = make_oracle(oracle_mod, cfg, [CU_MASS])
    o.set_xform(SHEAR)
    o.set_eam(tab)
    o.build_neighbors(tab.rc, 0.3)
    o.compute_forces("eam")
    assert np.array_equal(sim.order(), o.order())
    for x, y in zip(sim.neighbors(), o.neighbors()):
        assert np.array_equal(x, y)
    _eam_forces(sim, tab)
    ref = o.state()
    assert frel(sim.forces(), o.forces()) <= 1e-11
    assert abs(sim.ep().sum() - ref["ep"].sum()) <= 1e-12 * abs(ref["ep"].sum())
    vir, vref = sim.virial().sum(0), ref["virial"].sum(0)
    assert np.abs(vir - vref).max() <= 1e-10 * np.abs(vref).max()
    _eam_forces(sim, tab, mixed=True)
    assert frel(sim.forces(), o.forces()) <= 1e-5


def _jittered_ar(seed, cells):
    from exastamp_amd import inputs
    cfg = lj_c1_config(seed=seed, cells=cells)
    return inputs.Config(cfg.r + inputs.rng(seed + 90).normal(scale=0.08, size=cfg.r.shape), cfg.v, cfg.type,
                         cfg.lo, cfg.hi)


def test_sheared_lj_matches_oracle(oracle_mod):
    eps, sig, rc = lj_params()
    cfg = _jittered_ar(14, (7, 7, 7))
    sim = make_sim(cfg, [AR_MASS])
    sim.set_xform(SHEAR)
    sim.chunk_neighbors(rcut_max=rc, rcut_inc=0.3)
    assert sim.list_info()["kind"] == "block"
    sim.zero_force_energy()
    sim.lj_compute_force(parameters={"epsilon": "0.0104 eV", "sigma": "3.4 ang"}, rcut="8.5 ang",
                         trigger_thermo_state=True, compute_virial=True)
    sim.force_to_accel()
    o = make_oracle(oracle_mod, cfg, [AR_MASS])
    o.set_xform(SHEAR)
    o.set_lj(eps, sig, rc)
    o.build_neighbors(rc, 0.3)
    o.compute_forces("lj")
    for x, y in zip(sim.neighbors(), o.neighbors()):
        assert np.array_equal(x, y)
    assert frel(sim.forces(), o.forces()) <= 1e-11
    ref = o.state()
    assert abs(sim.ep().sum() - ref["ep"].sum()) <= 1e-12 * abs(ref["ep"].sum())


def test_sheared_eam_nve_matches_oracle(oracle_mod):
    """100 fused NVE steps of 1 fs at 600 K under shear: the MD loop's drift maps each
    step by m^-1 and tests physical displacements for the rebuild."""
    tab = cu_tables()
    cfg = cu_config((8, 8, 8), T=600.0, jitter=0.05, seed=23)
    sim = make_sim(cfg, [CU_MASS])
    sim.set_xform(SHEAR)
    sim.eam_alloy_init(tab)
    st, th = sim.md_run("eam", 100, 0.001, 4.95, 0.3, thermo_every=1)
    o = make_oracle(oracle_mod, cfg, [CU_MASS])
    o.set_xform(SHEAR)
    o.set_eam(tab)
    nreb, th_ref = o.md_run("eam", 100, 0.001, 4.95, 0.3)
    assert st.rebuilds == nreb and nreb >= 2
    pe = np.array([t.potential_energy for t in th])
    ke = np.array([t.kinetic_energy for t in th])
    assert np.abs(pe - th_ref[:, 24]).max() <= 1e-9 * np.abs(th_ref[:, 24]).max()
    assert np.abs(ke - th_ref[:, 21:24].sum(1)).max() <= 1e-9 * th_ref[:, 21:24].sum(1).max()
    r, _ = sim.state()
    assert np.abs(r - o.state()["r"]).max() < 1e-8


@pytest.mark.parametrize("mode", ["INV_XFORM", "IDENTITY"])
def test_push_f_v_r_modes_match_oracle(oracle_mod, mode):
    from exastamp_amd import inputs
    cfg = cu_config((5, 5, 5), T=600.0, jitter=0.05, seed=24)
    cfg = inputs.Config(cfg.r, cfg.v * 20.0, cfg.type, cfg.lo, cfg.hi)
    sim = make_sim(cfg, [CU_MASS])
    sim.set_xform(SHEAR)
    o = make_oracle(oracle_mod, cfg, [CU_MASS])
    o.set_xform(SHEAR)
    sim.chunk_neighbors(rcut_max=4.95, rcut_inc=0.3)  # the reference positions of the displacement test
    o.build_neighbors(4.95, 0.3)
    sim.push_f_v_r(0.002, xform_mode=mode)
    o.push_f_v_r(0.002, xform_mode=mode)
    r, _ = sim.state()
    assert np.abs(r - o.state()["r"]).max() <= 1e-12
    L = cfg.hi[0] - cfg.lo[0]
    step = r - cfg.r
    step -= L * np.rint(step / L)  # the build wrapped atoms that started outside the box
    want = cfg.v * 0.002 if mode == "IDENTITY" else (cfg.v * 0.002) @ np.linalg.inv(SHEAR).T
    assert np.abs(step - want).max() <= 1e-12
    assert sim.particle_displ_over(0.05) == o.displ_over(0.05)


def test_rotation_rotates_forces_and_virial_and_keeps_the_volume():
    tab = cu_tables()
    cfg = cu_config((8, 7, 9), T=600.0, jitter=0.1, seed=25)
    R = _rot([1.0, -2.0, 0.7], 0.43)
    a, b = _eam_sim(cfg, tab, R), _eam_sim(cfg, tab, None)
    _eam_forces(a, tab)
    _eam_forces(b, tab)
    assert frel(a.forces(), b.forces() @ R.T) <= 1e-11
    ta, tb = a.simulation_thermodynamic_state(), b.simulation_thermodynamic_state()
    assert ta.volume == pytest.approx(tb.volume, rel=1e-14)
    assert np.abs(ta.virial - R @ tb.virial @ R.T).max() <= 1e-10 * np.abs(tb.virial).max()
    assert ta.pressure_pa == pytest.approx(tb.pressure_pa, rel=1e-9)
    s = np.diag([1.03, 0.97, 1.01])
    a.set_xform(s)
    assert a.simulation_thermodynamic_state().volume == pytest.approx(tb.volume * np.linalg.det(s), rel=1e-14)


def test_rotation_leaves_the_adjoints_unchanged():
    """dL/dtheta of L = sum g_i . F_i + gE E is the same for a rotated domain with
    rotated g (the rotated system has the same energies and rotated forces), for the
    EAM tables and for LJ (epsilon, sigma)."""
    from exastamp_amd import inputs
    tab = cu_tables(500)
    cfg = cu_config((6, 6, 6), T=600.0, jitter=0.1, seed=26)
    R = _rot([0.2, 1.0, -0.4], 0.9)
    g = inputs.rng(3).normal(size=(cfg.n, 3))
    out = []
    for xf, gg in ((None, g), (R, g @ R.T)):
        sim = _eam_sim(cfg, tab, xf)
        _eam_forces(sim, tab)
        out.append(sim.eam_adjoint(gg, [0.7]))
    assert np.abs(out[1] - out[0]).max() <= 1e-10 * np.abs(out[0]).max()

    eps, sig, rc = lj_params()
    cfg = _jittered_ar(15, (6, 6, 6))
    g = inputs.rng(4).normal(size=(cfg.n, 3))
    out = []
    for xf, gg in ((None, g), (R, g @ R.T)):
        sim = make_sim(cfg, [AR_MASS])
        sim.set_xform(xf)
        sim.chunk_neighbors(rcut_max=rc, rcut_inc=0.3)
        sim.zero_force_energy()
        sim.lj_compute_force(parameters={"epsilon": "0.0104 eV", "sigma": "3.4 ang"}, rcut="8.5 ang")
        out.append(sim.lj_adjoint(gg, [0.3]))
    assert np.abs(out[1] - out[0]).max() <= 1e-10 * np.abs(out[0]).max()


def test_xform_errors_and_stale_list():
    from exastamp_amd import _lib
    tab = cu_tables(500)
    cfg = cu_config((4, 4, 4), T=0.0, jitter=0.02, seed=27)
    sim = _eam_sim(cfg, tab, SHEAR)
    with pytest.raises(ValueError):
        sim.set_xform(np.diag([1.0, -1.0, 1.0]))
    bad = np.ascontiguousarray(np.diag([1.0, 1.0, 0.0]).ravel())
    with pytest.raises(_lib.XsError) as e:
        sim.ctx.xs_domain_xform(_lib.dptr(bad))
    assert e.value.code == 1  # XS_E_ARG
    sim.set_xform(np.eye(3))  # a new xform leaves the list stale until the next build
    with pytest.raises(_lib.XsError):
        _eam_forces(sim, tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    _eam_forces(sim, tab)
    with pytest.raises(ValueError):
        sim.push_f_v_r(0.001, xform_mode="XFORM")


def test_rotated_replica_batch_rotates_every_replica():
    """The xform applies to every replica of a batch (strained boxes of the adjoint
    batches): each replica's forces rotate and its energy is unchanged."""
    from exastamp_amd import inputs
    tab = cu_tables(500)
    cfg = inputs.strained_batch(5, (4, 4, 4), 3.615, [CU_MASS], seed=13)
    R = _rot([0.5, 0.5, -1.0], 0.7)
    a, b = _eam_sim(cfg, tab, R), _eam_sim(cfg, tab, None)
    _eam_forces(a, tab)
    _eam_forces(b, tab)
    assert frel(a.forces(), b.forces() @ R.T) <= 1e-11
    ea, eb = a.replica_energies(), b.replica_energies()
    assert np.abs(ea - eb).max() <= 1e-12 * np.abs(eb).max()


def test_sheared_mixed_nve_conserves_energy_like_fp64():
    """The mixed MD loop under shear (block passes on staged physical positions, the
    drift in its own kernel): its energy drift over 200 steps is within 2x the FP64
    run's, as without an xform."""
    tab = cu_tables()
    cfg = cu_config((8, 8, 8), T=600.0, jitter=0.05, seed=28)
    drift = {}
    for mixed in (False, True):
        sim = make_sim(cfg, [CU_MASS])
        sim.set_xform(SHEAR)
        sim.eam_alloy_init(tab)
        st, th = sim.md_run("eam", 200, 0.001, 4.95, 0.3, mixed=mixed, thermo_every=10)
        assert st.rebuilds >= 3
        e = np.array([t.total_energy for t in th])
        drift[mixed] = abs(e[-1] - e[0])
        scale = abs(th[0].potential_energy)
    assert drift[True] <= 2.0 * drift[False] + 1e-9 * scale
