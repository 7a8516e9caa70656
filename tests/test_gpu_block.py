"""Cell-block (LDS-staged) EAM and LJ paths against the global-list path and the oracle.

The block list is the global list in another encoding (16-bit halo slots instead of
(index, image) words): converted back it must be identical entry by entry.  The block
kernels evaluate the same pairs with the same per-pair arithmetic as the global-list
kernels but sum each atom's list in interleaved parts (two or four lanes per atom),
so per-atom sums differ from the global path only by FP64 summation order.
The global path is selected with XSMD_LIST=global, read when a context is created.

Tolerances: neighbour lists bit-exact.  FP64 form: per-atom outputs within 1e-10 of
their rms (the sums of ~50 pair terms of up to ~10x the net force reorder at
~1e-12).  Mixed form: within 1e-6 of the rms, because a last-bit change of the FP64
density can flip the FP32 rounding of rho before the embedding lookup, which moves
F'(rho) by one FP32 ulp (6e-8 relative) and the forces with it.  Trajectories of 40
NVE steps: 1e-9 A (FP64), 1e-6 A (mixed).  Against the oracle as in test_gpu_parity
(FP64 1e-11, mixed 1e-5 of rms|F|).
"""
import numpy as np
import pytest

from conftest import cu_tables, cuni_tables
from test_gpu_parity import AR_MASS, CU_MASS, cu_config, frel, lj_c1_config, lj_params, make_oracle, make_sim

pytestmark = pytest.mark.gpu

CUNI = [63.546, 58.693]


def _case(name):
    from exastamp_amd import inputs
    if name == "cu_jitter":
        return cu_config((9, 8, 7), jitter=0.12, seed=5), [CU_MASS], None, cu_tables()
    if name == "cu_one_cell":        # 7.2 A box: one cell per axis, stencil sees 3 images
        return cu_config((2, 2, 2), jitter=0.05, seed=6), [CU_MASS], None, cu_tables()
    if name == "cu_two_cells":       # two cells per axis: -1 and +1 neighbours are one cell
        return cu_config((3, 4, 3), jitter=0.05, seed=7), [CU_MASS], None, cu_tables()
    if name == "cuni_slab":          # two species, free surface along z
        cfg = inputs.alloy_slab_config((8, 8, 6), 3.57, CUNI, vacancy_frac=0.01, vacuum=10.0, T=500.0, seed=9)
        return cfg, CUNI, ["Cu", "Ni"], cuni_tables()
    if name == "replicas":           # several boxes, different strains
        return inputs.strained_batch(5, (4, 4, 4), 3.615, [CU_MASS], seed=13), [CU_MASS], None, cu_tables(500)
    if name == "cu_open_y":          # y not periodic: atoms on both y faces and beyond them
        c = cu_config((6, 7, 5), jitter=0.1, seed=11)
        r = c.r.copy()
        by_y = np.argsort(r[:, 1])
        r[by_y[:3], 1] -= np.array([0.2, 0.45, 0.7])          # below lo (the bottom plane is at 0)
        r[by_y[-3:], 1] += 1.8 + np.array([0.15, 0.4, 0.65])  # above hi (the top plane is at hi - a/2)
        assert (r[:, 1] < c.lo[0, 1]).sum() >= 3 and (r[:, 1] >= c.hi[0, 1]).sum() == 3
        return inputs.Config(r, c.v, c.type, c.lo, c.hi, periodic=(1, 0, 1)), [CU_MASS], None, cu_tables()
    raise KeyError(name)


CASES = ["cu_jitter", "cu_one_cell", "cu_two_cells", "cuni_slab", "replicas", "cu_open_y"]


def _sim(monkeypatch, kind, cfg, masses, names, tab):
    if kind == "global":
        monkeypatch.setenv("XSMD_LIST", "global")
    else:
        monkeypatch.delenv("XSMD_LIST", raising=False)
    sim = make_sim(cfg, masses, names)
    sim.eam_alloy_init(tab)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    assert sim.list_info()["kind"] == kind
    return sim


def _outputs(sim, rc, mixed):
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=rc, trigger_thermo_state=True, compute_virial=True, mixed=mixed)
    return sim.accelerations(), sim.ep(), sim.virial(), sim.rho_dEmb()


@pytest.mark.parametrize("name", CASES)
def test_block_list_equals_global_list(monkeypatch, name):
    cfg, masses, names, tab = _case(name)
    b = _sim(monkeypatch, "block", cfg, masses, names, tab)
    g = _sim(monkeypatch, "global", cfg, masses, names, tab)
    info = b.list_info()
    assert info["blocks"] >= 1 and info["halo_capacity"] >= 1
    assert np.array_equal(b.order(), g.order())
    for x, y in zip(b.neighbors(), g.neighbors()):
        assert np.array_equal(x, y)


def _close(x, y, tol=1e-10):
    scale = np.sqrt(np.mean(y.astype(np.float64) ** 2))
    return np.abs(x - y).max() <= tol * max(scale, 1e-300)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("mixed", [False, True])
def test_block_forces_match_global(monkeypatch, name, mixed):
    cfg, masses, names, tab = _case(name)
    b = _sim(monkeypatch, "block", cfg, masses, names, tab)
    g = _sim(monkeypatch, "global", cfg, masses, names, tab)
    for x, y in zip(_outputs(b, tab.rc, mixed), _outputs(g, tab.rc, mixed)):
        assert _close(x, y, 1e-6 if mixed else 1e-10)


@pytest.mark.parametrize("name", ["cu_one_cell", "cu_two_cells", "cu_open_y"])
def test_block_small_boxes_vs_oracle(oracle_mod, monkeypatch, name):
    """Boxes of one and two cells per axis, and a box open along y with atoms beyond
    its faces (clamped into the edge cells, whose boxes are unbounded outward), against
    the oracle directly."""
    cfg, masses, names, tab = _case(name)
    sim = _sim(monkeypatch, "block", cfg, masses, names, tab)
    _outputs(sim, tab.rc, False)
    sim.force_to_accel()
    o = make_oracle(oracle_mod, cfg, masses)
    o.set_eam(tab)
    o.build_neighbors(tab.rc, 0.3)
    o.compute_forces("eam")
    assert np.array_equal(sim.order(), o.order())
    for x, y in zip(sim.neighbors(), o.neighbors()):
        assert np.array_equal(x, y)
    assert frel(sim.forces(), o.forces()) <= 1e-11
    ref = o.state()["ep"].sum()
    assert abs(sim.ep().sum() - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("mixed", [False, True])
def test_block_md_run_matches_global(monkeypatch, mixed):
    """40 fused NVE steps with rebuilds: the same trajectory on both list kinds."""
    tab = cu_tables()
    cfg = cu_config((6, 6, 6), T=1200.0, seed=4)
    out = []
    for kind in ("block", "global"):
        if kind == "global":
            monkeypatch.setenv("XSMD_LIST", "global")
        else:
            monkeypatch.delenv("XSMD_LIST", raising=False)
        sim = make_sim(cfg, [CU_MASS])
        sim.eam_alloy_init(tab)
        stats = sim.md_run("eam", 40, 0.002, tab.rc, 0.3, mixed=mixed)
        assert sim.list_info()["kind"] == kind
        out.append((sim.state(), stats))
    (ra, va), sa = out[0]
    (rb, vb), sb = out[1]
    assert np.abs(ra - rb).max() <= (1e-6 if mixed else 1e-9)
    assert _close(va, vb, 1e-6 if mixed else 1e-9)


def _lj_sim(monkeypatch, kind, cfg):
    if kind == "global":
        monkeypatch.setenv("XSMD_LIST", "global")
    else:
        monkeypatch.delenv("XSMD_LIST", raising=False)
    eps, sig, rc = lj_params()
    sim = make_sim(cfg, [AR_MASS])
    sim.lj_parameters(parameters={"epsilon": eps, "sigma": sig}, rcut=rc)
    return sim


def _lj_jitter_config():
    from exastamp_amd import inputs
    cfg = lj_c1_config(seed=21, cells=(9, 8, 7))
    return inputs.Config(cfg.r + inputs.rng(5).normal(scale=0.08, size=cfg.r.shape), cfg.v, cfg.type, cfg.lo,
                         cfg.hi)


@pytest.mark.parametrize("mixed", [False, True])
def test_block_lj_forces_match_global(monkeypatch, mixed):
    """LJ on the block list: forces, energies and virials as on the global list."""
    eps, sig, rc = lj_params()
    cfg = _lj_jitter_config()
    out = []
    for kind in ("block", "global"):
        sim = _lj_sim(monkeypatch, kind, cfg)
        sim.chunk_neighbors(rcut_max=rc, rcut_inc=0.3)
        assert sim.list_info()["kind"] == kind
        sim.zero_force_energy()
        sim.lj_compute_force(parameters={"epsilon": eps, "sigma": sig}, rcut=rc, trigger_thermo_state=True,
                             compute_virial=True, mixed=mixed)
        out.append((sim.accelerations(), sim.ep(), sim.virial()))
    for x, y in zip(*out):
        assert _close(x, y, 1e-6 if mixed else 1e-10)


def _lj_open_z_config():
    """The jittered argon crystal, open along z, with atoms moved beyond both z faces."""
    from exastamp_amd import inputs
    c = _lj_jitter_config()
    r = c.r.copy()
    by_z = np.argsort(r[:, 2])
    r[by_z[:2], 2] -= np.array([0.3, 0.9])
    r[by_z[-2:], 2] += 2.63 + np.array([0.2, 0.7])  # the top plane is at hi - a/2
    assert (r[:, 2] >= c.hi[0, 2]).sum() == 2
    return inputs.Config(r, c.v, c.type, c.lo, c.hi, periodic=(1, 1, 0))


@pytest.mark.parametrize("open_z", [False, True])
def test_block_lj_vs_oracle(monkeypatch, oracle_mod, open_z):
    """LJ on the block list against the oracle (FP64 1e-11 of rms|F|), periodic and
    open along z with atoms beyond the faces."""
    eps, sig, rc = lj_params()
    cfg = _lj_open_z_config() if open_z else _lj_jitter_config()
    sim = _lj_sim(monkeypatch, "block", cfg)
    sim.chunk_neighbors(rcut_max=rc, rcut_inc=0.3)
    assert sim.list_info()["kind"] == "block"
    sim.zero_force_energy()
    sim.lj_compute_force(parameters={"epsilon": eps, "sigma": sig}, rcut=rc, trigger_thermo_state=True,
                         compute_virial=True)
    sim.force_to_accel()
    o = make_oracle(oracle_mod, cfg, [AR_MASS])
    o.set_lj(eps, sig, rc)
    o.build_neighbors(rc, 0.3)
    o.compute_forces("lj")
    assert frel(sim.forces(), o.forces()) <= 1e-11
    ref = o.state()
    assert abs(sim.ep().sum() - ref["ep"].sum()) <= 1e-12 * abs(ref["ep"].sum())
    vir, vref = sim.virial().sum(0), ref["virial"].sum(0)
    assert np.abs(vir - vref).max() <= 1e-12 * np.abs(vref).max()


def test_block_lj_md_run_open_z_vs_oracle(monkeypatch, oracle_mod):
    """60 FP64 LJ NVE steps (dt 4 fs) in the argon crystal open along z with atoms
    beyond both faces: equal rebuild counts (eight), per-step potential energy within
    1e-9 and positions within 1e-8 A of the oracle's trajectory."""
    eps, sig, rc = lj_params()
    cfg = _lj_open_z_config()
    sim = _lj_sim(monkeypatch, "block", cfg)
    st, th = sim.md_run("lj", 60, 0.004, rc, 0.3, thermo_every=1)
    assert sim.list_info()["kind"] == "block"
    o = make_oracle(oracle_mod, cfg, [AR_MASS])
    o.set_lj(eps, sig, rc)
    nreb, th_ref = o.md_run("lj", 60, 0.004, rc, 0.3)
    assert st.rebuilds == nreb and nreb >= 1
    pe = np.array([t.potential_energy for t in th])
    assert np.abs(pe - th_ref[:, 24]).max() <= 1e-9 * np.abs(th_ref[:, 24]).max()
    r, _ = sim.state()
    assert np.abs(r - o.state()["r"]).max() < 1e-8


@pytest.mark.parametrize("mixed", [False, True])
def test_block_lj_md_run_matches_global(monkeypatch, mixed):
    """40 fused LJ NVE steps (drift epilogue on the block list) against the global list."""
    eps, sig, rc = lj_params()
    cfg = lj_c1_config(seed=23, cells=(8, 8, 8))
    out = []
    for kind in ("block", "global"):
        sim = _lj_sim(monkeypatch, kind, cfg)
        stats, _ = sim.md_run("lj", 40, 0.004, rc, 0.3, mixed=mixed)
        assert sim.list_info()["kind"] == kind
        out.append((sim.state(), stats))
    (ra, va), sa = out[0]
    (rb, vb), sb = out[1]
    assert sa.rebuilds == sb.rebuilds
    assert np.abs(ra - rb).max() <= (1e-6 if mixed else 1e-9)
    assert _close(va, vb, 1e-6 if mixed else 1e-9)


def test_block_adjoint_and_pair_count_use_converted_list(monkeypatch):
    """Operators that read the global list (pair count, adjoint) work after a block build
    and give the global-build results (pair counts exactly; the adjoint to 1e-12 of its
    largest entry, since its table gradients are accumulated with atomics)."""
    tab = cu_tables(500)
    cfg = cu_config((5, 5, 5), jitter=0.05, seed=12)
    res = []
    for kind in ("block", "global"):
        sim = _sim(monkeypatch, kind, cfg, [CU_MASS], None, tab)
        _outputs(sim, tab.rc, False)
        pc = sim.chunk_neighbors_stats(tab.rc)
        g = np.random.default_rng(1).normal(size=(cfg.n, 3))
        grad = sim.eam_adjoint(g, np.ones(1))
        res.append((pc, grad))
    assert res[0][0] == res[1][0] and res[0][0]["within_rcut"] > 0
    ga, gb = res[0][1], res[1][1]
    assert np.abs(ga - gb).max() <= 1e-12 * np.abs(gb).max()


def _env(monkeypatch, **kv):
    for k in ("XSMD_LIST", "XSMD_LIST_SORT", "XSMD_BLOCK_ATOMS", "XSMD_SORT_BITS"):
        monkeypatch.delenv(k, raising=False)
    for k, v in kv.items():
        monkeypatch.setenv(k, v)


def test_block_md_run_vs_oracle_energy(monkeypatch, oracle_mod):
    """100 mixed NVE steps on the default path (block passes, sorted lists, fused
    drift; Cu 600 K, the bench's density): the potential energy of every step within
    1e-6 of the FP64 oracle trajectory's."""
    _env(monkeypatch)
    tab = cu_tables()
    cfg = cu_config((8, 8, 8), T=600.0, jitter=0.0, seed=31)
    sim = make_sim(cfg, [CU_MASS])
    sim.eam_alloy_init(tab)
    _, th = sim.md_run("eam", 100, 0.001, tab.rc, 0.3, mixed=True, thermo_every=1)
    assert sim.list_info()["kind"] == "block"
    o = make_oracle(oracle_mod, cfg, [CU_MASS])
    o.set_eam(tab)
    _, th_ref = o.md_run("eam", 100, 0.001, tab.rc, 0.3)
    pe = np.array([t.potential_energy for t in th])
    assert np.abs(pe - th_ref[:, 24]).max() <= 1e-6 * np.abs(th_ref[:, 24]).max()


def test_block_md_run_open_box_vs_oracle(monkeypatch, oracle_mod):
    """60 FP64 NVE steps at 600 K in the box open along y, some of whose face atoms start
    beyond the faces (clamped into the edge cells at every rebuild; seven rebuilds):
    equal rebuild counts, per-step potential energy within 1e-9 and positions within
    1e-8 A of the oracle's trajectory."""
    _env(monkeypatch)
    cfg, masses, names, tab = _case("cu_open_y")
    sim = make_sim(cfg, masses)
    sim.eam_alloy_init(tab)
    st, th = sim.md_run("eam", 60, 0.001, tab.rc, 0.3, thermo_every=1)
    assert sim.list_info()["kind"] == "block"
    o = make_oracle(oracle_mod, cfg, masses)
    o.set_eam(tab)
    nreb, th_ref = o.md_run("eam", 60, 0.001, tab.rc, 0.3)
    assert st.rebuilds == nreb and nreb >= 1
    pe = np.array([t.potential_energy for t in th])
    assert np.abs(pe - th_ref[:, 24]).max() <= 1e-9 * np.abs(th_ref[:, 24]).max()
    r, _ = sim.state()
    ro = o.state()["r"]
    assert np.abs(r - ro).max() < 1e-8
    assert (ro[:, 1] < cfg.lo[0, 1]).sum() >= 3 and (ro[:, 1] >= cfg.hi[0, 1]).sum() >= 1


@pytest.mark.parametrize("name", ["cu_jitter", "cuni_slab", "replicas", "cu_open_y"])
def test_distance_sorted_lists(monkeypatch, name):
    """Block lists are reordered by distance after the build (XSMD_LIST_SORT=0 keeps
    the stencil order): the converted global list is the same entry by entry, and
    the passes agree to summation order (FP64 1e-10, mixed 1e-6 of the rms)."""
    cfg, masses, names, tab = _case(name)
    sims = {}
    for label, env in (("sorted", {}), ("stencil", {"XSMD_LIST_SORT": "0"})):
        _env(monkeypatch, **env)
        sim = make_sim(cfg, masses, names)
        sim.eam_alloy_init(tab)
        sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
        assert sim.list_info()["kind"] == "block"
        sims[label] = sim
    for x, y in zip(sims["sorted"].neighbors(), sims["stencil"].neighbors()):
        assert np.array_equal(x, y)
    for mixed in (False, True):
        for x, y in zip(_outputs(sims["sorted"], tab.rc, mixed), _outputs(sims["stencil"], tab.rc, mixed)):
            assert _close(x, y, 1e-6 if mixed else 1e-10)


def test_bricks_beyond_one_round(monkeypatch):
    """Bricks of more central atoms than the 256 atom slots of a pass workgroup walk
    their atoms in two rounds: FCC Cu 16^3 unit cells on an 11^3 cell grid (12.3 atoms
    per cell) with XSMD_BLOCK_ATOMS=340 gives 3x3x3-cell bricks of ~330 atoms.  Lists
    equal to the global build, forces as in the block-vs-global tests."""
    cfg, masses, names, tab = cu_config((16, 16, 16), jitter=0.08, seed=21), [CU_MASS], None, cu_tables()
    _env(monkeypatch, XSMD_BLOCK_ATOMS="340")
    b = _sim(monkeypatch, "block", cfg, masses, names, tab)
    _env(monkeypatch)
    g = _sim(monkeypatch, "global", cfg, masses, names, tab)
    info = b.list_info()
    assert info["central_max"] > 256, info
    for x, y in zip(b.neighbors(), g.neighbors()):
        assert np.array_equal(x, y)
    for mixed, tol in ((False, 1e-10), (True, 1e-6)):
        for x, y in zip(_outputs(b, tab.rc, mixed), _outputs(g, tab.rc, mixed)):
            assert _close(x, y, tol)
