"""Host-side logic: units, seeded inputs, table generation (no GPU)."""
import numpy as np
import pytest


def test_units_parse_deck_quantities():
    from exastamp_amd import units
    # 1 eV = e / (amu * 1e4) Da.A^2/ps^2 (CODATA-2022 amu) ~ 9648.53
    assert units.parse_quantity("0.0104 eV") == pytest.approx(0.0104 * 1.602176634e-19 / 1.66053906892e-23,
                                                             rel=1e-14)
    assert 9648.5 < units.EV < 9648.6
    assert units.parse_quantity("3.4 ang") == 3.4
    assert units.parse_quantity("39.948 Da") == 39.948
    assert units.parse_quantity("1.0e-3 ps") == 1e-3
    assert units.parse_quantity("2 fs") == pytest.approx(0.002)
    assert units.parse_quantity("50. m/s") == pytest.approx(0.5)  # A/ps
    assert units.parse_quantity(4.95) == 4.95
    with pytest.raises(ValueError):
        units.parse_quantity("3 furlongs")
    # kB T at 300 K in eV
    assert units.KB * 300 / units.EV == pytest.approx(0.025852, rel=1e-4)


def test_maxwell_boltzmann_exact_temperature_zero_momentum():
    from exastamp_amd import inputs, units
    cfg = inputs.lattice_config((6, 6, 6), 3.615, [63.546], T=600.0, seed=3)
    m = 63.546
    p = (m * cfg.v).sum(0)
    assert np.abs(p).max() < 1e-9 * m * np.abs(cfg.v).max() * cfg.n
    ke = 0.5 * m * (cfg.v ** 2).sum()
    assert units.temperature_from_ke(ke, cfg.n) == pytest.approx(600.0, rel=1e-12)


def test_inputs_are_seeded_and_deterministic():
    from exastamp_amd import inputs
    a = inputs.lattice_config((4, 4, 4), 3.615, [63.546], T=300.0, jitter=0.05, seed=11)
    b = inputs.lattice_config((4, 4, 4), 3.615, [63.546], T=300.0, jitter=0.05, seed=11)
    c = inputs.lattice_config((4, 4, 4), 3.615, [63.546], T=300.0, jitter=0.05, seed=12)
    assert np.array_equal(a.r, b.r) and np.array_equal(a.v, b.v)
    assert not np.array_equal(a.r, c.r)
    assert a.n == 256


def test_fcc_neighbour_shell_counts():
    """12, 6, 24, 12, 24 neighbours in the first five FCC shells (Cu: 42 within 4.95 A,
    54 within 5.25 A)."""
    from exastamp_amd.potentials import _fcc_shells
    d, c = _fcc_shells(3.615, 5.8)
    assert list(c[:5]) == [12, 6, 24, 12, 24]
    assert c[d < 4.95].sum() == 42 and c[d < 5.25].sum() == 54


def test_alloy_slab_and_batch_layouts():
    from exastamp_amd import inputs
    s = inputs.alloy_slab_config((8, 8, 8), 3.57, [63.546, 58.693], vacancy_frac=0.01, vacuum=10.0, seed=1)
    assert s.periodic == (1, 1, 0)
    assert 0.4 < s.type.mean() < 0.6
    assert 0.97 * 2048 < s.n < 2048
    assert s.r[:, 2].min() >= 10.0 and s.r[:, 2].max() <= s.hi[0, 2] - 10.0 + 1e-9
    b = inputs.strained_batch(6, (3, 3, 3), 3.615, [63.546], seed=2)
    assert b.nrep == 6 and b.n == 6 * 108
    L = b.hi[:, 0]
    assert np.all(np.abs(L / (3 * 3.615) - 1) <= 0.03 + 1e-12)
    parts = inputs.split_replicas(b)
    assert sum(p.n for p in parts) == b.n


def test_synthetic_tables_zero_pressure_lattice():
    """r0 of the Morse term is solved so that dE/da = 0 at the lattice constant."""
    from exastamp_amd import potentials
    t = potentials.johnson_tables(nr=2000, nrho=2000)
    assert t.nel == 1 and t.frho.shape == (1, 2000) and t.z2r.shape == (1, 2000)
    assert t.rhor[0, -1] == pytest.approx(0.0, abs=1e-12) and t.z2r[0, -1] == pytest.approx(0.0, abs=1e-12)
    theta = t.theta()
    assert theta.size == 6000
    t2 = t.with_theta(theta * 1.01)
    assert np.allclose(t2.frho, 1.01 * t.frho)


def test_atomic_mass_constant_from_recorded_density():
    """The reference's recorded AlCu B2 run (fixture) prints the density of 8000 Al +
    8000 Cu atoms in a 72 A cube as 3.221992473234 g/cm^3; that pins the atomic mass
    constant its unit conversions use (CODATA-2022) to the 12 digits printed."""
    from conftest import GOLDEN
    from exastamp_amd import units
    row = (GOLDEN / "alcu_b2_thermodynamic_state.csv").read_text().splitlines()[1].split()
    volume, rho = float(row[19]), float(row[20])
    assert volume == 72.0 ** 3
    mass_da = 8000 * 26.982 + 8000 * 63.546
    assert abs(mass_da * units.AMU_KG * 1e3 / (volume * 1e-24) / rho - 1.0) < 2e-12


def test_fit_loss_is_the_elementwise_loss():
    """fit_loss (one autograd node with its analytic backward) gives the value and the
    gradients of the elementwise form: force MSE + 0.1 per-atom energy MSE."""
    import torch
    from exastamp_amd.fitting import fit_loss
    g = torch.Generator().manual_seed(3)
    F = torch.randn(120, 3, dtype=torch.float64, generator=g, requires_grad=True)
    E = torch.randn(5, dtype=torch.float64, generator=g, requires_grad=True)
    Ft = torch.randn(120, 3, dtype=torch.float64, generator=g)
    Et = torch.randn(5, dtype=torch.float64, generator=g)
    n = [20, 30, 25, 25, 20]
    L = fit_loss(F, E, Ft, Et, n)
    (3.0 * L).backward()
    gF, gE = F.grad.clone(), E.grad.clone()
    F.grad = E.grad = None
    nt = torch.tensor(n, dtype=torch.float64)
    L2 = ((F - Ft) ** 2).mean() + 0.1 * (((E - Et) / nt) ** 2).mean()
    (3.0 * L2).backward()
    assert abs(float(L.detach()) - float(L2.detach())) <= 1e-15 * float(L2.detach())
    assert torch.allclose(gF, F.grad, rtol=1e-14, atol=0) and torch.allclose(gE, E.grad, rtol=1e-14, atol=1e-30)


def test_thermo_cell_geometry():
    """Box edges and angles of the thermo outputs: xform times the grid box edges
    (print_thermodynamic_state.cpp:252-262); a shear in the b-c plane tilts alpha."""
    import math
    from exastamp_amd.deck import cell_geometry
    lengths, angles = cell_geometry(np.eye(3), [3.0, 4.0, 5.0])
    assert lengths == (3.0, 4.0, 5.0) and angles == (90.0, 90.0, 90.0)
    t = 0.3
    lengths, angles = cell_geometry([[1, 0, 0], [0, 1, t], [0, 0, 1]], [3.0, 4.0, 5.0])
    assert lengths[2] == pytest.approx(5.0 * math.hypot(1.0, t), rel=1e-15)
    assert angles[0] == pytest.approx(math.degrees(math.atan2(1.0, t)), rel=1e-12)
    assert angles[1] == pytest.approx(90.0) and angles[2] == pytest.approx(90.0)


def test_deck_domain_accepts_a_deforming_xform():
    from exastamp_amd.deck import DeckError, _op_domain

    class R:
        domain = {}

    r = R()
    _op_domain(r, {"xform": [[1.0, 0.1, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]})
    assert np.allclose(r.domain["xform"][0], [1.0, 0.1, 0.0])
    with pytest.raises(DeckError):
        _op_domain(r, {"xform": [[1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [0.0, 0.0, 1.0]]})
