"""Pins for the oracle's deformed-domain path (exaNBody's Domain::xform through
LinearXForm, pair_potential_impl.hxx:388-413; the push's inv_xform mapping,
push_vec3_2nd_order_xform_minimization.cpp:34).  Positions are grid coordinates,
displacements, forces and velocities physical.

Each check fixes the transform by physics, not by restating it:
- a diagonal xform is the same system as an identity box stretched by it;
- a rotation leaves energies unchanged and rotates forces and the virial (an xform
  applied transposed would rotate them the other way);
- under shear, the force is minus the energy gradient in physical space
  (dE/ds = -M^T F), across the sheared periodic boundaries;
- the virial is minus the derivative of the energy under M -> (I + eps) M;
- the cell list equals brute force over the sheared images;
- an NVE run under a rotation follows the identity run in grid space (the push
  maps the physical step back by M^-1).
"""

import numpy as np
import pytest

from conftest import cu_tables

CU = 63.546
SHEAR = np.array([[1.02, 0.17, -0.06], [0.0, 0.97, 0.11], [0.0, 0.0, 1.01]])


def _rot(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * k @ k


def _cu(cells=(4, 4, 5), jitter=0.08, seed=5, T=0.0):
    from exastamp_amd import inputs
    return inputs.lattice_config(cells, 3.615, [CU], T=T, jitter=jitter, seed=seed)


def _run(cfg, xf=None, tab=None, skin=0.3, forces=True):
    import pyoracle
    tab = tab or cu_tables()
    o = pyoracle.Oracle(threads=4)
    o.set_config(cfg, [CU])
    o.set_xform(xf)
    o.set_eam(tab)
    o.build_neighbors(tab.rc, skin)
    if forces:
        o.compute_forces("eam")
    return o


def _energy(cfg, xf):
    return _run(cfg, xf).state()["ep"].sum()


def _list_set(o):
    c, ids, im = o.neighbors()
    off = np.concatenate([[0], np.cumsum(c)])
    return [sorted(zip(ids[off[a]:off[a + 1]].tolist(), map(tuple, im[off[a]:off[a + 1]].tolist())))
            for a in range(len(c))]


def test_set_xform_rejects_singular_and_reflecting_maps():
    import pyoracle
    o = pyoracle.Oracle()
    with pytest.raises(RuntimeError):
        o.set_xform(np.diag([1.0, 1.0, 0.0]))
    with pytest.raises(RuntimeError):
        o.set_xform(np.diag([1.0, -1.0, 1.0]))
    o.set_xform(None)


def test_diagonal_xform_is_a_stretched_box():
    from exastamp_amd import inputs
    cfg = _cu()
    d = np.array([1.031, 0.982, 1.013])
    a = _run(cfg, np.diag(d))
    b = _run(inputs.Config(cfg.r * d, cfg.v, cfg.type, cfg.lo * d, cfg.hi * d))
    sa, sb = a.state(), b.state()
    rms = np.sqrt((b.forces() ** 2).sum(1).mean())
    assert np.abs(a.forces() - b.forces()).max() <= 1e-11 * rms
    assert np.abs(sa["ep"] - sb["ep"]).max() <= 1e-12 * np.abs(sb["ep"]).max()
    assert np.abs(sa["virial"] - sb["virial"]).max() <= 1e-11 * np.abs(sb["virial"]).max()
    assert _list_set(a) == _list_set(b)


def test_rotation_xform_rotates_forces_and_virial():
    cfg = _cu()
    R = _rot([1.0, 2.0, 0.5], 0.37)
    a, b = _run(cfg, R), _run(cfg)
    sa, sb = a.state(), b.state()
    rms = np.sqrt((b.forces() ** 2).sum(1).mean())
    assert np.abs(a.forces() - b.forces() @ R.T).max() <= 1e-11 * rms
    assert np.abs(a.forces() - b.forces() @ R).max() > 1e-3 * rms  # the transposed map is wrong
    assert np.abs(sa["ep"] - sb["ep"]).max() <= 1e-12 * np.abs(sb["ep"]).max()
    W = np.einsum("ab,nbc,dc->nad", R, sb["virial"], R)
    assert np.abs(sa["virial"] - W).max() <= 1e-11 * np.abs(sb["virial"]).max()
    assert _list_set(a) == _list_set(b)


def test_sheared_force_is_minus_physical_energy_gradient():
    from exastamp_amd import inputs
    cfg = _cu()
    o = _run(cfg, SHEAR)
    F = o.forces()
    g = F @ SHEAR  # (M^T F)_a = sum_b M_ba F_b = -dE/ds_a
    L = cfg.hi[0] - cfg.lo[0]
    edge = np.argmax(np.abs(cfg.r[:, 0] - cfg.lo[0, 0] - 0.5 * L[0]))  # an atom at the x boundary
    h = 1e-5
    for i in (0, 57, int(edge)):
        for a in range(3):
            rp, rm = cfg.r.copy(), cfg.r.copy()
            rp[i, a] += h
            rm[i, a] -= h
            ep = _energy(inputs.Config(rp, cfg.v, cfg.type, cfg.lo, cfg.hi), SHEAR)
            em = _energy(inputs.Config(rm, cfg.v, cfg.type, cfg.lo, cfg.hi), SHEAR)
            assert g[i, a] == pytest.approx(-(ep - em) / (2 * h), rel=1e-6, abs=1e-6 * np.abs(g).max())


def test_sheared_virial_is_minus_deformation_derivative():
    cfg = _cu()
    W = _run(cfg, SHEAR).state()["virial"].sum(0)
    h = 1e-6
    for a in range(3):
        for b in range(3):
            e = np.zeros((3, 3))
            e[a, b] = h
            ep = _energy(cfg, (np.eye(3) + e) @ SHEAR)
            em = _energy(cfg, (np.eye(3) - e) @ SHEAR)
            assert W[a, b] == pytest.approx(-(ep - em) / (2 * h), rel=1e-6, abs=1e-7 * np.abs(W).max())


def test_sheared_cell_list_equals_brute_force():
    import pyoracle
    cfg = _cu((4, 5, 5), 0.15, 2)
    o = pyoracle.Oracle()
    o.set_config(cfg, [CU])
    o.set_xform(SHEAR)
    o.build_neighbors(4.95, 0.3)
    c1, i1, m1 = o.neighbors()
    c2, i2, m2 = o.brute_neighbors(5.25)
    assert np.array_equal(c1, c2) and c1.sum() > 40 * len(c1)
    off = np.concatenate([[0], np.cumsum(c1)])
    for a in range(len(c1)):
        s = slice(off[a], off[a + 1])
        k1 = np.lexsort((m1[s, 0], m1[s, 1], m1[s, 2], i1[s]))
        assert np.array_equal(i1[s][k1], i2[s]) and np.array_equal(m1[s][k1], m2[s])


def test_rotated_nve_run_follows_the_identity_run_in_grid_space():
    from exastamp_amd import inputs
    cfg = _cu(T=600.0)
    R = _rot([0.3, -1.0, 2.0], 0.61)
    rot = inputs.Config(cfg.r, cfg.v @ R.T, cfg.type, cfg.lo, cfg.hi)
    runs = []
    for c, xf in ((cfg, None), (rot, R)):
        import pyoracle
        o = pyoracle.Oracle(threads=4)
        o.set_config(c, [CU])
        o.set_xform(xf)
        o.set_eam(cu_tables())
        nreb, th = o.md_run("eam", 60, 0.001, 4.95, 0.3)
        runs.append((nreb, th, o.state()))
    (n0, th0, s0), (n1, th1, s1) = runs
    assert n0 == n1 and n0 >= 2
    assert np.abs(s1["r"] - s0["r"]).max() <= 1e-9
    assert np.abs(s1["v"] - s0["v"] @ R.T).max() <= 1e-9 * np.abs(s0["v"]).max()
    assert np.abs(th1[:, 24] - th0[:, 24]).max() <= 1e-9 * np.abs(th0[:, 24]).max()
