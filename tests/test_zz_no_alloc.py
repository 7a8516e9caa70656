"""xs_thermo_noalloc (the allocation-free form of xs_thermo) against xs_thermo,
through the C ABI: same 27 values, bit for bit, on two calls in a row whose second
system is smaller than the first (the fixed partial-sum array keeps the first
call's rows past the second call's blocks; they must not be read)."""
import ctypes as C

import numpy as np
import pytest

from conftest import cu_tables

pytestmark = pytest.mark.gpu
CU = 63.546


def _state(sim, tab, cfg):
    sim.set_config(cfg)
    sim.chunk_neighbors(rcut_max=tab.rc, rcut_inc=0.3)
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=tab.rc, trigger_thermo_state=True, compute_virial=True)
    sim.force_to_accel()


def _both(sim):
    import torch
    from exastamp_amd import _lib
    host = np.zeros(27)
    sim.ctx.xs_thermo(_lib.dptr(host))
    dev = torch.full((27,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    sim.ctx.xs_thermo_noalloc(C.c_void_p(dev.data_ptr()))
    sim.ctx.xs_synchronize()
    return host, dev.cpu().numpy()


def test_thermo_noalloc_matches_thermo():
    from exastamp_amd import inputs
    from exastamp_amd.operators import Simulation
    tab = cu_tables(500)
    sim = Simulation(0)
    sim.species([("Cu", CU)])
    sim.eam_alloy_init(tab)
    # 40^3 FCC cells = 256 000 atoms: 1000 partial blocks; then 5^3 cells = 500 atoms: 2 blocks
    for cells, seed in (((40, 40, 40), 3), ((5, 5, 5), 4)):
        cfg = inputs.lattice_config(cells, 3.615, [CU], T=600.0, jitter=0.05, seed=seed)
        _state(sim, tab, cfg)
        host, dev = _both(sim)
        print(f"{cfg.n} atoms: PE {host[24]:.6e}, KE {host[21:24].sum():.6e}, n {dev[26]:.0f}")
        assert dev[26] == cfg.n
        np.testing.assert_array_equal(dev, host)
    sim.close()


def test_thermo_noalloc_rejects_null_output():
    from exastamp_amd.operators import Simulation, XsError
    sim = Simulation(0)
    with pytest.raises(XsError, match="null output"):
        sim.ctx.xs_thermo_noalloc(None)
    sim.close()
