"""exaStamp operators of the short-range force path, over libxsmd's C ABI.

Each method of ``Simulation`` is named after the exaStamp/exaNBody operator it
stands in for and takes that operator's slots as keyword arguments (units as
deck quantities such as ``"0.0104 eV"`` are accepted wherever the reference
takes a Quantity).  Where the reference aborts through fatal_error() this
raises ``XsError`` (from the C ABI) or ``ValueError`` (slot validation).

    sim = Simulation()
    sim.species([("Cu", 63.546)])
    sim.set_config(cfg)
    sim.eam_alloy_init(parameters=tables)            # or {"file": "Cu.eam.alloy"}
    sim.chunk_neighbors(rcut_max=4.95, rcut_inc=0.3)
    sim.zero_force_energy()
    sim.eam_alloy_force(rcut=4.95, trigger_thermo_state=True, compute_virial=True)
    sim.force_to_accel()
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import XS_ENERGY, XS_MIXED, XS_SYMMETRIC, XS_VIRIAL, XsError
from .inputs import Config
from .potentials import SetflTables
from .units import EV, KB, PRESSURE_UNIT_PA, parse_quantity

# EamJohnsonParameters (johnson.h:30-51): the order of xs_johnson_force's parameter array
JOHNSON_PARAMETERS = ("re", "fe", "rhoe", "alpha", "beta", "A", "B", "kappa", "lambda", "Fn0", "Fn1", "Fn2", "Fn3",
                      "F0", "F1", "F2", "F3", "Fo", "eta")


@dataclasses.dataclass
class ThermodynamicState:
    """simulation_thermodynamic_state outputs (simulation_thermodynamic_state.cpp:165-225)."""
    raw: np.ndarray                 # the 27 reduced doubles
    volume: float

    @property
    def virial(self) -> np.ndarray:
        return self.raw[0:9].reshape(3, 3)

    @property
    def kinetic_tensor(self) -> np.ndarray:
        return self.raw[9:18].reshape(3, 3)

    @property
    def momentum(self) -> np.ndarray:
        return self.raw[18:21]

    @property
    def kinetic_energy(self) -> float:
        return float(self.raw[21:24].sum())

    @property
    def potential_energy(self) -> float:
        return float(self.raw[24])

    @property
    def mass(self) -> float:
        return float(self.raw[25])

    @property
    def particle_count(self) -> int:
        return int(self.raw[26])

    @property
    def total_energy(self) -> float:
        return self.kinetic_energy + self.potential_energy

    @property
    def temperature(self) -> float:
        """Kelvin: 2 (KE - p^2/2M) / (3 N kB)."""
        ke = self.raw[21:24] - 0.5 * self.momentum ** 2 / self.mass
        return float(2.0 * ke.sum() / (3.0 * self.particle_count * KB))

    @property
    def pressure_pa(self) -> float:
        """Hydrostatic pressure (kinetic tensor + virial) / V, in Pa."""
        kt = 2.0 * (self.kinetic_tensor - 0.5 * np.outer(self.momentum, self.momentum) / self.mass)
        p = (np.trace(kt) + np.trace(self.virial)) / (3.0 * self.volume)
        return float(p * PRESSURE_UNIT_PA)

    def per_atom_ev(self, e: float) -> float:
        return e / EV / self.particle_count


class Simulation:
    def __init__(self, device: int = 0):
        self.ctx = _lib.Context(device)
        self.species_names: list[str] = []
        self.masses: list[float] = []
        self.n = 0
        self.cfg: Config | None = None
        self.rcut_max = 0.0
        self.lj_params: tuple[float, float] | None = None
        self.lj_rcut = 0.0
        self.eam: SetflTables | None = None
        self.eam_knots = 0          # knots of the loaded tables (the length of theta)
        self._types_set = None      # the species mask eam_alloy_force last sent (not re-sent unchanged)
        self.xform = np.eye(3)      # domain xform (grid -> physical)

    # --- set-up -------------------------------------------------------------
    def species(self, spec):
        """particle_type_add_properties / ParticleSpecies: [(name, mass), ...]."""
        self.species_names = [s[0] for s in spec]
        self.masses = [parse_quantity(s[1]) for s in spec]
        m = np.asarray(self.masses, dtype=np.float64)
        self.ctx.xs_species_set(len(self.masses), _lib.dptr(m))

    def set_config(self, cfg: Config):
        """Load atoms (what setup_system + init_particles produce)."""
        r = np.ascontiguousarray(cfg.r, dtype=np.float64)
        v = np.ascontiguousarray(cfg.v, dtype=np.float64)
        cols = [np.ascontiguousarray(r[:, a]) for a in range(3)] + [np.ascontiguousarray(v[:, a]) for a in range(3)]
        t = np.ascontiguousarray(cfg.type, dtype=np.uint8)
        rep = None if cfg.rep is None else np.ascontiguousarray(cfg.rep, dtype=np.uint16)
        lo = np.ascontiguousarray(cfg.lo, dtype=np.float64).ravel()
        hi = np.ascontiguousarray(cfg.hi, dtype=np.float64).ravel()
        per = (C.c_int * 3)(*[int(p) for p in cfg.periodic])
        self.ctx.xs_system_set(cfg.n, *[_lib.dptr(c) for c in cols], _lib.ptr(t, C.c_uint8),
                               _lib.ptr(rep, C.c_uint16), cfg.nrep, _lib.dptr(lo), _lib.dptr(hi), per)
        self.n = cfg.n
        self.cfg = cfg
        self._keep = cols  # keep host buffers alive until the copy completed

    def set_xform(self, m=None):
        """Domain xform (exaNBody Domain::set_xform): the 3x3 map from grid coordinates
        (the positions of set_config and state) to physical ones; None = identity.
        Pair displacements, forces, velocities and the virial are physical; the
        thermodynamic volume is det(xform) times the grid box volume.  The neighbour
        list must be rebuilt (chunk_neighbors) before the next force operator."""
        a = np.eye(3) if m is None else np.asarray(m, dtype=np.float64).reshape(3, 3)
        if not np.all(np.isfinite(a)) or not np.linalg.det(a) > 0.0:
            raise ValueError("domain xform: det must be positive")
        flat = np.ascontiguousarray(a.ravel())
        self.ctx.xs_domain_xform(_lib.dptr(flat))
        self.xform = a.copy()

    # --- neighbour list -------------------------------------------------------
    def chunk_neighbors(self, rcut_max=None, rcut_inc=0.0, config=None):
        """chunk_neighbors (config_move_particles.msp:54-68); rcut_inc is the skin."""
        config = config or {}
        chunk = int(config.get("chunk_size", 1))
        rmax = parse_quantity(rcut_max) if rcut_max is not None else self.rcut_max
        if rmax <= 0:
            raise ValueError("chunk_neighbors needs rcut_max > 0 (run a force operator first or pass it)")
        self.ctx.xs_chunk_neighbors(rmax, parse_quantity(rcut_inc), chunk)

    def particle_displ_over(self, threshold) -> bool:
        out = C.c_int(0)
        self.ctx.xs_particle_displ_over(parse_quantity(threshold), C.byref(out))
        return bool(out.value)

    # --- operators ------------------------------------------------------------
    def zero_force_energy(self):
        self.ctx.xs_zero_force_energy()

    def force_to_accel(self):
        self.ctx.xs_force_to_accel()

    @staticmethod
    def _flags(trigger_thermo_state, compute_virial, mixed):
        f = 0
        if trigger_thermo_state:
            f |= XS_ENERGY
            if compute_virial:
                f |= XS_VIRIAL
        if mixed:
            f |= XS_MIXED
        return f

    def lj_compute_force(self, parameters, rcut, ghost=False, compute_virial=True,
                         trigger_thermo_state=True, mixed=False, symmetric=False, type=None, **_slots):
        """lj_compute_force{parameters: {epsilon, sigma}, rcut} (pair_potential_impl.hxx:108-121).
        The reference always fills ep (and the virial when the grid has one); here
        they are filled when trigger_thermo_state (and compute_virial) are set.  As
        there (pair_potential_impl.hxx:143-158), a single-material operator on a system
        of several species needs its `type` slot; every pair then takes the parameters."""
        if len(self.species_names) > 1 and type is None:
            raise ValueError(f"lj_compute_force: exactly 1 atom type expected (single material) but found "
                             f"{len(self.species_names)}; give the type slot")
        if type is not None and type not in self.species_names:
            raise ValueError(f"lj_compute_force: unknown type {type!r}")
        eps = parse_quantity(parameters["epsilon"])
        sig = parse_quantity(parameters["sigma"])
        rc = parse_quantity(rcut)
        self.rcut_max = max(self.rcut_max, rc)
        self.lj_params = (eps, sig)
        self.lj_rcut = rc
        flags = self._flags(trigger_thermo_state, compute_virial, mixed) | (XS_SYMMETRIC if symmetric else 0)
        self.ctx.xs_lj_compute_force(eps, sig, rc, flags)

    def lj_adjoint(self, dL_dF, dL_dE, parameters=None, rcut=None) -> np.ndarray:
        """(dL/depsilon, dL/dsigma) in internal units given dL/dF (n, 3; forces in internal
        units) and dL/dE per replica, for lj_compute_force with these parameters and rcut
        (default: those of the last lj_compute_force / lj_parameters call) at the current
        positions."""
        if parameters is None:
            if self.lj_params is None:
                raise ValueError("lj_adjoint needs LJ parameters (run lj_compute_force first or pass them)")
            eps, sig = self.lj_params
        else:
            eps, sig = parse_quantity(parameters["epsilon"]), parse_quantity(parameters["sigma"])
        rc = parse_quantity(rcut) if rcut is not None else self.lj_rcut
        g = np.ascontiguousarray(np.asarray(dL_dF, dtype=np.float64).reshape(self.n, 3))
        ge = np.ascontiguousarray(np.asarray(dL_dE, dtype=np.float64).reshape(-1))
        if ge.size != self.cfg.nrep:
            raise ValueError("dL_dE needs one value per replica")
        out = np.zeros(2)
        self.ctx.xs_lj_adjoint(eps, sig, rc, _lib.dptr(g), _lib.dptr(ge), _lib.dptr(out), 0)
        return out

    def lj_compute_force_symetric(self, parameters, rcut, **kw):
        """lj_compute_force_symetric (pair_potential_singlemat_symetric.cpp:101-332):
        the same forces, each pair evaluated once with Newton's third law."""
        self.lj_compute_force(parameters, rcut, symmetric=True, **kw)

    def lj_multi_force(self, rcut, parameters=(), common_parameters=None, compute_virial=True,
                       trigger_thermo_state=True, mixed=False, **_slots):
        """lj_multi_force{rcut, common_parameters: {epsilon, sigma}, parameters: [{type_a,
        type_b, rcut, parameters: {epsilon, sigma}}, ...]} (pair_potential_impl.hxx:282-320,
        pair_potential_force_op_multiparam.h:229-283): every unordered pair of species takes
        its entry (in either order) or common_parameters with the operator's rcut."""
        rc = parse_quantity(rcut)
        common = common_parameters or {"epsilon": 0.0, "sigma": 0.0}
        names = list(self.species_names)
        nt = len(names)
        npair = nt * (nt + 1) // 2
        eps = np.full(npair, parse_quantity(common["epsilon"]), dtype=np.float64)
        sig = np.full(npair, parse_quantity(common["sigma"]), dtype=np.float64)
        rcs = np.full(npair, rc, dtype=np.float64)
        for e in parameters or ():
            try:
                a, b = names.index(e["type_a"]), names.index(e["type_b"])
            except ValueError:
                raise ValueError(f"lj_multi_force: unknown species in {e}") from None
            lo, hi = min(a, b), max(a, b)
            pid = hi * (hi + 1) // 2 + lo
            eps[pid] = parse_quantity(e["parameters"]["epsilon"])
            sig[pid] = parse_quantity(e["parameters"]["sigma"])
            rcs[pid] = parse_quantity(e["rcut"])
        self.rcut_max = max(self.rcut_max, rc, float(rcs.max()))
        self.ctx.xs_lj_multi_force(nt, _lib.dptr(eps), _lib.dptr(sig), _lib.dptr(rcs),
                                   self._flags(trigger_thermo_state, compute_virial, mixed))

    def lj_parameters(self, parameters, rcut):
        """Set the LJ parameters slot without computing (for md_run)."""
        eps = parse_quantity(parameters["epsilon"])
        sig = parse_quantity(parameters["sigma"])
        rc = parse_quantity(rcut)
        self.rcut_max = max(self.rcut_max, rc)
        self.lj_params = (eps, sig)
        self.lj_rcut = rc
        self.ctx.xs_lj_parameters(eps, sig, rc)

    def eam_alloy_init(self, parameters):
        """eam_alloy_init{parameters: {file: ...}} or an in-memory SetflTables."""
        if isinstance(parameters, SetflTables):
            t = parameters
        else:
            f = parameters["file"] if isinstance(parameters, dict) else parameters
            from .potentials import read_setfl
            t = read_setfl(Path(f))
        nz = t.nel * (t.nel + 1) // 2
        if t.z2r.shape != (nz, t.nr):
            raise ValueError("z2r table count does not match the element count")
        fr = np.ascontiguousarray(t.frho, dtype=np.float64)
        rr = np.ascontiguousarray(t.rhor, dtype=np.float64)
        zr = np.ascontiguousarray(t.z2r, dtype=np.float64)
        self.ctx.xs_eam_alloy_tables(t.nel, t.nrho, t.drho, t.nr, t.dr, t.rc, _lib.dptr(fr), _lib.dptr(rr),
                                     _lib.dptr(zr))
        self.eam = t
        self.eam_knots = fr.size + rr.size + zr.size

    def eam_table_info(self) -> dict:
        """The FP32 s = r^2 tables of the mixed single-element force pass: pieces per table
        (0 = none, knot rows), the shortest distance they cover, and their largest FP32
        errors against the FP64 spline in FP32 ulps of the largest |value| or |slope/r|."""
        n = C.c_int(0)
        rlo = C.c_double(0.0)
        err = np.zeros(3)
        self.ctx.xs_eam_table_info(C.byref(n), C.byref(rlo), _lib.dptr(err))
        return {"s_pieces": int(n.value), "r_lo": float(rlo.value),
                "ulps": {"rho_slope_over_r": float(err[0]), "phi": float(err[1]), "phi_slope_over_r": float(err[2])}}

    def gather_ceiling(self, rcut, force_pass, reps=10, level=0) -> float:
        """Average ms of the gather-ceiling kernel of the mixed single-element density
        (force_pass False) or force pass (True) on the current block list: the pass's
        staging, list walk and table-row gathers without its spline arithmetic (level
        0); level 1 without the table-row gathers, 2 the list words only, 3 the staging
        only."""
        ms = C.c_double(0.0)
        self.ctx.xs_gather_ceiling(float(rcut), 1 if force_pass else 0, int(level), int(reps), C.byref(ms))
        return float(ms.value)

    JOHNSON_PARAMETERS = JOHNSON_PARAMETERS

    def johnson_force(self, rcut, parameters, compute_virial=False, **_slots):
        """johnson_force{rcut, parameters: {re, fe, rhoe, alpha, beta, A, B, kappa, lambda,
        Fn0..Fn3, F0..F3, Fo, eta}} (registered at eam_potential.cu:189): both passes."""
        self._johnson(rcut, parameters, True, True, compute_virial)

    def johnson_emb(self, rcut, parameters, **_slots):
        """johnson_emb (eam_potential.cu:190): the embedding pass only; F(rho_i) into ep,
        F'(rho_i) kept for johnson_force_reuse_emb."""
        self._johnson(rcut, parameters, True, False, False)

    def johnson_force_reuse_emb(self, rcut, parameters, compute_virial=False, **_slots):
        """johnson_force_reuse_emb (eam_potential.cu:191): the force pass on the F'(rho_i)
        of the last johnson_emb (an error if the atoms were reordered since)."""
        self._johnson(rcut, parameters, False, True, compute_virial)

    def _johnson(self, rcut, parameters, emb, force, compute_virial):
        """The Johnson operators: analytic Johnson EAM through the single-material
        template (eam_potential.cu:62-193, johnson.h).  Every parameter is converted with
        its unit, as the reference's YAML decoder does (johnson.h:185-205).  Energy is
        always accumulated (EmbOp and ForceOp add to ep unconditionally); the virial when
        compute_virial stands for a grid with a virial field.  FP64 only."""
        rc = parse_quantity(rcut)
        try:
            p = np.array([parse_quantity(parameters[k]) for k in self.JOHNSON_PARAMETERS], dtype=np.float64)
        except KeyError as e:
            raise ValueError(f"johnson_force: parameter {e.args[0]} missing") from None
        self.rcut_max = max(self.rcut_max, rc)
        self.ctx.xs_johnson_force(_lib.dptr(p), rc, int(emb), int(force),
                                  XS_ENERGY | (XS_VIRIAL if compute_virial else 0))

    def eam_alloy_force(self, rcut, eam_rho=True, eam_rho2emb=True, eam_ghost=True, eam_force=True,
                        eam_symmetry=False, compute_virial=False, trigger_thermo_state=False,
                        mixed=False, types=None, **_slots):
        """eam_alloy_force (eam_potential_multimat.cu:88-109 slots).  eam_ghost is
        accepted and has no effect (one GPU holds every atom; images are resolved in
        the pair kernels).  eam_symmetry selects the Newton-symmetric evaluation
        (each pair once from a half list, the partner's share added atomically):
        the same quantities up to FP64 summation order.  types (eam_potential_multimat.cu:90,
        154-166): an empty list means every species; otherwise only pairs of two listed
        species contribute, and atoms of an unlisted species get rho = F' = 0 and no force."""
        mask = self._types_mask(types)
        if mask != self._types_set:
            self.ctx.xs_eam_alloy_types(mask)
            self._types_set = mask
        rc = parse_quantity(rcut)
        self.rcut_max = max(self.rcut_max, rc)
        flags = self._flags(trigger_thermo_state, compute_virial, mixed) | (XS_SYMMETRIC if eam_symmetry else 0)
        self.ctx.xs_eam_alloy_force(rc, int(bool(eam_rho)), int(bool(eam_rho2emb)), int(bool(eam_force)), flags)

    def _types_mask(self, types) -> int:
        """types slot -> enable bits per species (bit t = species t); empty = all."""
        if not types:
            return 0xFFFFFFFF
        unknown = [t for t in types if t not in self.species_names]
        if unknown:
            raise ValueError(f"eam_alloy_force: unknown types {unknown}")
        return sum(1 << self.species_names.index(t) for t in set(types))

    def eam_adjoint(self, dL_dF, dL_dE, mixed=False, rcut=None):
        """dL/dtheta for the knots (setfl order: F, rho, z2r tables) given dL/dF
        (n, 3; forces in internal units) and dL/dE per replica.  Needs a preceding
        eam_alloy_force at the current positions (its rho/F' checkpoint)."""
        t = self.eam
        if t is None:
            raise ValueError("eam_adjoint needs eam_alloy_init first")
        g = np.ascontiguousarray(np.asarray(dL_dF, dtype=np.float64).reshape(self.n, 3))
        ge = np.ascontiguousarray(np.asarray(dL_dE, dtype=np.float64).reshape(-1))
        if ge.size != self.cfg.nrep:
            raise ValueError("dL_dE needs one value per replica")
        out = np.zeros(t.theta().size)
        rc = parse_quantity(rcut) if rcut is not None else t.rc
        self.ctx.xs_eam_adjoint(rc, _lib.dptr(g), _lib.dptr(ge), _lib.dptr(out), XS_MIXED if mixed else 0)
        return out

    def replica_energies(self) -> np.ndarray:
        """Total potential energy of each replica (from the per-atom ep)."""
        ep = self.ep()
        rep = self.cfg.rep if self.cfg.rep is not None else np.zeros(self.n, dtype=np.int64)
        return np.bincount(rep, weights=ep, minlength=self.cfg.nrep)

    def push_f_v_r(self, dt, dt_scale=1.0, xform_mode="INV_XFORM"):
        """push_f_v_r{dt_scale, xform_mode}: INV_XFORM maps the physical step into grid
        space by the inverse xform, IDENTITY adds it as is."""
        modes = {"INV_XFORM": _lib.XS_XFORM_INV, "IDENTITY": _lib.XS_XFORM_IDENTITY}
        if xform_mode not in modes:
            raise ValueError(f"push_f_v_r: xform_mode {xform_mode!r} (INV_XFORM or IDENTITY)")
        self.ctx.xs_push_f_v_r(parse_quantity(dt), float(dt_scale), modes[xform_mode])

    def push_f_v(self, dt, dt_scale=1.0, xform_mode="IDENTITY"):
        """push_f_v{dt_scale, xform_mode}: v += a dt in physical space (the scheme's
        IDENTITY mode; a mapped velocity kick is not part of this path)."""
        if xform_mode != "IDENTITY" and not np.array_equal(self.xform, np.eye(3)):
            raise ValueError("push_f_v: only xform_mode IDENTITY under a non-identity xform")
        self.ctx.xs_push_f_v(parse_quantity(dt), float(dt_scale))

    def simulation_thermodynamic_state(self) -> ThermodynamicState:
        out = np.zeros(27)
        self.ctx.xs_thermo(_lib.dptr(out))
        return ThermodynamicState(out, self.volume())

    def volume(self) -> float:
        """det(xform) times the grid box volume (simulation_thermodynamic_state.cpp:202-211)."""
        c = self.cfg
        if c is None:
            return 0.0
        return float(np.prod(c.hi - c.lo, axis=1).sum()) * float(np.linalg.det(self.xform))

    # --- fused NVE loop -------------------------------------------------------
    def md_run(self, potential, nsteps, dt, rcut, skin, mixed=False, thermo_every=0):
        """verlet_nve for nsteps (with first_iteration); returns (stats, thermo list)."""
        pot = {"lj": _lib.XS_POT_LJ, "eam": _lib.XS_POT_EAM}[potential]
        nth = (nsteps // thermo_every + 1) if thermo_every > 0 else 0
        out = np.zeros(max(nth, 1) * 27)
        st = _lib.RunStats()
        flags = XS_MIXED if mixed else 0
        self.ctx.xs_md_run(pot, int(nsteps), parse_quantity(dt), parse_quantity(rcut), parse_quantity(skin),
                           flags, int(thermo_every), _lib.dptr(out) if nth else None, C.byref(st))
        states = [ThermodynamicState(out[27 * k:27 * k + 27].copy(), self.volume()) for k in range(nth)]
        return st, states

    # --- read back (id order) --------------------------------------------------
    def state(self):
        n = self.n
        arrs = [np.zeros(n) for _ in range(6)]
        self.ctx.xs_system_get(*[_lib.dptr(a) for a in arrs])
        return np.stack(arrs[:3], axis=1), np.stack(arrs[3:], axis=1)

    def set_velocities(self, v):
        cols = [np.ascontiguousarray(v[:, a], dtype=np.float64) for a in range(3)]
        self.ctx.xs_velocity_set(*[_lib.dptr(c) for c in cols])

    def accelerations(self) -> np.ndarray:
        out = np.zeros(3 * self.n)
        self.ctx.xs_field_get(_lib.XS_FIELD_ACCEL, _lib.dptr(out))
        return out.reshape(3, self.n).T.copy()

    def forces(self) -> np.ndarray:
        """Accelerations times masses (valid after force_to_accel)."""
        m = np.asarray(self.masses)[self.cfg.type]
        return self.accelerations() * m[:, None]

    def ep(self) -> np.ndarray:
        out = np.zeros(self.n)
        self.ctx.xs_field_get(_lib.XS_FIELD_EP, _lib.dptr(out))
        return out

    def virial(self) -> np.ndarray:
        out = np.zeros(9 * self.n)
        self.ctx.xs_field_get(_lib.XS_FIELD_VIRIAL, _lib.dptr(out))
        return out.reshape(self.n, 3, 3)

    def rho_dEmb(self) -> np.ndarray:
        out = np.zeros(self.n)
        self.ctx.xs_field_get(_lib.XS_FIELD_RHO_DEMB, _lib.dptr(out))
        return out

    def chunk_neighbors_stats(self, rcut) -> dict:
        """List entries and entries within rcut at the current positions (the
        chunk_neighbors_stats diagnostic, data/samples/perf/microjet_lj.msp:25-28)."""
        e, w = C.c_int64(0), C.c_int64(0)
        self.ctx.xs_pair_count(parse_quantity(rcut), C.byref(e), C.byref(w))
        return {"entries": e.value, "within_rcut": w.value}

    def list_info(self) -> dict:
        """{'kind': 'global'|'block', 'blocks': bricks, 'halo_capacity': LDS slots,
        'central_max': most central atoms of a brick, 'brick': cells per brick (x, y, z)}."""
        kind, blocks, cap, cen = C.c_int(0), C.c_int64(0), C.c_int(0), C.c_int(0)
        brick = (C.c_int * 3)()
        self.ctx.xs_list_info(C.byref(kind), C.byref(blocks), C.byref(cap), C.byref(cen), brick)
        return {"kind": ("global", "block")[kind.value], "blocks": blocks.value, "halo_capacity": cap.value,
                "central_max": cen.value, "brick": tuple(brick)}

    def neighbors(self):
        """(counts[n], ids[total], imgs[total,3]) in id space, entries in list order."""
        counts = np.zeros(self.n, dtype=np.int64)
        total = C.c_int64(0)
        self.ctx.xs_neighbors_get(_lib.ptr(counts, C.c_int64), C.byref(total), None, None)
        ids = np.zeros(total.value, dtype=np.int64)
        imgs = np.zeros((total.value, 3), dtype=np.int32)
        self.ctx.xs_neighbors_get(_lib.ptr(counts, C.c_int64), C.byref(total), _lib.ptr(ids, C.c_int64),
                                  _lib.ptr(imgs, C.c_int32))
        return counts, ids, imgs

    def order(self) -> np.ndarray:
        perm = np.zeros(self.n, dtype=np.int64)
        self.ctx.xs_order_get(_lib.ptr(perm, C.c_int64))
        return perm

    def close(self):
        self.ctx.close()


__all__ = ["Simulation", "ThermodynamicState", "XsError"]
