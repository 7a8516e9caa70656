"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
OR_POT_LJ, OR_POT_EAM, OR_POT_LJ_MULTI, OR_POT_JOHNSON = 0, 1, 2, 3
_POT = {"lj": OR_POT_LJ, "eam": OR_POT_EAM, "lj_multi": OR_POT_LJ_MULTI, "johnson": OR_POT_JOHNSON}
OR_ENERGY, OR_VIRIAL = 1, 2

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built: run `make -C oracle`")
        lib = C.CDLL(str(LIB_PATH))
        sig = {
            "or_interpolate": (C.c_int, [C.c_int, C.c_double, _D, _D]),
            "or_setfl_header": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _D,
                                          C.POINTER(C.c_int), _D, _D]),
            "or_setfl_data": (C.c_int, [C.c_char_p, _D, _D, _D, _D]),
            "or_create": (C.c_void_p, []),
            "or_destroy": (None, [C.c_void_p]),
            "or_set_threads": (C.c_int, [C.c_void_p, C.c_int]),
            "or_set_system": (C.c_int, [C.c_void_p, C.c_int64, _D, _D, _D, _D, _D, _D, C.POINTER(C.c_uint8),
                                        _D, _D, C.POINTER(C.c_int), C.c_int, _D]),
            "or_set_lj": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
            "or_set_lj_multi": (C.c_int, [C.c_void_p, C.c_int, _D, _D, _D]),
            "or_set_johnson": (C.c_int, [C.c_void_p, _D, C.c_double]),
            "or_johnson_eval": (None, [_D, C.c_double, _D]),
            "or_johnson_embed": (None, [_D, C.c_double, _D, _D]),
            "or_set_eam": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                     _D, _D, _D]),
            "or_set_eam_types": (C.c_int, [C.c_void_p, C.c_uint]),
            "or_build_neighbors": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
            "or_compute_forces": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "or_eam_pass": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
            "or_zero_force_energy": (C.c_int, [C.c_void_p]),
            "or_force_to_accel": (C.c_int, [C.c_void_p]),
            "or_push_f_v_r": (C.c_int, [C.c_void_p, C.c_double]),
            "or_push_f_v_r_mode": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
            "or_set_xform": (C.c_int, [C.c_void_p, _D]),
            "or_push_f_v": (C.c_int, [C.c_void_p, C.c_double]),
            "or_displ_over": (C.c_int, [C.c_void_p, C.c_double]),
            "or_thermo": (C.c_int, [C.c_void_p, _D]),
            "or_md_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _D]),
            "or_md_steps": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _D]),
            "or_n": (C.c_int64, [C.c_void_p]),
            "or_get_state": (C.c_int, [C.c_void_p] + [_D] * 12),
            "or_neighbor_count": (C.c_int64, [C.c_void_p, _I64]),
            "or_neighbor_get": (C.c_int, [C.c_void_p, _I64, _I32]),
            "or_get_order": (C.c_int, [C.c_void_p, _I64]),
            "or_brute_neighbor_count": (C.c_int64, [C.c_void_p, C.c_double, _I64]),
            "or_brute_neighbor_get": (C.c_int, [C.c_void_p, C.c_double, _I64, _I32]),
            "or_eam_rho": (C.c_double, [C.c_void_p, C.c_double, C.c_int, C.c_int]),
            "or_eam_pair": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _D, _D]),
            "or_eam_embed": (None, [C.c_void_p, C.c_double, C.c_int, _D, _D]),
            "or_lj_energy": (None, [C.c_void_p, C.c_double, _D, _D]),
            "or_eV": (C.c_double, []),
            "or_kB": (C.c_double, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def interpolate(f: np.ndarray, delta: float) -> np.ndarray:
    """Spline rows (n+1, 8) of the reference's interpolate for knots f."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros((f.size + 1, 8))
    rc = load().or_interpolate(f.size, delta, _d(f), _d(out))
    if rc:
        raise ValueError("or_interpolate failed")
    return out


class Oracle:
    def __init__(self, threads: int = 0):
        self.lib = load()
        self.h = C.c_void_p(self.lib.or_create())
        if threads:
            self.lib.or_set_threads(self.h, threads)
        self.n = 0

    def close(self):
        if self.h:
            self.lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc, what):
        if rc < 0 or (rc != 0 and what != "md_run"):
            raise RuntimeError(f"oracle {what} failed ({rc})")
        return rc

    def set_config(self, cfg, masses):
        if cfg.nrep != 1:
            raise ValueError("the oracle holds one periodic box")
        r = np.ascontiguousarray(cfg.r, dtype=np.float64)
        v = np.ascontiguousarray(cfg.v, dtype=np.float64)
        cols = [np.ascontiguousarray(r[:, a]) for a in range(3)] + [np.ascontiguousarray(v[:, a]) for a in range(3)]
        t = np.ascontiguousarray(cfg.type, dtype=np.uint8)
        lo = np.ascontiguousarray(cfg.lo[0], dtype=np.float64)
        hi = np.ascontiguousarray(cfg.hi[0], dtype=np.float64)
        per = (C.c_int * 3)(*[int(p) for p in cfg.periodic])
        m = np.ascontiguousarray(masses, dtype=np.float64)
        self._ok(self.lib.or_set_system(self.h, cfg.n, *[_d(c) for c in cols], t.ctypes.data_as(C.POINTER(C.c_uint8)),
                                        _d(lo), _d(hi), per, m.size, _d(m)), "set_system")
        self.n = cfg.n
        self.masses = m
        self.types = t.copy()

    def set_lj(self, eps, sigma, rcut):
        self._ok(self.lib.or_set_lj(self.h, eps, sigma, rcut), "set_lj")

    def set_lj_multi(self, ntypes, eps, sigma, rcut):
        """Per unordered type pair (entry b*(b+1)/2 + a for a <= b) LJ parameters."""
        e, s, r = (np.ascontiguousarray(x, dtype=np.float64) for x in (eps, sigma, rcut))
        self._ok(self.lib.or_set_lj_multi(self.h, ntypes, _d(e), _d(s), _d(r)), "set_lj_multi")

    def set_johnson(self, params, rcut):
        """params: the 19 Johnson parameters (internal units, EamJohnsonParameters order)."""
        self._jp = np.ascontiguousarray(params, dtype=np.float64)
        assert self._jp.size == 19
        self._ok(self.lib.or_set_johnson(self.h, _d(self._jp), rcut), "set_johnson")

    def set_eam(self, t):
        fr = np.ascontiguousarray(t.frho, dtype=np.float64)
        rr = np.ascontiguousarray(t.rhor, dtype=np.float64)
        zr = np.ascontiguousarray(t.z2r, dtype=np.float64)
        self._ok(self.lib.or_set_eam(self.h, t.nel, t.nrho, t.drho, t.nr, t.dr, t.rc, _d(fr), _d(rr), _d(zr)),
                 "set_eam")

    def set_eam_types(self, mask=0xFFFFFFFF):
        """eam_alloy_force's types slot as enable bits per species (bit t = species t)."""
        self._ok(self.lib.or_set_eam_types(self.h, mask), "set_eam_types")

    def build_neighbors(self, rcut, skin):
        self._ok(self.lib.or_build_neighbors(self.h, rcut, skin), "build_neighbors")

    def compute_forces(self, potential, flags=OR_ENERGY | OR_VIRIAL):
        pot = _POT.get(potential, potential)
        self._ok(self.lib.or_compute_forces(self.h, pot, flags), "compute_forces")

    def eam_pass(self, rho, rho2emb, force, flags=0):
        self._ok(self.lib.or_eam_pass(self.h, int(rho), int(rho2emb), int(force), flags), "eam_pass")

    def zero_force_energy(self):
        self.lib.or_zero_force_energy(self.h)

    def force_to_accel(self):
        self.lib.or_force_to_accel(self.h)

    def push_f_v_r(self, dt, xform_mode="INV_XFORM"):
        if xform_mode not in ("INV_XFORM", "IDENTITY"):
            raise ValueError(f"push_f_v_r: xform_mode {xform_mode}")
        self.lib.or_push_f_v_r_mode(self.h, dt, int(xform_mode == "INV_XFORM"))

    def set_xform(self, m=None):
        """Domain xform: 3x3 grid -> physical map (None = identity); invalidates the list."""
        a = None if m is None else np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(9))
        self._ok(self.lib.or_set_xform(self.h, None if a is None else _d(a)), "set_xform")

    def push_f_v(self, dt):
        self.lib.or_push_f_v(self.h, dt)

    def displ_over(self, thr) -> bool:
        return bool(self.lib.or_displ_over(self.h, thr))

    def thermo(self) -> np.ndarray:
        out = np.zeros(27)
        self.lib.or_thermo(self.h, _d(out))
        return out

    def md_run(self, potential, nsteps, dt, rcut, skin, thermo=True):
        pot = _POT[potential]
        out = np.zeros((nsteps + 1) * 27) if thermo else None
        nreb = self._ok(self.lib.or_md_run(self.h, pot, nsteps, dt, rcut, skin, _d(out)), "md_run")
        return nreb, (out.reshape(nsteps + 1, 27) if thermo else None)

    def md_steps(self, potential, nsteps, dt, rcut, skin):
        """Continue for nsteps (no initial build/force); returns the rebuild count."""
        pot = _POT[potential]
        rc = self.lib.or_md_steps(self.h, pot, nsteps, dt, rcut, skin, None)
        if rc < 0:
            raise RuntimeError(f"oracle md_steps failed ({rc})")
        return rc

    def state(self):
        n = self.n
        a = [np.zeros(n) for _ in range(10)]
        vir = np.zeros(9 * n)
        rho = np.zeros(n)
        self.lib.or_get_state(self.h, *[_d(x) for x in a], _d(vir), _d(rho))
        return {
            "r": np.stack(a[0:3], 1), "v": np.stack(a[3:6], 1), "a": np.stack(a[6:9], 1), "ep": a[9],
            "virial": vir.reshape(n, 3, 3), "rho_dEmb": rho,
        }

    def forces(self):
        s = self.state()
        return s["a"] * self.masses[self.types][:, None]

    def neighbors(self):
        counts = np.zeros(self.n, dtype=np.int64)
        tot = self.lib.or_neighbor_count(self.h, counts.ctypes.data_as(_I64))
        ids = np.zeros(tot, dtype=np.int64)
        imgs = np.zeros((tot, 3), dtype=np.int32)
        self.lib.or_neighbor_get(self.h, ids.ctypes.data_as(_I64), imgs.ctypes.data_as(_I32))
        return counts, ids, imgs

    def brute_neighbors(self, rlist):
        counts = np.zeros(self.n, dtype=np.int64)
        tot = self.lib.or_brute_neighbor_count(self.h, rlist, counts.ctypes.data_as(_I64))
        ids = np.zeros(tot, dtype=np.int64)
        imgs = np.zeros((tot, 3), dtype=np.int32)
        self.lib.or_brute_neighbor_get(self.h, rlist, ids.ctypes.data_as(_I64), imgs.ctypes.data_as(_I32))
        return counts, ids, imgs

    def order(self):
        perm = np.zeros(self.n, dtype=np.int64)
        self.lib.or_get_order(self.h, perm.ctypes.data_as(_I64))
        return perm

    # scalar evaluators
    def eam_rho(self, r, itype=0, jtype=0):
        return self.lib.or_eam_rho(self.h, r, itype, jtype)

    def eam_pair(self, r, fpi, fpj, itype=0, jtype=0):
        a, b = C.c_double(), C.c_double()
        self.lib.or_eam_pair(self.h, r, fpi, fpj, itype, jtype, C.byref(a), C.byref(b))
        return a.value, b.value

    def eam_embed(self, rho, itype=0):
        a, b = C.c_double(), C.c_double()
        self.lib.or_eam_embed(self.h, rho, itype, C.byref(a), C.byref(b))
        return a.value, b.value

    def lj_energy(self, r):
        a, b = C.c_double(), C.c_double()
        self.lib.or_lj_energy(self.h, r, C.byref(a), C.byref(b))
        return a.value, b.value


def setfl_read(path):
    lib = load()
    nel, nrho, nr = C.c_int(), C.c_int(), C.c_int()
    drho, dr, rc = C.c_double(), C.c_double(), C.c_double()
    p = str(path).encode()
    if lib.or_setfl_header(p, C.byref(nel), C.byref(nrho), C.byref(drho), C.byref(nr), C.byref(dr), C.byref(rc)):
        raise IOError(path)
    ne = nel.value
    fr = np.zeros(ne * nrho.value)
    rr = np.zeros(ne * nr.value)
    zr = np.zeros(ne * (ne + 1) // 2 * nr.value)
    m = np.zeros(ne)
    if lib.or_setfl_data(p, _d(fr), _d(rr), _d(zr), _d(m)):
        raise IOError(path)
    return dict(nel=ne, nrho=nrho.value, drho=drho.value, nr=nr.value, dr=dr.value, rc=rc.value,
                frho=fr.reshape(ne, -1), rhor=rr.reshape(ne, -1), z2r=zr.reshape(-1, nr.value), masses=m)


def johnson_eval(params, r):
    """rho(r), rho'(r), phi(r), phi'(r) of the Johnson functions (johnson.h:59-124)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros(4)
    load().or_johnson_eval(_d(p), float(r), _d(out))
    return out


def johnson_embed(params, rho):
    """F(rho), F'(rho) of the Johnson embedding (johnson.h:131-167)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    F, Fp = C.c_double(), C.c_double()
    load().or_johnson_embed(_d(p), float(rho), C.byref(F), C.byref(Fp))
    return F.value, Fp.value
