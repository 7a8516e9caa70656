"""ctypes binding of libxsmd.so (the C ABI declared in include/xsmd.h).

The library must have been built in-tree (``python -m exastamp_amd.build``);
there is no fallback: loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libxsmd.so"

XS_ENERGY = 1
XS_VIRIAL = 2
XS_MIXED = 4
XS_CONTINUE = 8
XS_SYMMETRIC = 16
XS_POT_LJ = 0
XS_POT_EAM = 1
XS_FIELD_EP = 0
XS_FIELD_VIRIAL = 1
XS_FIELD_RHO_DEMB = 2
XS_FIELD_ACCEL = 3
XS_XFORM_IDENTITY = 0
XS_XFORM_INV = 1

PROF_CLASSES = ["neighbors", "eam_rho", "eam_force", "lj", "drift", "thermo", "adjoint"]

STATUS = {0: "XS_OK", 1: "XS_E_ARG", 2: "XS_E_STATE", 3: "XS_E_HIP", 4: "XS_E_ALLOC", 5: "XS_E_IO",
          6: "XS_E_LIMIT"}


class XsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class RunStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("rebuilds", C.c_int64), ("max_neighbors", C.c_int64),
                ("seconds", C.c_double)]


# name -> (restype, argtypes)
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_U8 = C.POINTER(C.c_uint8)
_U16 = C.POINTER(C.c_uint16)
_CTX = C.c_void_p
SIGNATURES = {
    "xs_version": (C.c_int, []),
    "xs_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "xs_ctx_destroy": (C.c_int, [_CTX]),
    "xs_last_error": (C.c_char_p, [_CTX]),
    "xs_synchronize": (C.c_int, [_CTX]),
    "xs_ctx_set_stream": (C.c_int, [_CTX, C.c_void_p]),
    "xs_ctx_own_stream": (C.c_int, [_CTX]),
    "xs_device_memory": (C.c_int, [_CTX, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "xs_species_set": (C.c_int, [_CTX, C.c_int, _D]),
    "xs_system_set": (C.c_int, [_CTX, C.c_int64, _D, _D, _D, _D, _D, _D, _U8, _U16, C.c_int, _D, _D,
                                C.POINTER(C.c_int)]),
    "xs_system_size": (C.c_int64, [_CTX]),
    "xs_system_get": (C.c_int, [_CTX, _D, _D, _D, _D, _D, _D]),
    "xs_field_get": (C.c_int, [_CTX, C.c_int, _D]),
    "xs_velocity_set": (C.c_int, [_CTX, _D, _D, _D]),
    "xs_chunk_neighbors": (C.c_int, [_CTX, C.c_double, C.c_double, C.c_int]),
    "xs_particle_displ_over": (C.c_int, [_CTX, C.c_double, C.POINTER(C.c_int)]),
    "xs_neighbors_get": (C.c_int, [_CTX, _I64, _I64, _I64, _I32]),
    "xs_order_get": (C.c_int, [_CTX, _I64]),
    "xs_pair_count": (C.c_int, [_CTX, C.c_double, _I64, _I64]),
    "xs_list_info": (C.c_int, [_CTX, C.POINTER(C.c_int), _I64, C.POINTER(C.c_int), C.POINTER(C.c_int),
                               C.POINTER(C.c_int)]),
    "xs_zero_force_energy": (C.c_int, [_CTX]),
    "xs_force_to_accel": (C.c_int, [_CTX]),
    "xs_lj_compute_force": (C.c_int, [_CTX, C.c_double, C.c_double, C.c_double, C.c_uint]),
    "xs_lj_parameters": (C.c_int, [_CTX, C.c_double, C.c_double, C.c_double]),
    "xs_lj_multi_force": (C.c_int, [_CTX, C.c_int, _D, _D, _D, C.c_uint]),
    "xs_eam_alloy_tables": (C.c_int, [_CTX, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                                      _D, _D, _D]),
    "xs_eam_alloy_load": (C.c_int, [_CTX, C.c_char_p, _D, C.POINTER(C.c_int)]),
    "xs_eam_alloy_force": (C.c_int, [_CTX, C.c_double, C.c_int, C.c_int, C.c_int, C.c_uint]),
    "xs_eam_alloy_types": (C.c_int, [_CTX, C.c_uint]),
    "xs_eam_table_info": (C.c_int, [_CTX, C.POINTER(C.c_int), _D, _D]),
    "xs_johnson_force": (C.c_int, [_CTX, _D, C.c_double, C.c_int, C.c_int, C.c_uint]),
    "xs_push_f_v_r": (C.c_int, [_CTX, C.c_double, C.c_double, C.c_int]),
    "xs_domain_xform": (C.c_int, [_CTX, _D]),
    "xs_push_f_v": (C.c_int, [_CTX, C.c_double, C.c_double]),
    "xs_thermo": (C.c_int, [_CTX, _D]),
    "xs_thermo_noalloc": (C.c_int, [_CTX, C.c_void_p]),
    "xs_eam_adjoint": (C.c_int, [_CTX, C.c_double, _D, _D, _D, C.c_uint]),
    "xs_eam_adjoint_dev": (C.c_int, [_CTX, C.c_double, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_uint]),
    "xs_eam_theta_dev": (C.c_int, [_CTX, C.c_void_p]),
    "xs_forces_energies_dev": (C.c_int, [_CTX, C.c_void_p, C.c_void_p, C.c_double]),
    "xs_lj_adjoint": (C.c_int, [_CTX, C.c_double, C.c_double, C.c_double, _D, _D, _D, C.c_uint]),
    "xs_spline_adjoint": (C.c_int, [C.c_int, C.c_double, _D, _D]),
    "xs_stream_copy": (C.c_int, [_CTX, C.c_size_t, C.c_int, _D]),
    "xs_gather_ceiling": (C.c_int, [_CTX, C.c_double, C.c_int, C.c_int, C.c_int, _D]),
    "xs_profile_enable": (C.c_int, [_CTX, C.c_int]),
    "xs_profile_get": (C.c_int, [_CTX, _D, _I64]),
    "xs_md_run": (C.c_int, [_CTX, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint, C.c_int,
                            _D, C.POINTER(RunStats)]),
}

_lib = None


def load() -> C.CDLL:
    """Load libxsmd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built: run `python -m exastamp_amd.build`")
        lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    """Owns one xs_ctx on a device; checks every status."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.xs_ctx_create(device, C.byref(h))
        if rc != 0:
            raise XsError(rc, f"xs_ctx_create(device={device}) failed (is a GPU visible?)")
        self.h = h

    def check(self, rc: int):
        if rc != 0:
            msg = self.lib.xs_last_error(self.h)
            raise XsError(rc, msg.decode() if msg else "")

    def __getattr__(self, name):
        # self.xs_foo(args...) -> checked call of xs_foo(ctx, args...)
        if not name.startswith("xs_"):
            raise AttributeError(name)
        f = getattr(self.lib, name)

        def call(*args):
            return self.check(f(self.h, *args))
        return call

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.xs_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
