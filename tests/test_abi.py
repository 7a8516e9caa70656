"""The C-ABI library loads and exports every entry point include/xsmd.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import re

import pytest
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "xsmd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ("xs_ctx_create", "xs_chunk_neighbors", "xs_lj_compute_force", "xs_eam_alloy_force",
                 "xs_push_f_v_r", "xs_push_f_v", "xs_thermo", "xs_md_run"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    from exastamp_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    from exastamp_amd import _lib
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_binding_loads_with_signatures():
    from exastamp_amd import _lib
    lib = _lib.load()
    assert lib.xs_version() >= 1


def test_spline_adjoint_is_the_transpose_of_interpolate():
    """Dot-product test: <G, M f> == <M^T G, f> with M the oracle's knot -> (c3..c6)
    map (interpolate) and M^T the library's xs_spline_adjoint (host code, no GPU)."""
    import ctypes as C
    import numpy as np
    import pyoracle
    from exastamp_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(7)
    for n, delta in ((5, 0.1), (6, 0.3), (40, 0.01), (501, 0.002)):
        f = rng.normal(size=n)
        G = rng.normal(size=(n + 1, 4))
        G[0] = 0.0
        rows = pyoracle.interpolate(f, delta)
        lhs = float((G * rows[:, 3:7]).sum())
        out = np.zeros(n)
        Gc = np.ascontiguousarray(G)
        assert lib.xs_spline_adjoint(n, delta, Gc.ctypes.data_as(C.POINTER(C.c_double)),
                                     out.ctypes.data_as(C.POINTER(C.c_double))) == 0
        assert float(out @ f) == pytest.approx(lhs, rel=1e-12, abs=1e-12)
