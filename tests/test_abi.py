"""The C ABI library loads and exports every entry point of include/hetd.h
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

from hetd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "hetd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hetd_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert set(declared()) == set(capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("lib/libhetd.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_host_side_setup_entry_points():
    # pure host functions of the ABI: impurity potential and its table
    L = capi.load()
    v = ctypes.c_double()
    assert L.hetd_vimp(b"Xe_He", 2.0, 1.0, 0.0, ctypes.byref(v)) == 0
    assert abs(v.value - 21450.4867067095) < 1e-9
    assert L.hetd_vimp(b"Xe_He", 2.0, 2.55, 2793.9365463, ctypes.byref(v)) == 0
    assert v.value == 2793.9365463
    assert L.hetd_vimp(b"no_such_potential", 2.0, 2.0, 0.0, ctypes.byref(v)) == capi.HETD_ERR_ARG
    from hetd import build_potion
    tab, d = build_potion("Xe_He", 2.55, 2793.9365463, 32, 16.0)
    assert len(tab) == 3201 and abs(d - 64.0 / 3200) < 1e-15
    import oracle
    otab, od = oracle.build_potion("Xe_He", 2.55, 2793.9365463, 32, 16.0)
    assert od == d and (abs(otab - tab) <= 1e-13 * abs(otab).max()).all()


def test_default_params_match_python_mirror():
    L = capi.load()
    p = capi.Params()
    L.hetd_default_params(ctypes.byref(p))
    q = capi.default_params()
    for name, _ in capi.Params._fields_:
        assert getattr(p, name) == getattr(q, name), name


@pytest.mark.parametrize("field,value,needle", [
    ("npd", 4, "npd must be odd"),
    ("icon", 7, "icon must be"),
    ("nx", 2048, "above 1024"),
])
def test_failed_init_leaves_an_inert_context(field, value, needle):
    # the configuration checks come before any HIP call, so this runs
    # without a GPU: the context reports why and refuses every other call
    L = capi.load()
    p = capi.default_params(nx=32, ny=32, nz=32, xmax=6.4, ymax=6.4, zmax=6.4)
    setattr(p, field, value)
    h = ctypes.c_void_p()
    assert L.hetd_init(ctypes.byref(p), ctypes.byref(h)) == capi.HETD_ERR_ARG
    assert h.value
    assert needle in L.hetd_last_error(h).decode()
    d = ctypes.c_double()
    assert L.hetd_poten(h) == capi.HETD_ERR_STATE
    assert L.hetd_steprk(h, 1e-4) == capi.HETD_ERR_STATE
    assert L.hetd_steppc(h, 1e-4, ctypes.byref(d), ctypes.byref(d)) == capi.HETD_ERR_STATE
    assert L.hetd_norm(h, ctypes.byref(d)) == capi.HETD_ERR_STATE
    assert L.hetd_evolve(h, 1e-4, 1, None) == capi.HETD_ERR_STATE
    assert L.hetd_set_uext(h, None) == capi.HETD_ERR_STATE
    assert L.hetd_history_size(h) == -1
    arr = (ctypes.c_void_p * 1)(h)
    assert L.hetd_pool_evolve(arr, 1, 1e-4, 1) == capi.HETD_ERR_STATE
    assert needle in L.hetd_last_error(h).decode()
    assert L.hetd_finalize(h) == capi.HETD_OK


def test_python_binding_checks_before_the_device():
    # both checks come before any HIP call
    import numpy as np
    from hetd import Context
    p = capi.default_params(nx=16, ny=16, nz=16, xmax=3.2, ymax=3.2, zmax=3.2)
    with pytest.raises(ValueError, match="unknown impurity potential"):
        Context(p, selec="no_such_potential")
    with pytest.raises(ValueError, match="expected"):
        Context._host(np.zeros((16, 16, 15)), np.complex128, (16, 16, 16), "psi")
    a = Context._host(np.zeros((16, 16, 16)), np.complex128, (16, 16, 16), "psi")
    assert a.dtype == np.complex128 and a.flags.c_contiguous
