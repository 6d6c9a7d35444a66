"""ctypes mirror of include/hetd.h (structures, constants, prototypes).

The product's compute path is lib/libhetd.so (HIP kernels for gfx950); this
module only declares its C ABI for Python callers.  Loading fails loudly if
the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HETD_OK = 0
HETD_ERR_ARG = 1
HETD_ERR_HIP = 2
HETD_ERR_GUARD = 3
HETD_ERR_STATE = 4

HETD_FP64 = 0
HETD_MIXED = 1
HETD_CORE_OT = 0
HETD_CORE_OTC = 1

F_PSI, F_DEN, F_POT4, F_DENCG, F_DELJ4, F_CCOR, F_DXDEN, F_DYDEN, F_DZDEN, F_UIMP, \
    F_PENALTY, F_UALPHAS, F_DENALF, F_FVLJ, F_WCGK, F_PC, F_Q, F_FDEN = range(18)
COMPLEX_FIELDS = (F_PSI, F_PC, F_Q)
HALF_FIELDS = (F_FVLJ, F_WCGK)
HALF_COMPLEX_FIELDS = (F_FDEN,)
TRAJ_STRIDE = 11
K_OFF, K_UPDATE, K_XPASS, K_YPASS, K_ZPASS, K_FORCE = range(6)
K_NAMES = {K_UPDATE: "update", K_XPASS: "xpass", K_YPASS: "ypass", K_ZPASS: "zpass", K_FORCE: "force"}

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libhetd.so")


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("npd", C.c_int32), ("icon", C.c_int32), ("core", C.c_int32),
        ("lsolid", C.c_int32), ("precision", C.c_int32),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("eps4", C.c_double), ("sigma4", C.c_double), ("h4", C.c_double),
        ("cp4", C.c_double), ("cpp4", C.c_double), ("h2o2m4", C.c_double),
        ("alphas", C.c_double), ("l_alpha", C.c_double), ("den0s", C.c_double),
        ("c_solid", C.c_double), ("beta_solid", C.c_double), ("den_m", C.c_double),
        ("mimpur", C.c_double),
        ("Lambdah", C.c_double), ("txmean", C.c_double), ("txsurf", C.c_double),
        ("tymean", C.c_double), ("tysurf", C.c_double), ("tzmean", C.c_double),
        ("tzsurf", C.c_double),
    ]


class Energies(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("etot4", "ekin4", "elj4", "ecor4", "esolid", "ealphas",
                 "ekinx", "eHeX", "eimpu", "etot")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def default_params(**kw) -> Params:
    """Reference defaults (modules.f90:130-150, 204-216; DFT4He3d.f90:52-145).

    Mirrors hetd_default_params() so that the oracle can be driven without
    loading the product library.
    """
    p = Params()
    p.nx = p.ny = p.nz = 64
    p.npd = 13
    p.icon = 13
    p.core = HETD_CORE_OT
    p.lsolid = 1
    p.precision = HETD_FP64
    p.xmax = p.ymax = p.zmax = 64.0
    p.eps4, p.sigma4, p.h4 = 10.22, 2.556, 2.190323
    p.cp4, p.cpp4, p.h2o2m4 = -2.41186e4, 1.85850e6, 6.05969638298
    p.alphas, p.l_alpha, p.den0s = 54.31, 1.0, 0.04
    p.c_solid, p.beta_solid, p.den_m = 3.1577504e4, 40.0, 0.37
    p.mimpur = 131.29
    p.Lambdah = 2.0
    p.txmean = p.tymean = p.tzmean = 1000.0
    p.txsurf = p.tysurf = p.tzsurf = 2.0
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        if k == "core" and isinstance(v, str):
            v = HETD_CORE_OTC if v.strip().upper() == "OTC" else HETD_CORE_OT
        setattr(p, k, v)
    return p


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


def _proto(lib):
    P = C.POINTER(Params)
    lib.hetd_default_params.argtypes = [P]
    lib.hetd_default_params.restype = None
    lib.hetd_init.argtypes = [P, C.POINTER(_vp)]
    lib.hetd_finalize.argtypes = [_vp]
    lib.hetd_last_error.argtypes = [_vp]
    lib.hetd_last_error.restype = C.c_char_p
    lib.hetd_vimp.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, _dp]
    lib.hetd_build_potion.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_int32,
                                      C.c_double, _dp, C.POINTER(C.c_int32), _dp]
    lib.hetd_set_potion.argtypes = [_vp, _dp, C.c_int32, C.c_double]
    lib.hetd_vsurf.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, _dp]
    lib.hetd_set_uext.argtypes = [_vp, _dp]
    lib.hetd_set_surface.argtypes = [_vp, C.c_int32, C.c_double]
    lib.hetd_set_state.argtypes = [_vp, _dp, _dp, _dp]
    lib.hetd_potenimp.argtypes = [_vp, _dp]
    lib.hetd_poten.argtypes = [_vp]
    lib.hetd_forceimp.argtypes = [_vp, _dp, _dp]
    lib.hetd_set_aimp.argtypes = [_vp, _dp]
    lib.hetd_steprk.argtypes = [_vp, C.c_double]
    lib.hetd_steppc.argtypes = [_vp, C.c_double, _dp, _dp]
    lib.hetd_energy.argtypes = [_vp, C.POINTER(Energies)]
    lib.hetd_norm.argtypes = [_vp, _dp]
    lib.hetd_moments.argtypes = [_vp, C.c_double, _dp]
    lib.hetd_history_size.argtypes = [_vp]
    lib.hetd_history_size.restype = C.c_int64
    lib.hetd_save_history.argtypes = [_vp, _vp]
    lib.hetd_load_history.argtypes = [_vp, _vp, C.c_int64]
    lib.hetd_evolve.argtypes = [_vp, C.c_double, C.c_int32, _dp]
    lib.hetd_autotune.argtypes = [_vp, C.c_double, C.c_int32, C.POINTER(C.c_int32)]
    lib.hetd_pool_evolve.argtypes = [C.POINTER(_vp), C.c_int32, C.c_double, C.c_int32]
    lib.hetd_get_field.argtypes = [_vp, C.c_int32, _dp]
    lib.hetd_get_imp.argtypes = [_vp, _dp, _dp, _dp]
    lib.hetd_fft_r2c.argtypes = [_vp, _dp, _dp]
    lib.hetd_fft_c2r.argtypes = [_vp, _dp, _dp]
    lib.hetd_deriv.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32, _dp, _dp]
    lib.hetd_stream.argtypes = [_vp]
    lib.hetd_stream.restype = _vp
    lib.hetd_timing.argtypes = [_vp, C.c_int32]
    lib.hetd_timing.restype = C.c_int
    lib.hetd_timing_read.argtypes = [_vp, _dp, C.POINTER(C.c_int64), _dp]
    lib.hetd_timing_read.restype = C.c_int
    for name in ("hetd_init", "hetd_finalize", "hetd_vimp", "hetd_build_potion", "hetd_vsurf",
                 "hetd_set_uext", "hetd_set_surface",
                 "hetd_set_potion", "hetd_set_state", "hetd_potenimp", "hetd_poten",
                 "hetd_forceimp", "hetd_set_aimp", "hetd_steprk", "hetd_steppc",
                 "hetd_energy", "hetd_norm", "hetd_moments", "hetd_save_history", "hetd_load_history",
                 "hetd_evolve", "hetd_pool_evolve", "hetd_autotune", "hetd_get_field",
                 "hetd_get_imp", "hetd_fft_r2c", "hetd_fft_c2r", "hetd_deriv"):
        getattr(lib, name).restype = C.c_int
    return lib


EXPORTED = ("hetd_default_params", "hetd_init", "hetd_finalize", "hetd_last_error",
            "hetd_vimp", "hetd_build_potion", "hetd_set_potion", "hetd_vsurf", "hetd_set_uext",
            "hetd_set_surface", "hetd_set_state",
            "hetd_potenimp", "hetd_poten", "hetd_forceimp", "hetd_set_aimp",
            "hetd_steprk", "hetd_steppc", "hetd_energy", "hetd_norm", "hetd_moments", "hetd_evolve",
            "hetd_history_size", "hetd_save_history", "hetd_load_history", "hetd_pool_evolve", "hetd_autotune",
            "hetd_get_field", "hetd_get_imp", "hetd_fft_r2c", "hetd_fft_c2r",
            "hetd_deriv", "hetd_stream", "hetd_timing", "hetd_timing_read")

_LIB = None


def load(path: str | None = None):
    """Load lib/libhetd.so (built by __graft_entry__.build()); raise if absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("libhetd.so not built at %s: run __graft_entry__.build()" % p)
    lib = _proto(C.CDLL(p))
    if path is None:
        _LIB = lib
    return lib
