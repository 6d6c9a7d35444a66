"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this module, and only as the checker.  The interface mirrors
hetd.Context so that parity tests read the same on both sides.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "4hetddft-isotropic_amd"))
from hetd import capi  # noqa: E402  (structures only)

LIB_PATH = os.path.join(HERE, "liboracle.so")
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(capi.Params)
        L.orc_create.argtypes = [P]
        L.orc_create.restype = _vp
        L.orc_destroy.argtypes = [_vp]
        L.orc_destroy.restype = None
        L.orc_error.argtypes = [_vp]
        L.orc_error.restype = C.c_char_p
        L.orc_set_potion.argtypes = [_vp, _dp, C.c_int, C.c_double]
        L.orc_set_state.argtypes = [_vp, _dp, _dp, _dp]
        L.orc_init.argtypes = [_vp]
        L.orc_potenimp.argtypes = [_vp, _dp]
        L.orc_poten.argtypes = [_vp]
        L.orc_forceimp.argtypes = [_vp, _dp, _dp]
        L.orc_set_aimp.argtypes = [_vp, _dp]
        L.orc_steprk.argtypes = [_vp, C.c_double]
        L.orc_steppc.argtypes = [_vp, C.c_double, _dp, _dp]
        L.orc_energy.argtypes = [_vp, C.POINTER(capi.Energies)]
        L.orc_norm.argtypes = [_vp, _dp]
        L.orc_moments.argtypes = [_vp, C.c_double, _dp]
        L.orc_evolve.argtypes = [_vp, C.c_double, C.c_int, _dp]
        L.orc_get_field.argtypes = [_vp, C.c_int, _dp]
        L.orc_get_imp.argtypes = [_vp, _dp, _dp, _dp]
        L.orc_fft_r2c.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_fft_c2r.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_deriv.argtypes = [C.c_int] * 6 + [C.c_int, C.c_double, C.c_int, _dp, _dp]
        L.orc_si.argtypes = [C.c_double]
        L.orc_si.restype = C.c_double
        L.orc_vlj_hat.argtypes = [C.c_double] * 5
        L.orc_vlj_hat.restype = C.c_double
        L.orc_vimp.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_vimp.restype = C.c_double
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def _v3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


SELEC = {"null": 0, "Xe_He": 1}


def build_potion(selec, r_cutoff, umax, nmax, xmaxbox):
    """Table of potenimpini (potenimpini.f90:11-25) from the oracle's V."""
    npot = 100 * nmax + 1
    rmax = 4.0 * xmaxbox
    d = rmax / float(npot - 1)
    L = lib()
    s = SELEC[selec]
    tab = np.array([L.orc_vimp(s, float(i) * d, r_cutoff, umax) for i in range(npot)])
    return tab, d


class OracleCtx:
    """Same methods as hetd.Context, computed by the C restatement."""

    def __init__(self, params: "capi.Params", selec="null", r_cutoff=2.0, umax=0.0, potion=None):
        """potion: (table, spacing) of V(r) given from outside (a potential the
        oracle does not restate; the table itself is pinned against the
        reference's Select_Pot on its own), else built from the oracle's V."""
        self.L = lib()
        self.p = params
        self.h = self.L.orc_create(C.byref(params))
        self.shape = (params.nz, params.ny, params.nx)
        self.n = params.nx * params.ny * params.nz
        self.nh = (params.nx // 2 + 1) * params.ny * params.nz
        if potion is not None:
            tab, d = np.ascontiguousarray(potion[0], dtype=np.float64), float(potion[1])
        else:
            tab, d = build_potion(selec, r_cutoff, umax, max(params.nx, params.ny, params.nz),
                                  max(params.xmax, params.ymax, params.zmax))
        self.potion, self.delinter = tab, d
        self._chk(self.L.orc_set_potion(self.h, _p(tab), len(tab), d))

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("oracle error %d: %s" % (rc, self.L.orc_error(self.h).decode()))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, psi, r, v):
        a = np.ascontiguousarray(psi, dtype=np.complex128).view(np.float64)
        self._chk(self.L.orc_set_state(self.h, _p(a), _p(_v3(r)), _p(_v3(v))))

    def init(self):
        self._chk(self.L.orc_init(self.h))

    def potenimp(self, r):
        self._chk(self.L.orc_potenimp(self.h, _p(_v3(r))))

    def poten(self):
        self._chk(self.L.orc_poten(self.h))

    def forceimp(self, r):
        F = np.zeros(3)
        self._chk(self.L.orc_forceimp(self.h, _p(_v3(r)), _p(F)))
        return F

    def set_aimp(self, a):
        self._chk(self.L.orc_set_aimp(self.h, _p(_v3(a))))

    def steprk(self, dt):
        self._chk(self.L.orc_steprk(self.h, dt))

    def steppc(self, dt):
        a, b = C.c_double(), C.c_double()
        self._chk(self.L.orc_steppc(self.h, dt, C.byref(a), C.byref(b)))
        return a.value, b.value

    def energy(self):
        e = capi.Energies()
        self._chk(self.L.orc_energy(self.h, C.byref(e)))
        return e.as_dict()

    def norm(self):
        v = C.c_double()
        self._chk(self.L.orc_norm(self.h, C.byref(v)))
        return v.value

    def moments(self, npart):
        out = np.zeros(10)
        self._chk(self.L.orc_moments(self.h, npart, _p(out)))
        return {"cm": out[0:3].copy(), "r2": out[3:6].copy(), "n": out[6], "L": out[7:10].copy()}

    def evolve(self, dt, nsteps):
        t = np.zeros((nsteps, capi.TRAJ_STRIDE))
        self._chk(self.L.orc_evolve(self.h, dt, nsteps, _p(t)))
        return t

    def field(self, which):
        if which in capi.COMPLEX_FIELDS:
            out = np.zeros(self.shape, dtype=np.complex128)
            self._chk(self.L.orc_get_field(self.h, which, _p(out.view(np.float64))))
        elif which in capi.HALF_COMPLEX_FIELDS:
            out = np.zeros((self.p.nz, self.p.ny, self.p.nx // 2 + 1), dtype=np.complex128)
            self._chk(self.L.orc_get_field(self.h, which, _p(out.view(np.float64))))
        elif which in capi.HALF_FIELDS:
            out = np.zeros((self.p.nz, self.p.ny, self.p.nx // 2 + 1))
            self._chk(self.L.orc_get_field(self.h, which, _p(out)))
        else:
            out = np.zeros(self.shape)
            self._chk(self.L.orc_get_field(self.h, which, _p(out)))
        return out

    def imp(self):
        r, v, a = np.zeros(3), np.zeros(3), np.zeros(3)
        self._chk(self.L.orc_get_imp(self.h, _p(r), _p(v), _p(a)))
        return r, v, a


def fft_r2c(f):
    nz, ny, nx = f.shape
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
    lib().orc_fft_r2c(nx, ny, nz, _p(f), _p(out.view(np.float64)))
    return out


def fft_c2r(F, nx):
    nz, ny, nxh = F.shape
    F = np.ascontiguousarray(F, dtype=np.complex128)
    out = np.zeros((nz, ny, nx))
    lib().orc_fft_c2r(nx, ny, nz, _p(F.view(np.float64)), _p(out))
    return out


def deriv(f, npd, icon, id_, iv, h):
    nz, ny, nx = f.shape
    cplx = np.iscomplexobj(f)
    f = np.ascontiguousarray(f, dtype=np.complex128 if cplx else np.float64)
    out = np.zeros_like(f)
    fv = f.view(np.float64) if cplx else f
    ov = out.view(np.float64) if cplx else out
    rc = lib().orc_deriv(nx, ny, nz, npd, icon, id_, iv, h, int(cplx), _p(fv), _p(ov))
    if rc:
        raise ValueError("orc_deriv rc=%d" % rc)
    return out
