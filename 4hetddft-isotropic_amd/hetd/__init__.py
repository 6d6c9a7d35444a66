"""Python binding of the MI355X real-time propagation step (lib/libhetd.so).

``Context`` wraps the C ABI of include/hetd.h one call per reference
routine (poten, potenimp, forceimp, steprk, steppc, energy); the fields live
in device memory between calls.  The Fortran driver (fortran/) is the
reference-compatible host; this module serves tests and benchmarks.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import default_params  # noqa: F401

__all__ = ["Context", "default_params", "capi", "build_potion"]

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def _v3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def build_potion(selec, r_cutoff, umax, nmax, xmaxbox):
    """Impurity table of potenimpini (potenimpini.f90:11-25) via the library."""
    L = capi.load()
    npot, d = C.c_int32(), C.c_double()
    s = selec.encode()
    rc = L.hetd_build_potion(s, r_cutoff, umax, nmax, xmaxbox, None, C.byref(npot), C.byref(d))
    if rc:
        raise ValueError("unknown impurity potential %r" % selec)
    tab = np.zeros(npot.value)
    rc = L.hetd_build_potion(s, r_cutoff, umax, nmax, xmaxbox, _p(tab), C.byref(npot), C.byref(d))
    if rc:
        raise ValueError("unknown impurity potential %r" % selec)
    return tab, d.value


class Context:
    """Device context for one grid (hetd_init ... hetd_finalize)."""

    def __init__(self, params, selec="null", r_cutoff=2.0, umax=0.0, potion=None):
        self.L = capi.load()
        self.p = params
        if potion is None:  # an unknown selec raises before the device is touched
            potion = build_potion(selec, r_cutoff, umax, max(params.nx, params.ny, params.nz),
                                  max(params.xmax, params.ymax, params.zmax))
        h = C.c_void_p()
        rc = self.L.hetd_init(C.byref(params), C.byref(h))
        self.h = h
        if rc:
            msg = self.L.hetd_last_error(h).decode() if h.value else ""
            self.close()
            raise RuntimeError("hetd_init failed (%d): %s" % (rc, msg))
        self.shape = (params.nz, params.ny, params.nx)
        self.hshape = (params.nz, params.ny, params.nx // 2 + 1)
        tab, d = potion
        self.potion, self.delinter = tab, d
        self._chk(self.L.hetd_set_potion(self.h, _p(tab), len(tab), d))

    @staticmethod
    def _host(a, dtype, shape, what):
        # the C side reads exactly prod(shape) values: check before handing over
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.shape != tuple(shape):
            raise ValueError("%s: shape %s, expected %s (nz, ny, nx)" % (what, a.shape, tuple(shape)))
        return a

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("hetd error %d: %s" % (rc, self.L.hetd_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.hetd_finalize(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- reference call surface -------------------------------------------
    def set_state(self, psi, r, v):
        a = self._host(psi, np.complex128, self.shape, "psi").view(np.float64)
        self._chk(self.L.hetd_set_state(self.h, _p(a), _p(_v3(r)), _p(_v3(v))))

    def init(self):
        """The driver's first potenimp/poten/forceimp (DFT4He3d.f90:641-644)."""
        r, v, a = self.imp()
        self.potenimp(r)
        self.poten()
        F = self.forceimp(r)
        self.set_aimp(F / (self.p.mimpur * 0.02061484))

    def potenimp(self, r):
        self._chk(self.L.hetd_potenimp(self.h, _p(_v3(r))))

    def poten(self):
        self._chk(self.L.hetd_poten(self.h))

    def forceimp(self, r):
        F = np.zeros(3)
        self._chk(self.L.hetd_forceimp(self.h, _p(_v3(r)), _p(F)))
        return F

    def set_aimp(self, a):
        self._chk(self.L.hetd_set_aimp(self.h, _p(_v3(a))))

    def steprk(self, dt):
        self._chk(self.L.hetd_steprk(self.h, dt))

    def steppc(self, dt):
        a, b = C.c_double(), C.c_double()
        self._chk(self.L.hetd_steppc(self.h, dt, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_uext(self, uext):
        """External potential of the energy (energy.f90:124); None removes it."""
        if uext is None:
            self._chk(self.L.hetd_set_uext(self.h, None))
            return
        a = self._host(uext, np.float64, self.shape, "uext")
        self._chk(self.L.hetd_set_uext(self.h, _p(a)))

    def set_surface(self, lsurf3d, zsurf):
        """lsurf: the He-TiO2 potential as uext (DFT4He3d.f90:758-784)."""
        self._chk(self.L.hetd_set_surface(self.h, int(lsurf3d), float(zsurf)))

    def energy(self):
        e = capi.Energies()
        self._chk(self.L.hetd_energy(self.h, C.byref(e)))
        return e.as_dict()

    def norm(self):
        v = C.c_double()
        self._chk(self.L.hetd_norm(self.h, C.byref(v)))
        return v.value

    def moments(self, npart):
        """dict: cm (3), r2 (3), n, L (3) -- the energy block's diagnostics."""
        out = np.zeros(10)
        self._chk(self.L.hetd_moments(self.h, npart, _p(out)))
        return {"cm": out[0:3].copy(), "r2": out[3:6].copy(), "n": out[6], "L": out[7:10].copy()}

    def save_history(self):
        """The integrator history as bytes (hetd_save_history)."""
        n = self.L.hetd_history_size(self.h)
        buf = C.create_string_buffer(n)
        self._chk(self.L.hetd_save_history(self.h, buf))
        return buf.raw

    def load_history(self, blob):
        """Restore a history saved by save_history (after set_state)."""
        buf = C.create_string_buffer(blob, len(blob))
        self._chk(self.L.hetd_load_history(self.h, buf, len(blob)))

    def evolve(self, dt, nsteps, traj=True):
        t = np.zeros((max(nsteps, 1), capi.TRAJ_STRIDE))
        self._chk(self.L.hetd_evolve(self.h, dt, nsteps, _p(t) if traj else None))
        return t[:nsteps]

    def autotune(self, dt, steps):
        """Pick the fastest of the kernel variants that give the same bits
        (hetd_autotune); the state is unchanged.  Returns the choice bits."""
        ch = C.c_int32()
        self._chk(self.L.hetd_autotune(self.h, dt, steps, C.byref(ch)))
        return ch.value

    def field(self, which):
        if which in capi.COMPLEX_FIELDS:
            out = np.zeros(self.shape, dtype=np.complex128)
            self._chk(self.L.hetd_get_field(self.h, which, _p(out.view(np.float64))))
        elif which in capi.HALF_COMPLEX_FIELDS:
            out = np.zeros(self.hshape, dtype=np.complex128)
            self._chk(self.L.hetd_get_field(self.h, which, _p(out.view(np.float64))))
        elif which in capi.HALF_FIELDS:
            out = np.zeros(self.hshape)
            self._chk(self.L.hetd_get_field(self.h, which, _p(out)))
        else:
            out = np.zeros(self.shape)
            self._chk(self.L.hetd_get_field(self.h, which, _p(out)))
        return out

    def imp(self):
        r, v, a = np.zeros(3), np.zeros(3), np.zeros(3)
        self._chk(self.L.hetd_get_imp(self.h, _p(r), _p(v), _p(a)))
        return r, v, a

    def fft_r2c(self, f):
        f = self._host(f, np.float64, self.shape, "f")
        out = np.zeros(self.hshape, dtype=np.complex128)
        self._chk(self.L.hetd_fft_r2c(self.h, _p(f), _p(out.view(np.float64))))
        return out

    def fft_c2r(self, F):
        F = self._host(F, np.complex128, self.hshape, "F")
        out = np.zeros(self.shape)
        self._chk(self.L.hetd_fft_c2r(self.h, _p(F.view(np.float64)), _p(out)))
        return out

    def timing(self, kclass):
        """Start timing kernel class kclass (capi.K_*) with HIP events."""
        self._chk(self.L.hetd_timing(self.h, kclass))

    def timing_read(self):
        """(total ms, launches, algorithmic bytes) since timing() was called."""
        t, n, b = C.c_double(), C.c_int64(), C.c_double()
        self._chk(self.L.hetd_timing_read(self.h, C.byref(t), C.byref(n), C.byref(b)))
        return t.value, n.value, b.value

    def stream(self):
        return self.L.hetd_stream(self.h)

    def deriv(self, f, id_, iv):
        cplx = np.iscomplexobj(f)
        f = self._host(f, np.complex128 if cplx else np.float64, self.shape, "f")
        out = np.zeros_like(f)
        fv = f.view(np.float64) if cplx else f
        ov = out.view(np.float64) if cplx else out
        self._chk(self.L.hetd_deriv(self.h, id_, iv, int(cplx), _p(fv), _p(ov)))
        return out


def vsurf(lsurf3d, x, y, z):
    """He-TiO2 surface potential (morse.f90) at (x, y, z) above the surface, K."""
    v = C.c_double()
    if capi.load().hetd_vsurf(int(lsurf3d), float(x), float(y), float(z), C.byref(v)):
        raise ValueError("hetd_vsurf")
    return v.value


def pool_evolve(contexts, dt, nsteps):
    """Advance every Context in `contexts` by nsteps iterations (hetd_pool_evolve):
    the same states as calling ctx.evolve(dt, nsteps, traj=False) on each, with
    the members sharing the GPU."""
    if not contexts:
        return
    L = contexts[0].L
    arr = (C.c_void_p * len(contexts))(*[c.h.value for c in contexts])
    rc = L.hetd_pool_evolve(arr, len(contexts), dt, nsteps)
    if rc != 0:
        msgs = [c.L.hetd_last_error(c.h).decode() for c in contexts]
        raise RuntimeError("hetd_pool_evolve error %d: %s" % (rc, "; ".join(m for m in msgs if m)))
