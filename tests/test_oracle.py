"""Pins the CPU oracle (oracle/oracle.c) -- CPU only.

Each check is chosen so that a plausible mistake fails it: golden vectors
written by the reference itself (tests/golden), the reference's own
finite-difference table, brute-force / independent FFTs, closed forms and
known-answer values.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import golden_util as G
from hetd import capi
from hetd.initstate import woods_saxon_state

TOL = 1e-10  # relative, the FP64 parity bar


# ---------------------------------------------------------------- golden


def test_oracle_source_compiles(tmp_path):
    """The oracle's C source as it stands compiles (the tests load a prebuilt
    liboracle.so, which could hide a broken source)."""
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    r = subprocess.run(["gcc", "-O0", "-fPIC", "-std=c11", "-Wall", "-Wextra", "-Wno-unused-parameter",
                        "-shared", "-o", str(tmp_path / "liboracle_check.so"), os.path.join(here, "oracle.c"),
                        "-lm"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "warning" not in r.stderr, r.stderr

@pytest.mark.parametrize("name", G.NAMES)
def test_oracle_matches_reference_golden(oracle_mod, name):
    cfg, d = G.load(name)
    p = G.params_for(cfg, capi)
    ctx = oracle_mod.OracleCtx(p, selec=cfg["selec"], r_cutoff=cfg["r_cutoff"], umax=cfg["umax"])
    ctx.set_state(d["psi0"], d["rimp0"], d["vimp0"])
    ctx.init()
    for key, fid in (("pot4", capi.F_POT4), ("dencg", capi.F_DENCG), ("delj4", capi.F_DELJ4),
                     ("ccor", capi.F_CCOR), ("dxden", capi.F_DXDEN), ("dyden", capi.F_DYDEN),
                     ("dzden", capi.F_DZDEN), ("uimp", capi.F_UIMP), ("penalty", capi.F_PENALTY),
                     ("ualphas", capi.F_UALPHAS), ("denalf", capi.F_DENALF)):
        if "init_%s_planes" % key not in d:
            continue
        if key == "ccor" and cfg["core"] == "OTC":
            continue  # term_alfa reuses sto1 (term_alfa.f90:83-86): not the correlation term
        assert G.rel(G.planes(ctx.field(fid)), d["init_%s_planes" % key]) < TOL, key
    assert G.rel(G.planes(ctx.field(capi.F_WCGK)), d["init_wcgk_planes"]) < 1e-14
    assert G.rel(G.planes(ctx.field(capi.F_FDEN)), d["init_fden_planes"]) < 1e-12  # FFTW sign/scale
    # the reference's LJ form factor loses ~6 digits to cancellation at the
    # largest |p| (fforma.f90:94-101); the oracle evaluates it in long double
    assert G.rel(G.planes(ctx.field(capi.F_FVLJ)), d["init_fvlj_planes"]) < 1e-9
    e = ctx.energy()
    ref = d["init_energy"]
    for i, k in enumerate(("etot4", "ekin4", "elj4", "ecor4", "esolid", "ealphas", "ekinx", "eHeX", "etot")):
        assert abs(e[k] - ref[i]) <= TOL * max(abs(ref[i]), 1.0), k
    r, v, a = ctx.imp()
    assert np.abs(a - d["init_aimp"]).max() <= TOL * np.abs(d["init_aimp"]).max()
    # trajectory: RK x3 then PC (DFT4He3d.f90:840-852)
    traj = ctx.evolve(G.DT, cfg["niter"])
    st = d["steps"]
    scale = np.maximum(np.abs(st[:, 1:12]).max(axis=0), 1e-30)
    err = (np.abs(traj - st[:, 1:12]).max(axis=0) / scale)
    # errHe / errimp measure predictor-corrector differences, i.e. differences
    # of nearly equal states: compared to 1e-8 relative plus an absolute floor
    # of 1e-15 of the state itself
    psimax = np.abs(d["psi0"]).max()
    assert np.all(np.abs(traj[:, 0] - st[:, 1]) <= 1e-8 * np.abs(st[:, 1]) + 1e-15 * psimax), err
    assert np.all(np.abs(traj[:, 1] - st[:, 2]) <= 1e-6 * np.abs(st[:, 2]) + 1e-15), err
    assert err[2:].max() < TOL, err
    assert G.rel(G.planes(ctx.field(capi.F_PSI)), d["final_planes"]) < TOL
    en = d["energy"][-1]
    e = ctx.energy()
    assert abs(ctx.norm() - en[1]) < TOL * en[1]
    assert abs(e["etot4"] - en[2]) < TOL * abs(en[2])
    assert abs(e["etot"] - en[10]) < TOL * abs(en[10])


# ------------------------------------------------------ finite differences

def test_fd_weights_equal_reference_table(oracle_mod):
    with open(os.path.join(G.GOLDEN, "fd_coeffs.json")) as f:
        tab = json.load(f)
    for npd in (3, 5, 7, 9, 11, 13, 15, 25):
        for id_ in (1, 2):
            n = npd + 4
            f = np.zeros((1, 1, n))
            pos = n // 2
            f[0, 0, pos] = 1.0
            df = oracle_mod.deriv(f, npd, 8, id_, 1, 1.0)[0, 0]
            m = npd // 2
            got = [df[pos + m - j] for j in range(npd)]
            want = tab["%d_%d" % (npd, id_)]
            assert np.allclose(got, want, rtol=2e-16, atol=1e-300), (npd, id_)


def test_fd_weights_closed_form_m6():
    # d_k = (-1)^{k+1} (6!)^2/(k (6-k)!(6+k)!), c_k = 2 d_k/k (first and
    # second derivative, 13 points)
    with open(os.path.join(G.GOLDEN, "fd_coeffs.json")) as f:
        tab = json.load(f)
    from math import factorial
    d = [Fraction((-1) ** (k + 1) * factorial(6) ** 2, k * factorial(6 - k) * factorial(6 + k)) for k in range(1, 7)]
    assert [float(x) for x in d] == [6 / 7, -15 / 56, 5 / 63, -1 / 56, 1 / 385, -1 / 5544]
    row1 = tab["13_1"]
    assert np.allclose(row1[7:], [float(x) for x in d], rtol=1e-15)
    row2 = tab["13_2"]
    assert np.allclose(row2[7:], [float(2 * x / (k + 1)) for k, x in enumerate(d)], rtol=1e-15)
    assert abs(row2[6] - (-5369 / 1800)) < 1e-15


@pytest.mark.parametrize("npd", [9, 13])
@pytest.mark.parametrize("iv", [1, 2, 3])
def test_fd_exact_on_polynomials(oracle_mod, npd, iv):
    n = [20, 17, 18]
    h = 0.37
    m = npd // 2
    coords = [np.arange(k) * h for k in n]
    z, y, x = np.meshgrid(coords[2], coords[1], coords[0], indexing="ij")
    t = (x, y, z)[iv - 1]
    deg = npd - 1
    f = (t / 3.0) ** deg + 0.5 * t ** 2 - t
    d1 = deg * (t / 3.0) ** (deg - 1) / 3.0 + t - 1.0
    d2 = deg * (deg - 1) * (t / 3.0) ** (deg - 2) / 9.0 + 1.0
    for id_, want in ((1, d1), (2, d2)):
        got = oracle_mod.deriv(f, npd, 13, id_, iv, h)
        sl = [slice(None)] * 3
        sl[3 - iv] = slice(m, n[iv - 1] - m)
        sl = tuple(sl)
        assert np.allclose(got[sl], want[sl], rtol=1e-9, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("icon", [13, 8, 18])
def test_fd_boundary_brute_force(oracle_mod, icon):
    rng = np.random.default_rng(3)
    f = rng.standard_normal((5, 4, 16)) + 1j * rng.standard_normal((5, 4, 16))
    npd, h = 9, 0.5
    with open(os.path.join(G.GOLDEN, "fd_coeffs.json")) as fh:
        c = np.array(json.load(fh)["9_2"]) / h ** 2
    want = np.zeros_like(f)
    n, m = 16, 4
    for i in range(n):
        for j in range(npd):
            q = i + j - m
            s = 1.0
            if q < 0 or q >= n:
                if icon == 13:
                    continue
                if icon == 18:
                    s = -1.0
                q %= n
            want[:, :, i] += c[j] * s * f[:, :, q]
    got = oracle_mod.deriv(f, npd, icon, 2, 1, h)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("shape", [(6, 8, 10), (9, 15, 16), (5, 4, 9), (27, 10, 12)])
def test_fft_against_numpy(oracle_mod, shape):
    rng = np.random.default_rng(1)
    f = rng.standard_normal(shape)
    F = oracle_mod.fft_r2c(f)
    assert np.allclose(F, np.fft.rfftn(f), rtol=1e-13, atol=1e-12)
    g = oracle_mod.fft_c2r(F, shape[2]) / f.size
    assert np.allclose(g, f, rtol=1e-13, atol=1e-13)
    # c2r of a non-Hermitian half spectrum: canonical semantics (imaginary
    # parts of k=0 / Nyquist ignored along x), the same as numpy's irfftn
    G_ = F + 0.3 * (rng.standard_normal(F.shape) + 1j * rng.standard_normal(F.shape))
    assert np.allclose(oracle_mod.fft_c2r(G_, shape[2]), np.fft.irfftn(G_, s=shape, axes=(0, 1, 2)) * f.size,
                       rtol=1e-12, atol=1e-11)


def test_fft_brute_force_dft(oracle_mod):
    rng = np.random.default_rng(2)
    shape = (3, 5, 6)
    f = rng.standard_normal(shape)
    nz, ny, nx = shape
    F = np.zeros((nz, ny, nx // 2 + 1), dtype=complex)
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx // 2 + 1):
                s = 0
                for z in range(nz):
                    for y in range(ny):
                        for x in range(nx):
                            s += f[z, y, x] * np.exp(-2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
                F[kz, ky, kx] = s
    assert np.allclose(oracle_mod.fft_r2c(f), F, rtol=1e-12, atol=1e-12)


# -------------------------------------------------------- special functions

def test_sine_integral_against_scipy(oracle_mod):
    from scipy.special import sici
    L = oracle_mod.lib()
    for x in np.concatenate([np.linspace(0.01, 40, 300), [4.0, 4.0000001, 100.0, 1e3]]):
        assert abs(L.orc_si(x) - sici(x)[0]) < 4e-16 * max(1.0, abs(sici(x)[0])), x


def test_lj_form_factor_quadrature_and_b4(oracle_mod):
    from scipy.integrate import quad
    L = oracle_mod.lib()
    eps, sig, h = 10.22, 2.556, 2.190323
    v0 = L.orc_vlj_hat(0.0, eps, sig, h, 0.0)
    assert abs(v0 - (-718.99)) < 0.01  # b4 (modules.f90:212)
    for p in (0.05, 0.2, 0.5, 1.0):
        k = 2 * np.pi * p

        def integrand(r):
            return 4 * np.pi * r * r * 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6) * np.sin(k * r) / (k * r)

        val = quad(integrand, h, 60.0, limit=4000, epsabs=1e-13, epsrel=1e-13)[0]
        tail = quad(integrand, 60.0, 400.0, limit=8000, epsabs=1e-14)[0]
        assert abs(L.orc_vlj_hat(p, eps, sig, h, 0.0) - (val + tail)) < 1e-8 * 720, p


def test_xe_he_known_answers(oracle_mod):
    with open(os.path.join(G.GOLDEN, "kat.json")) as f:
        kat = json.load(f)
    L = oracle_mod.lib()
    # the quoted values carry 15 and 11 significant digits: agree to within
    # half a unit of the last quoted digit
    for (r, v), half_ulp in zip(kat["V_Xe_He"], (0.5e-10, 0.5e-7)):
        assert abs(L.orc_vimp(1, r, 1.0, 0.0) - v) <= half_ulp
    assert L.orc_vimp(1, 2.0, 2.55, 2793.9365463) == 2793.9365463
    assert L.orc_vimp(0, 3.0, 2.55, 100.0) == 0.0


# ------------------------------------------------------ closed-form states

def _uniform_ctx(oracle_mod, rho0=0.0218, lsolid=1, core="OT"):
    p = capi.default_params(nx=16, ny=15, nz=18, xmax=8.0, ymax=7.5, zmax=9.0, npd=13, icon=8,
                            lsolid=lsolid, core=core)
    ctx = oracle_mod.OracleCtx(p)
    psi = np.full((18, 15, 16), np.sqrt(rho0), dtype=complex)
    ctx.set_state(psi, (0, 0, 0), (0, 0, 0))
    return p, ctx


def test_poten_uniform_density_closed_form(oracle_mod):
    rho = 0.0218
    p, ctx = _uniform_ctx(oracle_mod, rho)
    ctx.init()
    L = oracle_mod.lib()
    b = L.orc_vlj_hat(0.0, p.eps4, p.sigma4, p.h4, 0.0)
    t = np.tanh(p.beta_solid * (rho - p.den_m))
    pen = p.c_solid * (1 + t + p.beta_solid * rho * (1 - t * t))
    U = b * rho + rho ** 2 * (p.cp4 / 2 + p.cpp4 * rho / 3) + rho ** 2 * (p.cp4 + p.cpp4 * rho) + pen
    assert np.allclose(ctx.field(capi.F_POT4), U, rtol=1e-12)
    assert np.allclose(ctx.field(capi.F_DENCG), rho, rtol=1e-13)
    e = ctx.energy()
    vol = 16 * 15 * 18 * 1.0
    ev = b * rho ** 2 / 2 + rho ** 3 * (p.cp4 / 2 + p.cpp4 * rho / 3) + p.c_solid * rho * (1 + t)
    assert abs(e["etot4"] - ev * vol) < 1e-10 * abs(ev * vol)
    assert abs(e["ekin4"]) < 1e-12


def test_force_vanishes_for_symmetric_density(oracle_mod):
    p = capi.default_params(nx=24, ny=24, nz=24, xmax=12.0, ymax=12.0, zmax=12.0, npd=13, icon=13)
    ctx = oracle_mod.OracleCtx(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463)
    psi = woods_saxon_state(24, 24, 24, 12.0, 12.0, 12.0, 60, 1, phase_amp=0.0, center=(-0.5, -0.5, -0.5))
    ctx.set_state(psi, (-0.5, -0.5, -0.5), (0, 0, 0))  # grid is symmetric about -h/2
    ctx.init()
    F = ctx.forceimp((-0.5, -0.5, -0.5))
    assert np.abs(F).max() < 1e-9
    ctx.potenimp((1.3, -0.5, -0.5))
    ctx.poten()
    F = ctx.forceimp((1.3, -0.5, -0.5))
    assert abs(F[0]) > 1.0  # and the probe is sensitive


def test_moments_closed_forms(oracle_mod):
    """Centre of mass, <x^2> and <L> of the energy block (DFT4He3d.f90:868-947)
    against closed forms: a Gaussian on a grid symmetric about its centre,
    and an m = 1 vortex, whose <Lz> is the particle number (Lx = Ly = 0)."""
    n, h, s = 48, 0.3, 1.2
    xm = n * h / 2
    p = capi.default_params(nx=n, ny=n, nz=n, xmax=xm, ymax=xm, zmax=xm, npd=13, icon=13)
    x = -xm + h * np.arange(n)
    c = -h / 2  # the grid is symmetric about -h/2
    zz, yy, xx = np.meshgrid(x - c, x - c, x - c, indexing="ij")
    g = np.exp(-(xx ** 2 + yy ** 2 + zz ** 2) / (2 * s * s))
    ctx = oracle_mod.OracleCtx(p)
    # Gaussian: cm = c, sum(den x^2) dV = pi^(3/2) s^3 (s^2/2 + c^2)
    ctx.set_state(g + 0j, (0, 0, 0), (0, 0, 0))
    m = ctx.moments(1.0)
    nrm = np.pi ** 1.5 * s ** 3
    assert abs(m["n"] - nrm) < 1e-12 * nrm
    m = ctx.moments(m["n"])
    assert np.abs(m["cm"] - c).max() < 1e-13
    assert np.abs(m["r2"] - nrm * (s * s / 2 + c * c)).max() < 1e-11 * nrm
    assert np.abs(m["L"]).max() < 1e-12 * nrm  # real wavefunction: no current
    # vortex (x + i y) g about the centre: Lz = N, Lx = Ly = 0 (FD error ~ h^12)
    ctx.set_state((xx + 1j * yy) * g, (0, 0, 0), (0, 0, 0))
    N = ctx.moments(1.0)["n"]
    m = ctx.moments(N)
    assert abs(m["L"][2] - N) < 1e-7 * N
    assert np.abs(m["L"][:2]).max() < 1e-12 * N
    assert np.abs(m["cm"] - c).max() < 1e-13
    # the opposite circulation flips the sign
    ctx.set_state((xx - 1j * yy) * g, (0, 0, 0), (0, 0, 0))
    assert abs(ctx.moments(N)["L"][2] + N) < 1e-7 * N
    ctx.close()


def _scalar_rk_pc(fun, dt, nsteps, y0):
    """Scalar recurrences of RK-Gill (steprk.f90:32-92) for the first three
    steps and Hamming predictor-modifier-corrector (steppc.f90:45-134) after,
    for y' = fun(y)."""
    s2 = np.sqrt(2.0)
    a = [0.5, 1 - 1 / s2, 1 + 1 / s2, 1 / 6]
    b = [2.0, 1.0, 1.0, 2.0]
    c = [0.5, 1 - 1 / s2, 1 + 1 / s2, 0.5]
    y, q = complex(y0), 0j
    ys, fs = [y], []
    pc = 0j
    for n in range(nsteps):
        if n < 3:
            for j in range(4):
                f = fun(y)
                if j == 0:
                    fs.append(f)
                r = a[j] * (f - b[j] * q)
                q = q + 3 * r - c[j] * f
                y = y + dt * r
            ys.append(y)
        else:
            f = fun(y)
            fs.append(f)
            P = ys[-4] + 4 / 3 * dt * (2 * f - fs[-2] + 2 * fs[-3])
            M = P - 112 / 121 * pc
            pc = P
            C = 0.125 * (9 * ys[-1] - ys[-3] + 3 * dt * (fun(M) + 2 * f - fs[-2]))
            pc = pc - C
            y = C + 9 / 121 * pc
            ys.append(y)
    return ys


def test_integrators_reduce_to_scalar_recurrence(oracle_mod):
    # a spatially uniform state in a periodic box: the Laplacian vanishes and
    # U depends on |psi|^2 only, U(rho) = b rho + rho^2 (cp4/2 + cpp4 rho/3)
    # + rho^2 (cp4 + cpp4 rho), so the 3-D integrators must reproduce the
    # scalar recurrence for y' = -i U(|y|^2) y (timec = i)
    rho = 0.0218
    p, ctx = _uniform_ctx(oracle_mod, rho, lsolid=0)
    ctx.init()
    b = oracle_mod.lib().orc_vlj_hat(0.0, p.eps4, p.sigma4, p.h4, 0.0)

    def fun(y):
        r = abs(y) ** 2
        U = b * r + r * r * (p.cp4 / 2 + p.cpp4 * r / 3) + r * r * (p.cp4 + p.cpp4 * r)
        return -1j * U * y

    dt = 0.02
    nsteps = 9
    traj = ctx.evolve(dt, nsteps)
    psi = ctx.field(capi.F_PSI)
    ys = _scalar_rk_pc(fun, dt, nsteps, np.sqrt(rho))
    assert abs(psi[3, 4, 5] - ys[-1]) < 1e-12 * abs(ys[-1])
    assert np.abs(psi - psi[0, 0, 0]).max() < 1e-14
    assert np.all(traj[3:, 0] > 0)


def test_integrators_fourth_order(oracle_mod):
    errs = []
    for dt in (0.04, 0.02):
        ys = _scalar_rk_pc(lambda y: -3j * y, dt, int(round(1.2 / dt)), 1.0)
        errs.append(abs(ys[-1] - np.exp(-1j * 3.0 * 1.2)))
    order = np.log2(errs[0] / errs[1])
    assert 3.8 < order < 5.5, order  # a dropped or mis-signed term falls to order <= 2
