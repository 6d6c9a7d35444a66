"""Parity of the HIP path (lib/libhetd.so on gfx950) with the oracle and with
the reference's own golden vectors.  FP64 bar: 1e-10 relative."""
import sys

import numpy as np
import pytest

import golden_util as G
from hetd import Context, capi

pytestmark = pytest.mark.gpu
TOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_pair(oracle_mod, p, selec="null", r_cutoff=2.0, umax=0.0):
    ctx = Context(p, selec=selec, r_cutoff=r_cutoff, umax=umax)
    orc = oracle_mod.OracleCtx(p, selec=selec, r_cutoff=r_cutoff, umax=umax)
    return ctx, orc


# --------------------------------------------------------------------- FFT

@pytest.mark.parametrize("shape", [(8, 6, 10), (27, 10, 12), (16, 16, 16), (5, 9, 45), (20, 12, 36),
                                   (32, 24, 64), (40, 30, 50), (16, 64, 256), (256, 16, 64), (64, 256, 16),
                                   (15, 9, 256), (384, 6, 10), (6, 384, 12), (8, 6, 384), (162, 20, 216),
                                   (12, 288, 320), (192, 180, 150),
                                   # LDS Stockham passes on long lines: narrower column
                                   # blocks where 8 columns would not fit (750, 810, 960)
                                   (4, 750, 6), (810, 4, 6), (6, 4, 972), (960, 5, 4),
                                   # x lines with a radix-2 stage: fewer pairs per block
                                   (6, 5, 250), (4, 3, 810), (3, 4, 54)])
def test_fft_matches_numpy_and_oracle(oracle_mod, shape):
    say("fft", shape)
    nz, ny, nx = shape
    p = capi.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * 0.2, ymax=ny * 0.2, zmax=nz * 0.2, npd=3)
    with Context(p) as ctx:
        rng = np.random.default_rng(7)
        f = rng.standard_normal(shape)
        F = ctx.fft_r2c(f)
        ref = np.fft.rfftn(f)
        assert np.abs(F - ref).max() < 1e-12 * np.abs(ref).max()
        assert np.abs(F - oracle_mod.fft_r2c(f)).max() < 1e-12 * np.abs(ref).max()
        g = ctx.fft_c2r(F)  # includes *renor (fft.f90:108)
        assert np.abs(g - f).max() < 1e-13 * np.abs(f).max()
        # non-Hermitian input: one-dimensional c2r semantics along x
        H = F + 0.3 * (rng.standard_normal(F.shape) + 1j * rng.standard_normal(F.shape))
        want = oracle_mod.fft_c2r(H, nx) / f.size
        assert np.abs(ctx.fft_c2r(H) - want).max() < 1e-12 * np.abs(want).max()


# ------------------------------------------------------------- stencils

@pytest.mark.parametrize("npd,icon", [(13, 13), (9, 8), (13, 18), (3, 13), (25, 8)])
def test_derivatives_match_oracle(oracle_mod, npd, icon):
    say("deriv", npd, icon)
    shape = (27, 25, 30)
    p = capi.default_params(nx=30, ny=25, nz=27, xmax=6.0, ymax=5.0, zmax=5.4, npd=npd, icon=icon)
    h = (0.4, 0.4, 0.4)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(shape)
    fc = f + 1j * rng.standard_normal(shape)
    with Context(p) as ctx:
        for iv in (1, 2, 3):
            for id_ in (1, 2):
                for arr in (f, fc):
                    got = ctx.deriv(arr, id_, iv)
                    want = oracle_mod.deriv(arr, npd, icon, id_, iv, h[iv - 1])
                    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), (iv, id_)


# ------------------------------------------------- golden configurations

@pytest.mark.parametrize("name", G.NAMES)
def test_poten_energy_force_match_golden(oracle_mod, name):
    say("golden init", name)
    cfg, d = G.load(name)
    p = G.params_for(cfg, capi)
    ctx, orc = make_pair(oracle_mod, p, cfg["selec"], cfg["r_cutoff"], cfg["umax"])
    ctx.set_state(d["psi0"], d["rimp0"], d["vimp0"])
    ctx.init()
    orc.set_state(d["psi0"], d["rimp0"], d["vimp0"])
    orc.init()
    for key, fid in (("pot4", capi.F_POT4), ("dencg", capi.F_DENCG), ("delj4", capi.F_DELJ4),
                     ("ccor", capi.F_CCOR), ("dxden", capi.F_DXDEN), ("dyden", capi.F_DYDEN),
                     ("dzden", capi.F_DZDEN), ("uimp", capi.F_UIMP), ("penalty", capi.F_PENALTY),
                     ("ualphas", capi.F_UALPHAS), ("denalf", capi.F_DENALF)):
        if "init_%s_planes" % key not in d:
            continue
        if key == "ccor" and cfg["core"] == "OTC":
            continue  # term_alfa reuses sto1 (term_alfa.f90:83-86)
        got = ctx.field(fid)
        if key == "penalty":
            # The reference forms 1 + tanh(x) with tanh near -1 (densities far
            # below den_m), which keeps ~4 digits of a ~1e-8 K penalty; the
            # library uses the cancellation-free exp form of its mean field.
            # Compared on the scale at which it enters pot4.
            scale = np.abs(d["init_pot4_planes"]).max()
            assert np.abs(G.planes(got) - d["init_penalty_planes"]).max() < TOL * scale, key
            assert np.abs(got - orc.field(fid)).max() < TOL * scale, key
            continue
        assert G.rel(G.planes(got), d["init_%s_planes" % key]) < TOL, key
        assert G.rel(got, orc.field(fid)) < TOL, key
    if cfg["core"] != "OTC":
        assert G.rel(G.planes(ctx.field(capi.F_FDEN)), d["init_fden_planes"]) < 1e-12
    assert G.rel(ctx.field(capi.F_FVLJ), orc.field(capi.F_FVLJ)) < 1e-15
    assert G.rel(ctx.field(capi.F_WCGK), orc.field(capi.F_WCGK)) < 1e-15
    e = ctx.energy()
    ref = d["init_energy"]
    for i, k in enumerate(("etot4", "ekin4", "elj4", "ecor4", "esolid", "ealphas", "ekinx", "eHeX", "etot")):
        assert abs(e[k] - ref[i]) <= TOL * max(abs(ref[i]), 1.0), k
    r, v, a = ctx.imp()
    assert np.abs(a - d["init_aimp"]).max() <= TOL * max(np.abs(d["init_aimp"]).max(), 1e-300)
    assert abs(ctx.norm() - orc.norm()) < 1e-13 * orc.norm()


@pytest.mark.parametrize("name", G.NAMES)
def test_mixed_fields_track_golden(name):
    """Every field copy of the mixed path (FP32 fields widened on the way out)
    against the reference's FP64 fields: FP32 rounding, 1e-5 relative."""
    say("golden init mixed", name)
    cfg, d = G.load(name)
    p = G.params_for(cfg, capi, precision=capi.HETD_MIXED)
    with Context(p, selec=cfg["selec"], r_cutoff=cfg["r_cutoff"], umax=cfg["umax"]) as ctx:
        ctx.set_state(d["psi0"], d["rimp0"], d["vimp0"])
        ctx.init()
        for key, fid in (("pot4", capi.F_POT4), ("dencg", capi.F_DENCG), ("delj4", capi.F_DELJ4),
                         ("ccor", capi.F_CCOR), ("ualphas", capi.F_UALPHAS), ("denalf", capi.F_DENALF)):
            if "init_%s_planes" % key not in d or (key == "ccor" and cfg["core"] == "OTC"):
                continue
            assert G.rel(G.planes(ctx.field(fid)), d["init_%s_planes" % key]) < 1e-5, key
        if cfg["core"] != "OTC":
            assert G.rel(G.planes(ctx.field(capi.F_FDEN)), d["init_fden_planes"]) < 1e-5
        e = ctx.energy()
        for i, k in enumerate(("etot4", "ekin4", "elj4", "ecor4", "esolid", "ealphas", "ekinx", "eHeX", "etot")):
            assert abs(e[k] - d["init_energy"][i]) <= 1e-5 * max(abs(d["init_energy"][i]), 1.0), k


@pytest.mark.parametrize("name", G.NAMES)
def test_evolve_matches_golden_trajectory(oracle_mod, name):
    say("golden evolve", name)
    cfg, d = G.load(name)
    p = G.params_for(cfg, capi)
    with Context(p, selec=cfg["selec"], r_cutoff=cfg["r_cutoff"], umax=cfg["umax"]) as ctx:
        ctx.set_state(d["psi0"], d["rimp0"], d["vimp0"])
        ctx.init()
        traj = ctx.evolve(G.DT, cfg["niter"])
        st = d["steps"]
        psimax = np.abs(d["psi0"]).max()
        assert np.all(np.abs(traj[:, 0] - st[:, 1]) <= 1e-8 * np.abs(st[:, 1]) + 1e-15 * psimax)
        assert np.all(np.abs(traj[:, 1] - st[:, 2]) <= 1e-6 * np.abs(st[:, 2]) + 1e-15)
        scale = np.maximum(np.abs(st[:, 3:12]).max(axis=0), 1e-30)
        assert (np.abs(traj[:, 2:] - st[:, 3:12]).max(axis=0) / scale).max() < TOL
        assert G.rel(G.planes(ctx.field(capi.F_PSI)), d["final_planes"]) < TOL
        en = d["energy"][-1]
        e = ctx.energy()
        assert abs(ctx.norm() - en[1]) < TOL * en[1]
        assert abs(e["etot4"] - en[2]) < TOL * abs(en[2])
        assert abs(e["etot"] - en[10]) < TOL * abs(en[10])


def test_fine_grained_calls_match_oracle(oracle_mod):
    """steprk x3 / steppc through the per-routine entry points, the driver's
    potenimp/poten/forceimp/aimp sequence done by the caller."""
    say("fine-grained")
    cfg, d = G.load("ot_xe_32")
    p = G.params_for(cfg, capi)
    ctx, orc = make_pair(oracle_mod, p, cfg["selec"], cfg["r_cutoff"], cfg["umax"])
    for c in (ctx, orc):
        c.set_state(d["psi0"], d["rimp0"], d["vimp0"])
        c.init()
    m = p.mimpur * 0.02061484
    for it in range(6):
        if it < 3:
            ctx.steprk(G.DT)
            orc.steprk(G.DT)
            e1 = e2 = (0.0, 0.0)
        else:
            e1 = ctx.steppc(G.DT)
            e2 = orc.steppc(G.DT)
        for c in (ctx, orc):
            r, v, a = c.imp()
            c.potenimp(r)
            c.poten()
            F = c.forceimp(r)
            c.set_aimp(F / m)
        assert abs(e1[0] - e2[0]) <= 1e-8 * abs(e2[0]) + 1e-16
        assert G.rel(ctx.field(capi.F_PSI), orc.field(capi.F_PSI)) < TOL
        assert G.rel(ctx.field(capi.F_POT4), orc.field(capi.F_POT4)) < TOL
        assert G.rel(ctx.field(capi.F_PC), orc.field(capi.F_PC)) < 1e-8
        for x, y in zip(ctx.imp(), orc.imp()):
            assert np.abs(x - y).max() <= TOL * max(np.abs(y).max(), 1e-300)


@pytest.mark.parametrize("shape,core", [((720, 4, 6), "OT"), ((750, 6, 4), "OTC"), ((4, 864, 6), "OT")])
def test_poten_long_axes_match_oracle(oracle_mod, shape, core):
    """The mean field on grids whose long axis takes the LDS Stockham passes
    with narrowed column blocks: the z convolution above 640 points in FP64,
    and the lengths that need them for the butterfly count (750, 864)."""
    say("poten long axis", shape, core)
    nz, ny, nx = shape
    h = 0.5
    p = capi.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=9,
                            icon=8, core=core, lsolid=0 if core == "OTC" else 1)
    rng = np.random.default_rng(4)
    z = np.arange(nz)[:, None, None]
    rho = 0.02 + 0.004 * np.cos(2 * np.pi * 3 * z / nz) + 0.001 * rng.random(shape)
    psi = np.sqrt(rho) * np.exp(1j * rng.uniform(-0.1, 0.1, shape))
    ctx, orc = make_pair(oracle_mod, p, "Xe_He", 2.55, 2793.9365463)
    for c in (ctx, orc):
        c.set_state(psi, (0.2, -0.1, 3.0), (0.0, 0.0, 1.0))
        c.init()
    for f in (capi.F_POT4, capi.F_DENCG, capi.F_DELJ4):
        assert G.rel(ctx.field(f), orc.field(f)) < TOL, f
    e1, e2 = ctx.energy(), orc.energy()
    for k in e2:
        assert abs(e1[k] - e2[k]) <= TOL * max(abs(e2[k]), 1.0), k
    t1 = ctx.evolve(G.DT, 4)
    t2 = orc.evolve(G.DT, 4)
    assert np.abs(t1[:, 2:] - t2[:, 2:]).max() <= TOL * np.abs(t2[:, 2:]).max()
    assert G.rel(ctx.field(capi.F_PSI), orc.field(capi.F_PSI)) < TOL


@pytest.mark.parametrize("shape", [(750, 6, 4), (4, 960, 6)])
def test_mixed_poten_long_axes_tracks_fp64(shape):
    """The FP32 Stockham passes on the same long lines: the mixed mean field
    against the FP64 one (FP32 rounding, 1e-5 relative)."""
    say("mixed poten long axis", shape)
    nz, ny, nx = shape
    h = 0.5
    kw = dict(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=9, icon=8)
    rng = np.random.default_rng(5)
    rho = 0.02 + 0.002 * rng.random(shape)
    psi = np.sqrt(rho) * np.exp(1j * rng.uniform(-0.1, 0.1, shape))
    pot = {}
    for prec in (capi.HETD_FP64, capi.HETD_MIXED):
        with Context(capi.default_params(precision=prec, **kw), selec="Xe_He", r_cutoff=2.55,
                     umax=2793.9365463) as ctx:
            ctx.set_state(psi, (0.0, 0.0, 1.0), (0.0, 0.0, 0.0))
            ctx.init()
            pot[prec] = ctx.field(capi.F_POT4)
    assert G.rel(pot[capi.HETD_MIXED], pot[capi.HETD_FP64]) < 1e-5


@pytest.mark.parametrize("shape,npd,icon,core", [((48, 40, 36), 13, 13, "OT"), ((30, 32, 45), 9, 8, "OT"),
                                                 ((36, 36, 36), 13, 18, "OTC"), ((64, 64, 64), 13, 13, "OT"),
                                                 ((16, 64, 64), 9, 8, "OTC"),
                                                 # the generic stencil kernels (npd other than 9, 13)
                                                 ((24, 20, 30), 7, 13, "OT"), ((20, 24, 18), 11, 18, "OTC"),
                                                 ((18, 20, 24), 3, 8, "OT"), ((20, 18, 16), 25, 13, "OT"),
                                                 # an axis shorter than the stencil: the reference's
                                                 # derivative along it is zero (tiled kernels)
                                                 ((8, 40, 36), 13, 13, "OT"), ((36, 8, 40), 9, 8, "OT")])
def test_evolve_matches_oracle_other_grids(oracle_mod, shape, npd, icon, core):
    say("oracle evolve", shape, npd, icon, core)
    from hetd.initstate import woods_saxon_state
    nz, ny, nx = shape
    h = 0.6
    p = capi.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=npd,
                            icon=icon, core=core, lsolid=0 if core == "OTC" else 1, tzmean=nz * h / 2 - 2.0)
    psi = woods_saxon_state(nx, ny, nz, p.xmax, p.ymax, p.zmax, 150, 3)
    ctx, orc = make_pair(oracle_mod, p, "Xe_He", 2.55, 2793.9365463)
    for c in (ctx, orc):
        c.set_state(psi, (0.3, -0.2, 0.1), (1.0, 2.0, -3.0))
        c.init()
    t1 = ctx.evolve(G.DT, 7)
    t2 = orc.evolve(G.DT, 7)
    assert np.abs(t1[:, 2:] - t2[:, 2:]).max() <= TOL * np.abs(t2[:, 2:]).max()
    assert G.rel(ctx.field(capi.F_PSI), orc.field(capi.F_PSI)) < TOL
    e1, e2 = ctx.energy(), orc.energy()
    for k in e2:
        assert abs(e1[k] - e2[k]) <= TOL * max(abs(e2[k]), 1.0), k


@pytest.mark.parametrize("shape,icon", [((36, 32, 40), 8), ((40, 36, 32), 13)])
def test_mixed_precision_otc_tracks_fp64(shape, icon):
    """HETD_MIXED with the OTC core: the FP32 mean field plus the alpha_s
    term formed in FP64 from the FP64 density, against the FP64 path
    (relative energy error <= 1e-6, psi within 1e-5)."""
    say("mixed OTC", shape, icon)
    from hetd.initstate import woods_saxon_state
    nz, ny, nx = shape
    h = 0.6
    kw = dict(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=13, icon=icon,
              core="OTC", lsolid=0)
    psi = woods_saxon_state(nx, ny, nz, kw["xmax"], kw["ymax"], kw["zmax"], 150, 3)
    out = {}
    for prec in (capi.HETD_FP64, capi.HETD_MIXED):
        p = capi.default_params(precision=prec, **kw)
        with Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463) as ctx:
            ctx.set_state(psi, (0.3, -0.2, 0.1), (1.0, 2.0, -3.0))
            ctx.init()
            t = ctx.evolve(G.DT, 12)
            out[prec] = (t, ctx.field(capi.F_PSI), ctx.energy(), ctx.norm())
    t64, p64, e64, n64 = out[capi.HETD_FP64]
    t32, p32, e32, n32 = out[capi.HETD_MIXED]
    assert G.rel(p32, p64) < 1e-5
    for k in ("etot4", "etot", "ealphas"):
        assert abs(e32[k] - e64[k]) <= 1e-6 * abs(e64[k]), k
    assert abs(n32 - n64) <= 1e-9 * n64
    assert np.abs(t32[:, 2:5] - t64[:, 2:5]).max() < 1e-8 * max(np.abs(t64[:, 2:5]).max(), 1.0)


@pytest.mark.parametrize("selec,r_cutoff,umax", [("RB", 3.0, 1.0e3), ("Ba_plus_gs", 1.05, 6.836177e3),
                                                 ("Ag_gs", 2.0, 5.0e3)])
def test_evolve_other_impurity_potentials(oracle_mod, selec, r_cutoff, umax):
    """Other impurities of Select_Pot: the table (pinned bitwise against the
    reference's Select_Pot in test_impurity.py) is handed to both sides."""
    say("impurity", selec)
    from hetd import build_potion
    from hetd.initstate import woods_saxon_state
    nz, ny, nx = 36, 40, 48
    h = 0.6
    p = capi.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=13,
                            icon=13)
    tab = build_potion(selec, r_cutoff, umax, max(nx, ny, nz), max(p.xmax, p.ymax, p.zmax))
    ctx = Context(p, potion=tab)
    orc = oracle_mod.OracleCtx(p, potion=tab)
    psi = woods_saxon_state(nx, ny, nz, p.xmax, p.ymax, p.zmax, 150, 3)
    for c in (ctx, orc):
        c.set_state(psi, (0.5, -0.2, 3.0), (0.0, 1.0, -2.0))
        c.init()
    t1 = ctx.evolve(G.DT, 6)
    t2 = orc.evolve(G.DT, 6)
    assert np.abs(t1[:, 2:] - t2[:, 2:]).max() <= TOL * np.abs(t2[:, 2:]).max()
    assert G.rel(ctx.field(capi.F_PSI), orc.field(capi.F_PSI)) < TOL
    e1, e2 = ctx.energy(), orc.energy()
    for k in e2:
        assert abs(e1[k] - e2[k]) <= TOL * max(abs(e2[k]), 1.0), k
    ctx.close()


@pytest.mark.parametrize("lsurf3d,prec", [(0, capi.HETD_FP64), (1, capi.HETD_FP64), (1, capi.HETD_MIXED)])
def test_surface_potential_enters_the_energy_only(lsurf3d, prec):
    """lsurf (DFT4He3d.f90:758-784, energy.f90:124): etot4 and etot gain
    sum(uext*den)*dxyz with uext = min(3000, V_surf(x, y, z - Zsurf)); the
    state evolves exactly as without it."""
    say("surface", lsurf3d, prec)
    from hetd import vsurf
    from hetd.initstate import woods_saxon_state
    nz, ny, nx = 40, 36, 32
    h = 0.6
    p = capi.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=13,
                            icon=13, precision=prec)
    psi = woods_saxon_state(nx, ny, nz, p.xmax, p.ymax, p.zmax, 150, 3)
    a = Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463)
    b = Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463)
    zsurf = -9.0
    for c in (a, b):
        c.set_state(psi, (0.3, -0.2, 0.1), (1.0, 2.0, -3.0))
        c.init()
    b.set_surface(lsurf3d, zsurf)
    a.evolve(G.DT, 5, traj=False)
    b.evolve(G.DT, 5, traj=False)
    assert np.array_equal(a.field(capi.F_PSI), b.field(capi.F_PSI))
    x = -p.xmax + h * np.arange(nx)
    y = -p.ymax + h * np.arange(ny)
    z = -p.zmax + h * np.arange(nz)
    u = np.array([[[min(3000.0, vsurf(lsurf3d, xi, yj, zk - zsurf)) for xi in x] for yj in y] for zk in z])
    den = np.abs(a.field(capi.F_PSI)) ** 2
    extra = float(np.sum(u * den)) * h ** 3
    ea, eb = a.energy(), b.energy()
    assert abs((eb["etot4"] - ea["etot4"]) - extra) <= 1e-11 * (abs(extra) + abs(ea["etot4"]))
    assert abs((eb["etot"] - ea["etot"]) - extra) <= 1e-11 * (abs(extra) + abs(ea["etot"]))
    assert eb["ekin4"] == ea["ekin4"] and eb["eHeX"] == ea["eHeX"]
    b.set_uext(None)
    assert b.energy()["etot4"] == ea["etot4"]
    a.close()
    b.close()


@pytest.mark.parametrize("npd,icon,prec", [(13, 13, capi.HETD_FP64), (9, 8, capi.HETD_FP64), (7, 13, capi.HETD_FP64),
                                           (13, 13, capi.HETD_MIXED)])
def test_moments_match_oracle(oracle_mod, npd, icon, prec):
    """Centre of mass, <x2,y2,z2>, norm and <Lx,Ly,Lz> of the energy block
    (hetd_moments) against the oracle, on a droplet given some circulation."""
    say("moments", npd, icon, prec)
    from hetd.initstate import woods_saxon_state
    nz, ny, nx = 36, 40, 48
    h = 0.6
    kw = dict(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=npd, icon=icon)
    psi = woods_saxon_state(nx, ny, nz, kw["xmax"], kw["ymax"], kw["zmax"], 150, 3, center=(0.7, -0.4, 0.3))
    x = -kw["xmax"] + h * np.arange(nx)
    y = -kw["ymax"] + h * np.arange(ny)
    z = -kw["zmax"] + h * np.arange(nz)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    psi = psi * (1.0 + 0.2 * ((xx - 0.7) + 1j * (yy + 0.4)) + 0.1j * (zz - 0.3))
    orc = oracle_mod.OracleCtx(capi.default_params(**kw))
    orc.set_state(psi, (0, 0, 0), (0, 0, 0))
    want = orc.moments(150.0)
    with Context(capi.default_params(precision=prec, **kw)) as ctx:
        ctx.set_state(psi, (0, 0, 0), (0, 0, 0))
        got = ctx.moments(150.0)
    orc.close()
    for k in ("cm", "r2"):
        assert np.abs(got[k] - want[k]).max() <= TOL * np.abs(want[k]).max(), k
    assert abs(got["n"] - want["n"]) <= TOL * want["n"]
    assert np.abs(want["L"]).max() > 1.0  # the state does carry angular momentum
    assert np.abs(got["L"] - want["L"]).max() <= 1e-9 * np.abs(want["L"]).max()


def test_guard_reports_nonpositive_density():
    say("guard")
    p = capi.default_params(nx=16, ny=16, nz=16, xmax=4.0, ymax=4.0, zmax=4.0, npd=9)
    psi = np.full((16, 16, 16), 0.1 + 0j)
    psi[5, 5, 5] = np.nan
    with Context(p) as ctx:
        ctx.set_state(psi, (0, 0, 0), (0, 0, 0))
        ctx.init()
        with pytest.raises(RuntimeError, match="den not > 0"):
            ctx.steprk(G.DT)


def test_requests_the_context_cannot_serve_are_rejected():
    """Field copies that do not exist for the configuration, an unknown field,
    calls before the state, and a table of another size: each returns an
    error with a message and leaves the context usable."""
    say("rejections")
    p = capi.default_params(nx=16, ny=16, nz=16, xmax=4.0, ymax=4.0, zmax=4.0, npd=9, lsolid=0,
                            core=capi.HETD_CORE_OTC, icon=8)
    with Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463) as ctx:
        with pytest.raises(RuntimeError, match="no state"):
            ctx.steprk(G.DT)
        psi = np.full((16, 16, 16), 0.1 + 0j)
        ctx.set_state(psi, (0, 0, 0), (0, 0, 0))
        ctx.init()
        with pytest.raises(RuntimeError, match="no penalty without lsolid"):
            ctx.field(capi.F_PENALTY)
        with pytest.raises(RuntimeError, match="FDEN not kept with OTC"):
            ctx.field(capi.F_FDEN)
        with pytest.raises(RuntimeError, match="unknown field"):
            ctx.field(99)
        tab = np.zeros(len(ctx.potion) + 1)
        with pytest.raises(RuntimeError, match="table size changed"):
            ctx._chk(ctx.L.hetd_set_potion(ctx.h, tab.ctypes.data_as(capi.C.POINTER(capi.C.c_double)),
                                           len(tab), ctx.delinter))
        ctx.evolve(G.DT, 4)
        assert np.isfinite(ctx.energy()["etot"])
    p2 = capi.default_params(nx=16, ny=16, nz=16, xmax=4.0, ymax=4.0, zmax=4.0, npd=9)
    with Context(p2) as ctx:
        ctx.set_state(psi, (0, 0, 0), (0, 0, 0))
        with pytest.raises(RuntimeError, match="only with core OTC"):
            ctx.field(capi.F_UALPHAS)


# ------------------------------------------------------ mixed precision

@pytest.mark.parametrize("shape", [(64, 64, 64), (36, 40, 48)])
def test_mixed_precision_tracks_fp64(oracle_mod, shape):
    """HETD_MIXED: FP32 density, mean field and FFTs, FP64 integrator state,
    reductions and impurity.  Reported against the FP64 path (which matches
    the oracle to 1e-10): relative energy error <= 1e-6, psi within 1e-5."""
    say("mixed", shape)
    from hetd.initstate import woods_saxon_state
    nz, ny, nx = shape
    h = 0.6
    kw = dict(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=13, icon=13)
    psi = woods_saxon_state(nx, ny, nz, kw["xmax"], kw["ymax"], kw["zmax"], 150, 3)
    out = {}
    for prec in (capi.HETD_FP64, capi.HETD_MIXED):
        p = capi.default_params(precision=prec, **kw)
        with Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463) as ctx:
            ctx.set_state(psi, (0.3, -0.2, 0.1), (1.0, 2.0, -3.0))
            ctx.init()
            t = ctx.evolve(G.DT, 12)
            out[prec] = (t, ctx.field(capi.F_PSI), ctx.energy(), ctx.norm(), ctx.field(capi.F_POT4))
    t64, p64, e64, n64, u64 = out[capi.HETD_FP64]
    t32, p32, e32, n32, u32 = out[capi.HETD_MIXED]
    say("mixed errors: psi %.2e E %.2e norm %.2e pot4 %.2e" % (
        G.rel(p32, p64), abs(e32["etot4"] - e64["etot4"]) / abs(e64["etot4"]), abs(n32 - n64) / n64,
        G.rel(u32, u64)))
    assert G.rel(p32, p64) < 1e-5
    assert abs(e32["etot4"] - e64["etot4"]) <= 1e-6 * abs(e64["etot4"])
    assert abs(e32["etot"] - e64["etot"]) <= 1e-6 * abs(e64["etot"])
    assert abs(n32 - n64) <= 1e-9 * n64
    assert np.abs(t32[:, 2:5] - t64[:, 2:5]).max() < 1e-8 * max(np.abs(t64[:, 2:5]).max(), 1.0)
