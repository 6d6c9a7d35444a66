"""Graph-replayed steps and the ensemble entry point against eager launches.

hetd_evolve without a trajectory replays steady-state predictor-corrector
iterations from a captured HIP graph; hetd_pool_evolve interleaves several
contexts.  Both must give bitwise the states of the eager, one-context path
(which the parity tests hold to the oracle).  Replays are checked several
times in a row, with eager work (energy, norm, field copies) in between."""
import os
import sys

import numpy as np
import pytest

import golden_util as G
from hetd import Context, capi, pool_evolve
from hetd.initstate import woods_saxon_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(shape, prec, icon, graphs, n4=150, seed=3, r0=(0.3, -0.2, 0.1), v0=(1.0, 2.0, -3.0), xsplit=False):
    nz, ny, nx = shape
    h = 0.6
    p = capi.default_params(nx=nx, ny=ny, nz=nz, xmax=nx * h / 2, ymax=ny * h / 2, zmax=nz * h / 2, npd=13,
                            icon=icon, precision=prec)
    os.environ["HETD_GRAPHS"] = "1" if graphs else "0"
    os.environ["HETD_XSPLIT"] = "1" if xsplit else "0"
    try:
        ctx = Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463)
    finally:
        del os.environ["HETD_GRAPHS"]
        del os.environ["HETD_XSPLIT"]
    ctx.set_state(woods_saxon_state(nx, ny, nz, p.xmax, p.ymax, p.zmax, n4, seed), r0, v0)
    ctx.init()
    return ctx


def same_state(a, b):
    assert np.array_equal(a.field(capi.F_PSI), b.field(capi.F_PSI))
    for x, y in zip(a.imp(), b.imp()):
        assert np.array_equal(x, y)
    ea, eb = a.energy(), b.energy()
    for k in ea:
        assert ea[k] == eb[k], k


@pytest.mark.parametrize("shape,prec,icon", [((48, 40, 36), capi.HETD_FP64, 13),
                                             ((32, 32, 32), capi.HETD_MIXED, 8)])
def test_graph_replays_match_eager(shape, prec, icon):
    say("graphs", shape, prec, icon)
    eager = make(shape, prec, icon, graphs=False)
    graph = make(shape, prec, icon, graphs=True)
    for n in (9, 14, 6, 25, 12):
        eager.evolve(G.DT, n, traj=False)
        graph.evolve(G.DT, n, traj=False)
        assert eager.norm() == graph.norm()
        same_state(eager, graph)
    eager.close()
    graph.close()


def test_pool_matches_sequential():
    specs = [((32, 32, 32), capi.HETD_FP64, 13), ((40, 36, 48), capi.HETD_MIXED, 13),
             ((30, 32, 45), capi.HETD_FP64, 8)]
    say("pool", specs)
    seq = [make(s, p, i, graphs=True, seed=5 + k) for k, (s, p, i) in enumerate(specs)]
    pool = [make(s, p, i, graphs=True, seed=5 + k) for k, (s, p, i) in enumerate(specs)]
    for n in (17, 8):
        for c in seq:
            c.evolve(G.DT, n, traj=False)
        pool_evolve(pool, G.DT, n)
        for a, b in zip(seq, pool):
            same_state(a, b)
    for c in seq + pool:
        c.close()


@pytest.mark.parametrize("shape,prec,icon", [((36, 40, 32), capi.HETD_FP64, 13),
                                             ((32, 32, 40), capi.HETD_MIXED, 8)])
def test_history_restart_is_bitwise(shape, prec, icon):
    """A context rebuilt from psi and the saved integrator history continues
    bit for bit as the context it was saved from (graph replays included)."""
    say("history", shape, prec, icon)
    a = make(shape, prec, icon, graphs=True)
    a.evolve(G.DT, 11, traj=False)
    blob = a.save_history()
    psi = a.field(capi.F_PSI)
    r, v = a.imp()[:2]
    b = make(shape, prec, icon, graphs=True, seed=9)
    b.set_state(psi, r, v)
    b.load_history(blob)
    for n in (7, 12):
        a.evolve(G.DT, n, traj=False)
        b.evolve(G.DT, n, traj=False)
        same_state(a, b)
    # rejected blobs: truncated, not a history, another grid; the state is kept
    with pytest.raises(RuntimeError):
        b.load_history(blob[:-8])
    bad = bytearray(blob)
    bad[0:8] = b"NOT_HIST"
    with pytest.raises(RuntimeError, match="not a history"):
        b.load_history(bytes(bad))
    c = make((16, 16, 16), prec, icon, graphs=False)
    with pytest.raises(RuntimeError, match="another grid"):
        c.load_history(blob)
    c.close()
    a.evolve(G.DT, 3, traj=False)
    b.evolve(G.DT, 3, traj=False)
    same_state(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("shape,prec,icon", [((20, 18, 256), capi.HETD_FP64, 13), ((16, 24, 64), capi.HETD_MIXED, 8),
                                             ((30, 20, 16), capi.HETD_FP64, 13), ((24, 24, 256), capi.HETD_MIXED, 13),
                                             ((32, 64, 36), capi.HETD_FP64, 13), ((30, 16, 40), capi.HETD_MIXED, 8)])
def test_split_exchange_x_pass_is_bitwise(shape, prec, icon):
    """The x and y register passes with the LDS exchange split in real and
    imaginary halves (HETD_XSPLIT) compute the same bits as the complex
    exchange (x lengths 256, 64, 16; y lengths 64, 16)."""
    say("xsplit", shape, prec, icon)
    a = make(shape, prec, icon, graphs=True)
    b = make(shape, prec, icon, graphs=True, xsplit=True)
    for n in (5, 9):
        a.evolve(G.DT, n, traj=False)
        b.evolve(G.DT, n, traj=False)
        same_state(a, b)
    for w in (capi.F_POT4, capi.F_DENCG, capi.F_DELJ4):
        assert np.array_equal(a.field(w), b.field(w))
    a.close()
    b.close()


@pytest.mark.parametrize("shape,prec,icon,tile", [((36, 40, 48), capi.HETD_FP64, 13, "4"),
                                                  ((32, 24, 40), capi.HETD_MIXED, 8, "4"),
                                                  ((30, 32, 45), capi.HETD_FP64, 18, "5"),
                                                  ((32, 24, 40), capi.HETD_MIXED, 13, "5")])
def test_pingpong_update_variant_is_bitwise(shape, prec, icon, tile):
    """HETD_TILE_UPDATE=4 (update kernels with two LDS planes in turn, one
    barrier per plane) and 5 (the same with non-temporal operand streams)
    compute the same bits as the default update kernels."""
    say("update tile", tile, shape, prec, icon)
    a = make(shape, prec, icon, graphs=False)
    b = make(shape, prec, icon, graphs=False)
    for n in (5, 8):
        a.evolve(G.DT, n, traj=False)
        os.environ["HETD_TILE_UPDATE"] = tile
        try:
            b.evolve(G.DT, n, traj=False)
        finally:
            del os.environ["HETD_TILE_UPDATE"]
        same_state(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("shape,prec,icon", [((36, 40, 48), capi.HETD_FP64, 13), ((32, 24, 40), capi.HETD_MIXED, 8)])
def test_force_energy_tile_variants_agree(shape, prec, icon):
    """HETD_TILE_FORCE / HETD_TILE_ENERGY = 0..5 (the tile shapes of the
    standalone force and of the energy sums): the sums are taken over other
    tiles, so results agree to rounding (1e-12), not bit for bit."""
    say("force/energy tiles", shape, prec, icon)
    ref = make(shape, prec, icon, graphs=False)
    ref.evolve(G.DT, 6, traj=False)
    e_ref = ref.energy()
    r_ref = ref.imp()
    psi_ref = ref.field(capi.F_PSI)
    ref.close()
    for v in "012345":
        os.environ["HETD_TILE_FORCE"] = v
        os.environ["HETD_TILE_ENERGY"] = v
        try:
            c = make(shape, prec, icon, graphs=False)
            c.evolve(G.DT, 6, traj=False)
            e = c.energy()
            r = c.imp()
            psi = c.field(capi.F_PSI)
            c.close()
        finally:
            del os.environ["HETD_TILE_FORCE"]
            del os.environ["HETD_TILE_ENERGY"]
        for k in e_ref:
            assert abs(e[k] - e_ref[k]) <= 1e-12 * max(abs(e_ref[k]), 1.0), (v, k)
        for x, y in zip(r, r_ref):
            assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300), v
        assert G.rel(psi, psi_ref) < 1e-12, v


@pytest.mark.parametrize("shape,prec,icon", [((32, 32, 64), capi.HETD_FP64, 13), ((30, 32, 16), capi.HETD_MIXED, 8)])
def test_autotune_leaves_state_and_results_unchanged(shape, prec, icon):
    """hetd_autotune times variants that compute the same bits and restores
    the state: afterwards the context evolves bit for bit as an untuned one."""
    say("autotune", shape, prec, icon)
    a = make(shape, prec, icon, graphs=True)
    b = make(shape, prec, icon, graphs=True)
    for c in (a, b):
        c.evolve(G.DT, 5, traj=False)
    choice = b.autotune(G.DT, 4)
    assert 0 <= choice <= 5
    same_state(a, b)
    for n in (6, 9):
        a.evolve(G.DT, n, traj=False)
        b.evolve(G.DT, n, traj=False)
        same_state(a, b)
    a.close()
    b.close()
