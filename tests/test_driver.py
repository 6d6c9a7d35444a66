"""The Fortran driver bin/hetd_drive: the reference's namelist and mode-2
files in, the reference's trajectory files and checkpoint layout out."""
import os
import subprocess

import numpy as np
import pytest

import golden_util as G
from hetd.initstate import read_mode2, write_mode2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVE = os.path.join(ROOT, "4hetddft-isotropic_amd", "bin", "hetd_drive")


def namelist(cfg, **over):
    kv = dict(title="'golden'", mode=2, nx=cfg["nx"], ny=cfg["ny"], nz=cfg["nz"],
              xmax=repr(cfg["xmax"]), ymax=repr(cfg["ymax"]), zmax=repr(cfg["zmax"]),
              npd=cfg["npd"], icon=cfg["icon"], core4="'%s'" % cfg["core"],
              lsolid=".true." if cfg["lsolid"] else ".false.", selec="'%s'" % cfg["selec"],
              r_cutoff=repr(cfg["r_cutoff"]), umax=repr(cfg["umax"]), mimpur="131.29d0",
              deltatps="5.d-4", niter=cfg["niter"], pener=cfg["pener"], pcurr=1000000,
              filedenin="'den.in'", filedenout="'den.out'", iter0=0, time0=0, n4=cfg["n4"],
              tzmean=repr(cfg["tzmean"]), nthread=5, fftwplan="'FFTW_PATIENT'", afermi="1.5d0")
    kv.update(over)
    return "&input\n" + ",\n".join(" %s = %s" % (k, v) for k, v in kv.items()) + "\n/\n"


def run_drive(tmp_path, text):
    with open(os.path.join(tmp_path, "in.nml"), "w") as f:
        f.write(text)
    with open(os.path.join(tmp_path, "in.nml")) as fin:
        return subprocess.run([DRIVE], stdin=fin, cwd=tmp_path, capture_output=True, text=True, timeout=300)


def test_driver_built():
    assert os.path.exists(DRIVE), "bin/hetd_drive not built (run __graft_entry__.build())"


@pytest.mark.parametrize("over,msg", [(dict(nx=7), "NX is not correct"),
                                      (dict(deltat="1.d-4"), "Inconsistent deltat"),
                                      (dict(mode=0), "only mode=2")])
def test_driver_rejects_like_the_reference(tmp_path, over, msg):
    # these checks happen on the host before any device call
    cfg, _ = G.load("ot_xe_32")
    r = run_drive(str(tmp_path), namelist(cfg, **over))
    assert r.returncode != 0
    assert msg in (r.stdout + r.stderr)


@pytest.mark.gpu
def test_driver_reproduces_reference_run(tmp_path):
    cfg, d = G.load("ot_xe_32")
    write_mode2(os.path.join(tmp_path, "den.in"), d["psi0"], cfg["xmax"], cfg["ymax"], cfg["zmax"],
                d["rimp0"], d["vimp0"])
    r = run_drive(str(tmp_path), namelist(cfg))
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    rimp = np.loadtxt(os.path.join(tmp_path, "rimp.out"), comments="#")
    st = d["steps"]
    assert rimp.shape == (cfg["niter"] + 1, 4)
    # E18.10 output: 11 significant digits
    assert np.abs(rimp[1:, 1:] - st[:, 3:6]).max() < 2e-10 * np.abs(st[:, 3:6]).max() + 1e-12
    psi, grid, r_, v_ = read_mode2(os.path.join(tmp_path, "den.out"))
    assert G.rel(G.planes(psi), d["final_planes"]) < 1e-10
    assert "elapsed partie iterative" in r.stdout


@pytest.mark.gpu
def test_driver_restart_with_history_is_bitwise(tmp_path):
    """lhistory (an extension key): a run restarted from a checkpoint and its
    integrator history continues bit for bit as the uninterrupted run."""
    cfg, d = G.load("ot_xe_32")
    full, part = os.path.join(tmp_path, "full"), os.path.join(tmp_path, "part")
    for w in (full, part):
        os.makedirs(w)
        write_mode2(os.path.join(w, "den.in"), d["psi0"], cfg["xmax"], cfg["ymax"], cfg["zmax"],
                    d["rimp0"], d["vimp0"])
    r = run_drive(full, namelist(cfg, niter=16, pcurr=8, lhistory=".true."))
    assert r.returncode == 0, r.stderr
    ck = [f for f in sorted(os.listdir(full)) if f.endswith(".hist") and not f.startswith("den.out")]
    assert ck, os.listdir(full)
    first = ck[0][:-5]  # the checkpoint after 8 iterations
    os.replace(os.path.join(full, first), os.path.join(part, "den.ck"))
    os.replace(os.path.join(full, ck[0]), os.path.join(part, "den.ck.hist"))
    r2 = run_drive(part, namelist(cfg, niter=16, pcurr=1000000, lhistory=".true.", filedenin="'den.ck'",
                                  iter0=8, time0=repr(8 * 5e-4)))
    assert r2.returncode == 0, r2.stderr
    assert "Integrator history restored" in r2.stdout
    a = np.loadtxt(os.path.join(full, "rimp.out"), comments="#")
    b = np.loadtxt(os.path.join(part, "rimp.out"), comments="#")
    assert np.array_equal(a[-8:, 1:], b[-8:, 1:])
    psa = read_mode2(os.path.join(full, "den.out"))[0]
    psb = read_mode2(os.path.join(part, "den.out"))[0]
    assert np.array_equal(psa, psb)
