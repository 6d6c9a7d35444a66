"""Throughput of the real-time propagation step on one MI355X.

Workload: He-1000 droplet + Xe impurity on a 256^3 grid at 0.4 A, Orsay-Trento
functional with the solid penalty (the reference default), 13-point
stencils, zero halo (icon 13), dt = 0.5 fs.  A "step" is one steady-state
iteration of the reference loop (DFT4He3d.f90:843-852): Hamming
predictor-corrector with its middle mean field, then the driver's mean field
at the corrected state.  Inputs are resident in HBM when the timed region
starts; the first three (Runge-Kutta) iterations are part of the warm-up.

Prints one JSON line (the benchmark contract) with the roofline of the
dominant kernel class, measured with HIP events on the library's stream, and
a CPU baseline timed on this host: the reference's own Fortran routines
(oracle/_ref/ref_harness, flang + MKL FFTW3) when built, else the C oracle.
"""
import argparse
import ctypes as C
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "4hetddft-isotropic_amd")
sys.path.insert(0, PKG)

KTOPS = 7.63822291
HBM_PEAK = 8.0e12  # B/s, MI355X
METRIC = "Mpoint-steps/s, He-1000+imp 256^3: mixed vs FP64 GPU at ≤1e-6 ΔE/E; % HBM roof"


def model_bytes_per_point_step(nx, mixed=False):
    """Algorithmic HBM bytes per grid point per steady-state PC step:
    448 + 8 s_h (FP64) or 360 + 4 s_h (mixed), s_h = 16 (nx/2+1)/nx."""
    sh = 16.0 * (nx // 2 + 1) / nx
    return 360.0 + 4.0 * sh if mixed else 448.0 + 8.0 * sh


class HipEvents:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.L.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.L.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.L.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value

    def copy_calibration(self, mib, reps=3):
        """Device-to-device copies of `mib` MiB: the known-size reference
        for scaling FETCH_SIZE / WRITE_SIZE counters to bytes."""
        L = self.L
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipFree.argtypes = [C.c_void_p]
        nb = mib << 20
        a, b = C.c_void_p(), C.c_void_p()
        assert L.hipMalloc(C.byref(a), nb) == 0 and L.hipMalloc(C.byref(b), nb) == 0
        for _ in range(reps):
            assert L.hipMemcpy(b, a, nb, 3) == 0  # hipMemcpyDeviceToDevice
        assert L.hipDeviceSynchronize() == 0
        L.hipFree(a)
        L.hipFree(b)
        return nb


def cpu_baseline(n, nthreads, steps=2):
    """Reference CPU path on this host's cores, a bounded sample: `steps`
    predictor-corrector iterations of the same 256^3 workload."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if os.path.exists(harness):
        work = tempfile.mkdtemp(prefix="hetd_cpu_")
        try:
            nml = os.path.join(work, "nml")
            xm = 0.2 * n
            with open(nml, "w") as f:
                f.write("&harness\n nx=%d, ny=%d, nz=%d, xmax=%r, ymax=%r, zmax=%r, npd=13, icon=13,\n" % (
                    n, n, n, xm, xm, xm))
                f.write(" core4='OT', lsolid=.true., selec='Xe_He', r_cutoff=2.55d0, umax=2793.9365463d0,\n")
                f.write(" mimpur=131.29d0, deltatps=5.d-4, niter=%d, pener=100000, outdir='%s', n4=1000,\n" % (
                    steps, work))
                f.write(" nthread=%d, ldump_init=.false., mode=-1, nrk=0\n/\n" % nthreads)
            env = dict(os.environ, MKL_THREADING_LAYER="INTEL", OMP_NUM_THREADS=str(nthreads),
                       MKL_NUM_THREADS=str(nthreads), KMP_DUPLICATE_LIB_OK="TRUE")
            with open(nml) as fin:
                subprocess.run([harness], stdin=fin, cwd=work, env=env, check=True, timeout=600,
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            with open(os.path.join(work, "timing.txt")) as f:
                sec, nst = f.read().split()
            sec, nst = float(sec), int(nst)
            return {"value": n ** 3 * nst / sec / 1e6, "unit": "Mpoint-steps/s", "cores": nthreads,
                    "kind": "reference",
                    "sample": "%d predictor-corrector steps of He-1000+Xe %d^3 FP64, reference Fortran "
                              "routines (flang -O2 -fopenmp, MKL FFTW3), %.1f s" % (nst, n, sec)}
        except Exception as exc:  # fall through to the oracle
            print("reference harness failed: %r" % (exc,), file=sys.stderr)
        finally:
            shutil.rmtree(work, ignore_errors=True)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    from hetd import capi
    from hetd.initstate import woods_saxon_state
    m = 128
    p = capi.default_params(nx=m, ny=m, nz=m, xmax=0.2 * m, ymax=0.2 * m, zmax=0.2 * m, npd=13, icon=13)
    o = oracle.OracleCtx(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463)
    o.set_state(woods_saxon_state(m, m, m, p.xmax, p.ymax, p.zmax, 1000, 3), (0, 0, 0), (0, 0, 2 * KTOPS))
    o.init()
    t0 = time.perf_counter()
    o.steppc(5e-4 / KTOPS)
    o.potenimp(o.imp()[0])
    o.poten()
    o.set_aimp(o.forceimp(o.imp()[0]) / (p.mimpur * 0.02061484))
    sec = time.perf_counter() - t0
    o.close()
    del np
    return {"value": m ** 3 / sec / 1e6, "unit": "Mpoint-steps/s", "cores": 1, "kind": "port",
            "sample": "1 predictor-corrector step of He-1000+Xe %d^3 FP64, C oracle, %.1f s" % (m, sec)}


def run_precision(args, prec, psi, kclass, ev):
    """Warm up, then time args.steps PC iterations; returns a result dict."""
    from hetd import Context, capi
    n = args.n
    xm = 0.2 * n
    p = capi.default_params(nx=n, ny=n, nz=n, xmax=xm, ymax=xm, zmax=xm, npd=13, icon=13, precision=prec)
    t0 = time.perf_counter()
    ctx = Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463)
    ctx.set_state(psi, (0.0, 0.0, 0.0), (0.0, 0.0, 2.0 * KTOPS))
    ctx.init()
    dt = 5e-4 / KTOPS
    ctx.evolve(dt, max(args.warmup, 3), traj=False)
    tuned = ctx.autotune(dt, args.autotune) if args.autotune > 0 else None
    e0 = ctx.energy()
    setup_s = time.perf_counter() - t0
    stream = ctx.stream()
    a, b = ev.event(), ev.event()
    ctx.timing(kclass)
    ev.record(a, stream)
    w0 = time.perf_counter()
    ctx.evolve(dt, args.steps, traj=False)
    ev.record(b, stream)
    ms = ev.elapsed_ms(a, b)
    wall = time.perf_counter() - w0
    kms, klaunch, kbytes = ctx.timing_read()
    ctx.timing(capi.K_OFF)
    classes = {}
    for kc, name in capi.K_NAMES.items():
        ctx.timing(kc)
        ctx.evolve(dt, 2, traj=False)
        t, nl, by = ctx.timing_read()
        classes[name] = {"ms_per_step": round(t / 2, 4), "launches_per_step": nl // 2,
                         "GBps": round(by / (t * 1e-3) / 1e9, 1) if t > 0 else None}
    ctx.timing(capi.K_OFF)
    e1 = ctx.energy()
    norm = ctx.norm()
    ctx.close()
    nsteps_total = max(args.warmup, 3) + args.steps + 2 * len(capi.K_NAMES)
    return {"ms": ms, "wall": wall, "kms": kms, "klaunch": klaunch, "kbytes": kbytes, "classes": classes,
            "e0": e0, "e1": e1, "norm": norm, "setup_s": setup_s, "nsteps_total": nsteps_total, "tuned": tuned}


def accuracy_run(args, prec, psi):
    """Untimed: args.check_steps steps from the same state, recording the
    impurity trajectory every step and the energies at the end."""
    from hetd import Context, capi
    n = args.n
    xm = 0.2 * n
    p = capi.default_params(nx=n, ny=n, nz=n, xmax=xm, ymax=xm, zmax=xm, npd=13, icon=13, precision=prec)
    ctx = Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463)
    ctx.set_state(psi, (0.0, 0.0, 0.0), (0.0, 0.0, 2.0 * KTOPS))
    ctx.init()
    e0 = ctx.energy()
    traj = ctx.evolve(5e-4 / KTOPS, args.check_steps, traj=True)
    e1 = ctx.energy()
    norm = ctx.norm()
    r, v, _ = ctx.imp()
    ctx.close()
    return {"e0": e0, "e1": e1, "norm": norm, "traj": traj, "r": r, "v": v}


def ensemble_leg(args):
    """Batched ensemble of independent droplet + impurity runs of uneven size
    (mixed precision): aggregate throughput of hetd_pool_evolve against the
    same runs advanced one after another with hetd_evolve."""
    from hetd import Context, capi, pool_evolve
    from hetd.initstate import ensemble_members, woods_saxon_state
    members = ensemble_members(args.ensemble_runs)
    dt = 5e-4 / KTOPS
    ctxs = []
    t0 = time.perf_counter()
    for m in members:
        n, xm = m["n"], m["xmax"]
        p = capi.default_params(nx=n, ny=n, nz=n, xmax=xm, ymax=xm, zmax=xm, npd=13, icon=13,
                                precision=capi.HETD_MIXED)
        c = Context(p, selec="Xe_He", r_cutoff=2.55, umax=2793.9365463)
        c.set_state(woods_saxon_state(n, n, n, xm, xm, xm, m["n4"], m["seed"]), (0.0, 0.0, 0.0),
                    tuple(v * KTOPS for v in m["v"]))
        c.init()
        ctxs.append(c)
    setup_s = time.perf_counter() - t0
    pool_evolve(ctxs, dt, 6)  # Runge-Kutta start and the first PC steps
    k = args.ensemble_steps
    t0 = time.perf_counter()
    for c in ctxs:
        c.evolve(dt, k, traj=False)
    seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    pool_evolve(ctxs, dt, k)
    pool = time.perf_counter() - t0
    pts = sum(float(m["n"]) ** 3 for m in members)
    norms = [c.norm() / m["n4"] for c, m in zip(ctxs, members)]
    for c in ctxs:
        c.close()
    return {"runs": len(members), "steps_timed": k, "precision": "mixed",
            "grids": sorted(set(m["n"] for m in members)), "total_points": int(pts),
            "sequential_Mpoint_steps_per_s": round(pts * k / seq / 1e6, 1),
            "pool_Mpoint_steps_per_s": round(pts * k / pool / 1e6, 1),
            "pool_over_sequential": round(seq / pool, 3),
            "sequential_s": round(seq, 3), "pool_s": round(pool, 3), "setup_s": round(setup_s, 1),
            "norm_per_atom_range": [min(norms), max(norms)],
            "workload": "%d He-N + Xe runs, N uniform 200..3000 (seeds 100..), cubic grids 64..192 at 0.4 A, "
                        "impurity 0.5..4 A/ps in random directions, mixed precision, npd 13" % len(members)}


def compare_accuracy(a64, amx):
    """Mixed against FP64 after the same steps: relative energy errors (total
    and per component scale), norm, impurity position / velocity."""
    import numpy as np
    out = {"steps": int(len(a64["traj"]))}
    for k in ("etot", "etot4"):
        out["dE_over_E_" + k] = abs(amx["e1"][k] - a64["e1"][k]) / abs(a64["e1"][k])
    for k in ("ekin4", "elj4", "ecor4", "esolid", "ekinx", "eHeX"):
        out["dE_K_" + k] = amx["e1"][k] - a64["e1"][k]
    out["dnorm_over_norm"] = abs(amx["norm"] - a64["norm"]) / a64["norm"]
    out["max_dz_A"] = float(np.max(np.abs(amx["traj"][:, 4] - a64["traj"][:, 4])))
    out["dr_end_A"] = float(np.linalg.norm(amx["r"] - a64["r"]))
    out["dv_over_v_end"] = float(np.linalg.norm(amx["v"] - a64["v"]) / np.linalg.norm(a64["v"]))
    out["energy_drift_fp64"] = (a64["e1"]["etot"] - a64["e0"]["etot"]) / abs(a64["e0"]["etot"])
    out["energy_drift_mixed"] = (amx["e1"]["etot"] - amx["e0"]["etot"]) / abs(amx["e0"]["etot"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--kernel", default="update", choices=["update", "xpass", "ypass", "zpass", "force"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run device copies of MIB MiB (the byte reference for PMC traffic counters)")
    ap.add_argument("--ensemble-runs", type=int, default=32)
    ap.add_argument("--ensemble-steps", type=int, default=60,
                    help="timed steps of the batched-ensemble leg (0: skip)")
    ap.add_argument("--autotune", type=int, default=0, metavar="STEPS",
                    help="before timing, let hetd_autotune pick among the bit-identical kernel variants, "
                         "timing each over STEPS steps (0: off, the default kernels)")
    ap.add_argument("--check-steps", type=int, default=400,
                    help="untimed mixed-vs-FP64 accuracy run length (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the propagation step runs on one GPU; --gpus %d is not supported (independent "
              "runs go one process per GPU, with no communication)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    from hetd import capi
    from hetd.initstate import woods_saxon_state

    n = args.n
    xm = 0.2 * n
    psi = woods_saxon_state(n, n, n, xm, xm, xm, 1000, 3)
    kclass = {"update": capi.K_UPDATE, "xpass": capi.K_XPASS, "ypass": capi.K_YPASS,
              "zpass": capi.K_ZPASS, "force": capi.K_FORCE}[args.kernel]
    ev = HipEvents()
    if args.calibrate_copy:
        print("copy calibration: %d bytes per copy" % ev.copy_calibration(args.calibrate_copy), file=sys.stderr)
    res = {}
    if args.precision in ("both", "fp64"):
        res["fp64"] = run_precision(args, capi.HETD_FP64, psi, kclass, ev)
    if args.precision in ("both", "mixed"):
        res["mixed"] = run_precision(args, capi.HETD_MIXED, psi, kclass, ev)
    acc = None
    if len(res) == 2 and args.check_steps > 0:
        acc = compare_accuracy(accuracy_run(args, capi.HETD_FP64, psi), accuracy_run(args, capi.HETD_MIXED, psi))
    del psi
    ens = ensemble_leg(args) if args.ensemble_steps > 0 and args.ensemble_runs > 0 else None
    head = "mixed" if "mixed" in res else "fp64"
    r = res[head]
    pts = float(n) ** 3
    ms_step = r["ms"] / args.steps
    value = pts * args.steps / (r["ms"] * 1e-3) / 1e6
    avg_launch_s = r["kms"] * 1e-3 / max(r["klaunch"], 1)
    achieved = (r["kbytes"] / max(r["klaunch"], 1)) / avg_launch_s / 1e9
    per_prec = {}
    for k, v in res.items():
        mstep = v["ms"] / args.steps
        per_prec[k] = {"Mpoint_steps_per_s": round(pts / (mstep * 1e-3) / 1e6, 2), "ms_per_step": round(mstep, 4),
                       "step_hbm_roof_frac": round(model_bytes_per_point_step(n, k == "mixed") * pts /
                                                   (mstep * 1e-3) / HBM_PEAK, 4),
                       "model_bytes_per_point_step": round(model_bytes_per_point_step(n, k == "mixed"), 2),
                       "E_per_atom_end": v["e1"]["etot4"] / 1000.0, "norm_end": v["norm"],
                       "kernel_classes": v["classes"],
                       "variants": None if v["tuned"] is None else
                       {"split_exchange": bool(v["tuned"] & 1),
                        "update_kernels": ["default", "two LDS planes", "two LDS planes, non-temporal"][v["tuned"] >> 1]}}
    cmp_ = {}
    if len(res) == 2:
        a64, amx = res["fp64"], res["mixed"]
        cmp_ = {"speedup_mixed_over_fp64": round(a64["ms"] / amx["ms"], 3),
                "dE_over_E_mixed_vs_fp64_etot4": abs(amx["e1"]["etot4"] - a64["e1"]["etot4"]) / abs(a64["e1"]["etot4"]),
                "dE_over_E_mixed_vs_fp64_etot": abs(amx["e1"]["etot"] - a64["e1"]["etot"]) / abs(a64["e1"]["etot"]),
                "dnorm_over_norm": abs(amx["norm"] - a64["norm"]) / a64["norm"],
                "steps_compared": a64["nsteps_total"], "accuracy_run": acc}
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%s_%d.json" % (args.kernel, head, n))
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    cpu = None
    if not args.no_cpu:
        nth = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(n, nth)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 3),
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic (Woods-Saxon He-1000 droplet, seeded phase, Xe at the centre moving 200 m/s)",
        "config": {"workload": "He-1000+Xe %d^3 0.4A OT+solid npd13 icon13 dt0.5fs predictor-corrector steps" % n,
                   "grid": [n, n, n], "precision": head,
                   "precision_detail": ("f32 density, mean field and FFTs; f64 wave function, integrator "
                                        "state and reductions" if head == "mixed" else "f64 throughout"),
                   "wall_ms_per_step": round(r["wall"] * 1e3 / args.steps, 4),
                   "per_precision": per_prec, "mixed_vs_fp64": cmp_, "ensemble": ens},
        "roofline": {"bound": "hbm", "kernel": args.kernel, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved * 1e9 / HBM_PEAK, 4),
                     "traffic": traffic, "launches": r["klaunch"],
                     "bytes_per_launch": r["kbytes"] / max(r["klaunch"], 1),
                     "avg_launch_us": round(avg_launch_s * 1e6, 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
