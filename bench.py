#!/usr/bin/env python3
"""Benchmark of the MI355X dycore step (one GPU).

Workload: the dry convective-boundary-layer LES, 256 x 256 x 128 cells
(dx = dy = 20 m, dz = 10 m), 5th-order upwind advection, Smagorinsky Cs = 0.1,
Coriolis + geostrophic forcing, no-slip bottom with surface heat flux, FP64,
dt = 0.5 s, MRI ratio 16 (25 acoustic substeps per RK3 step).  Synthetic
seeded initial state.  A step is one full RK3 step (slow RHS, acoustic
substeps, boundary fills) with the state resident in HBM.

Prints one JSON line: cell-updates/s per RK3 step, the roofline of the
dominant kernel (HBM bytes per launch / average launch time, HIP events on the
library's stream) and the CPU oracle timed on a bounded sample.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fast_rhs_N_column_cbl256x256x128_traffic.json")


def measured_traffic(kernel, bytes_per_launch):
    """HBM bytes per launch of the timed kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement, if it was taken on this kernel and
    workload (same name and algorithmic byte count); otherwise None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel or abs(t.get("algorithmic_bytes_per_launch", -1) - bytes_per_launch) > 0.5:
        return None
    return t["traffic_bytes_per_launch"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--workload", choices=["cbl", "moist_terrain", "thermal", "ensemble", "supercell"], default="cbl",
                    help="cbl (default): the dry CBL LES; moist_terrain: moist flow over a hill on the "
                         "terrain-following path (stretched levels, WENO5, Kessler, Rayleigh), dt = 1 s; "
                         "thermal: the 64^3 dry rising thermal, C2, dt = 1 s, 6 substeps per stage; "
                         "ensemble: dry CBL members of uneven box sizes advanced as one batch; "
                         "supercell: Weisman-Klemp deep convection, the learned Kessler stand-in trained on "
                         "the plain method's runs and run against plain Kessler")
    ap.add_argument("--sc-seeds", type=int, default=2, help="supercell: training runs (seeds 0..n-1)")
    ap.add_argument("--sc-steps", type=int, default=600, help="supercell: steps of each training and evaluation run")
    ap.add_argument("--sc-every", type=int, default=2, help="supercell: rows harvested every n-th step")
    ap.add_argument("--sc-rows", type=int, default=4_000_000, help="supercell: row cap per training run")
    ap.add_argument("--sc-fit", type=int, default=3000, help="supercell: Adam steps")
    ap.add_argument("--members", type=int, default=48, help="ensemble members (--workload ensemble)")
    ap.add_argument("--ens-streams", type=int, default=2, help="ensemble: member groups, one stream each")
    ap.add_argument("--ens-box", type=int, nargs=3, default=None, metavar=("NX", "NY", "NZ"),
                    help="ensemble: every member of this one box size instead of the five drawn sizes")
    ap.add_argument("--unbatched-steps", type=int, default=2,
                    help="ensemble: steps timed with every member launched on its own, for comparison")
    ap.add_argument("--graph", action="store_true", help="replay the steps from a captured hipGraph")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="mixed: FP32 face values of the advected quantities on perturbations; the same steps "
                         "are then run in FP64 too and the relative L2 / Linf differences are reported")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default acoustic column kernel instead of timing its variants in the warm-up")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB (3 launches), the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counter readings")
    return ap.parse_args()


def calibrate_copy():
    import torch
    n = 1 << 27                                   # 2^27 doubles = 1 GiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": n * 8, "bytes_written": n * 8,
                      "launches": 3}))


def cpu_baseline(nx, ny, nz, budget_s):
    """Time the CPU oracle (test infrastructure, OpenMP) on a bounded sample of
    the same workload: a 64 x 64 x nz column block of the CBL case, as many
    steps as fit in ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from erf_amd import problems as P
    from oracle_lib import Oracle
    sx = min(nx, 64)
    sy = min(ny, 64)
    cfg, f = P.cbl(sx, sy, nz, seed=0)
    orc = Oracle(cfg)
    orc.load(f)
    orc.step(0.0, 0.5, 1)                     # warm-up
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.step(0.5 * (n + 1), 0.5, 1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": sx * sy * nz * n / dt, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": "%d steps of the CBL case on a %dx%dx%d block (OpenMP C++ oracle, -O2)" % (n, sx, sy, nz)}


def run_ensemble(a):
    """The batched ensemble: a.members dry CBL members (own seed, geostrophic
    speed in [5, 15] m/s, box size drawn from five shapes), every kernel of the
    step launched once for all members.  Timed batched, then with every member
    launched on its own; with --precision mixed the same steps run in FP64 and
    the per-member errors are reported."""
    import numpy as np
    import torch
    from erf_amd import config as C, problems as P
    from erf_amd.lib import Sim, Ensemble

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    dt = 0.5
    mixed = a.precision == "mixed"
    sims, refs, shapes = [], [], []
    sizes = (tuple(a.ens_box),) if a.ens_box else P.ENSEMBLE_SIZES
    for cfg, f in P.cbl_ensemble(a.members, seed=0, sizes=sizes):
        shapes.append((cfg.nx, cfg.ny, cfg.nz))
        if mixed:
            c64 = C.Config.from_buffer_copy(cfg)
            r = Sim(c64)
            r.load(f)
            refs.append(r)
            cfg.precision = C.PREC_MIXED
        s = Sim(cfg)
        s.load(f)
        sims.append(s)
        del f
    print("%d members uploaded" % len(sims), file=sys.stderr, flush=True)
    ens = Ensemble(sims)
    ens.set_streams(a.ens_streams)
    if a.graph:
        ens.set_graph(True)
    warm = a.warmup
    if not a.no_autotune:
        ens.autotune(True)           # one warm-up step per column variant
        warm = max(warm, 9)
    ens.step(0.0, dt, warm)
    t = warm * dt
    variant, var_ms = ens.column_variant()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ens.step(t, dt, a.steps)
    wall = time.perf_counter() - w0
    t += a.steps * dt
    launches, cells = ens.info()
    value = cells * a.steps / wall
    # the same members, each launched on its own (the launch pattern of one box per kernel)
    un = None
    if a.unbatched_steps > 0:
        for s in sims:
            s.set_column_variant(variant)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        for n in range(a.unbatched_steps):
            for s in sims:
                s.step_async(t + n * dt, dt, 1)
        for s in sims:
            s.sync()
        uw = time.perf_counter() - w0
        t += a.unbatched_steps * dt
        un = {"value": cells * a.unbatched_steps / uw, "steps": a.unbatched_steps,
              "ms_per_step": 1e3 * uw / a.unbatched_steps}
    total_steps = warm + a.steps + (a.unbatched_steps if un else 0)
    # sanity and per-member errors against FP64
    worst_rho = 1e30
    errs = {}
    if mixed:
        ens64 = Ensemble(refs)
        ens64.step(0.0, dt, total_steps)
        ens64.close()
    for m, s in enumerate(sims):
        nx, ny, nz = shapes[m]
        ng = s.L.ng
        sl = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
        cons = s.download(C.F_CONS)
        worst_rho = min(worst_rho, float(cons[0][sl].min()))
        if not mixed:
            continue
        ref = {C.F_CONS: refs[m].download(C.F_CONS)}
        got = {C.F_CONS: cons}
        for fid, comp, name in [(C.F_CONS, C.RHO, "rho"), (C.F_CONS, C.RHOTHETA, "rhotheta"), (C.F_XVEL, 0, "u"),
                                (C.F_YVEL, 0, "v"), (C.F_ZVEL, 0, "w")]:
            if fid not in got:
                got[fid] = s.download(fid)
                ref[fid] = refs[m].download(fid)
            x = got[fid][comp][sl]
            y = ref[fid][comp][sl]
            r = y - y.mean() if name in ("rho", "rhotheta") else y
            nrm = float(np.sqrt((r ** 2).mean()))
            errs.setdefault(name, []).append(float(np.sqrt(((x - y) ** 2).mean())) / nrm if nrm > 0 else 0.0)
    assert np.isfinite(worst_rho) and worst_rho > 0.05, "state blew up"
    ens.close()
    for s in sims + refs:
        s.close()
    hist = {}
    for sh in shapes:
        k = "%dx%dx%d" % sh
        hist[k] = hist.get(k, 0) + 1
    out = {
        "metric": "cell-updates/s per RK3 step (batched ensemble, %s GPU path)" % ("mixed" if mixed else "FP64"),
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": warm,
        "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if mixed else "f64", "data": "synthetic (seeded)",
        "config": {"workload": "%d dry CBL LES members (U5, Smagorinsky, Coriolis+geostrophic, dt=0.5 s, MRI ratio 16), "
                               "own seeds, geostrophic speed in [5, 15] m/s, box sizes drawn from %s"
                               % (len(shapes), ", ".join("%dx%dx%d" % s for s in sizes)),
                   "members": hist, "cells": cells, "precision": a.precision, "graph": bool(a.graph),
                   "streams": a.ens_streams},
        "launches_per_step": launches,
        "column_variant": {"variant": variant, "step_ms": var_ms},
        "unbatched": un,
        "wall_s": wall,
    }
    if un:
        out["batched_over_unbatched"] = value / un["value"]
    if mixed:
        out["error_vs_fp64_per_member"] = {k: {"max": max(v), "median": float(np.median(v))} for k, v in errs.items()}
    print(json.dumps(out))


def run_supercell(a):
    """Config of the learned microphysics: Weisman-Klemp deep convection
    (256x256x80, dx = 1 km, dz = 250 m, dt = 2 s).  Rows of Kessler's own
    tendencies are harvested from the plain method's runs (seeds 0..n-1), the
    network is trained on the GPU, and the held-out run (seed 8) is advanced
    with plain Kessler and with the network from the same state: both timed,
    the network's run compared with Kessler's."""
    import numpy as np
    import torch
    from erf_amd import config as C, problems as P, emulator as E
    from erf_amd.lib import Sim, MLP

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    nx, ny, nz = (a.nx, a.ny, a.nz) if a.nx != 256 or a.nz != 128 else (256, 256, 80)
    dt = 2.0
    case = lambda seed: P.supercell(nx, ny, nz, seed=seed)
    w0 = time.perf_counter()
    rows = E.harvest(seeds=tuple(range(a.sc_seeds)), nsteps=a.sc_steps, dt=dt, case=case, every=a.sc_every,
                     max_rows=a.sc_rows)
    held = E.harvest(seeds=(8,), nsteps=a.sc_steps, dt=dt, case=case, every=4 * a.sc_every,
                     max_rows=min(a.sc_rows, 1_000_000))
    t_harvest = time.perf_counter() - w0
    print("rows: %d training, %d held out (%.0f s)" % (rows.shape[0], held.shape[0], t_harvest),
          file=sys.stderr, flush=True)
    if rows.shape[0] < 1000 or held.shape[0] < 100:
        sys.exit("supercell: too few cells with active microphysics in %d steps (%d training, %d held-out rows); "
                 "the storm forms after about 150 steps: raise --sc-steps" % (a.sc_steps, rows.shape[0], held.shape[0]))
    sc = E.Scaling(rows)
    untrained = MLP(64)
    untrained.set_params(E.init_params(0))
    sc.apply(untrained)
    mse0 = E.mse(untrained, sc, held)
    w0 = time.perf_counter()
    net, sc, losses = E.fit(rows, steps=a.sc_fit, batch=8192, lr=2e-3, scaling=sc)
    t_fit = time.perf_counter() - w0
    mse1 = E.mse(net, sc, held)
    del rows, held
    cfg, f = case(8)
    cmx = C.Config.from_buffer_copy(cfg)
    cmx.precision = C.PREC_MIXED
    ka, em, mx = Sim(cfg), Sim(cfg), Sim(cmx)
    for s in (ka, em, mx):
        s.load(f)
    del f
    em.set_mlp(net)
    res = {}
    for name, s in (("kessler", ka), ("mlp", em), ("mixed", mx)):
        s.step(0.0, dt, a.sc_steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s.step(a.sc_steps * dt, dt, a.steps)
        res[name] = (time.perf_counter() - t1) / a.steps
    n_mlp, n_fb = em.mlp_counts()
    ng = ka.L.ng
    sl = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    slw = (slice(ng, ng + nz + 1), slice(ng, ng + ny), slice(ng, ng + nx))
    ck = ka.download(C.F_CONS)
    wk = ka.download(C.F_ZVEL)[0][slw]
    rho_k = ck[C.RHO][sl]

    def rel_errors(sim):
        ce = sim.download(C.F_CONS)
        out = {}
        for name, comp in (("rho", C.RHO), ("rhotheta", C.RHOTHETA), ("qv", C.RHOQ1), ("qc", C.RHOQ2),
                           ("qr", C.RHOQ3), ("w", None)):
            if comp is None:
                x, y = sim.download(C.F_ZVEL)[0][slw], wk
            else:
                x, y = ce[comp][sl], ck[comp][sl]
                if comp not in (C.RHO, C.RHOTHETA):
                    x, y = x / ce[C.RHO][sl], y / rho_k
            r = y - y.mean(axis=(1, 2), keepdims=True) if name in ("rho", "rhotheta") else y
            nrm = float(np.sqrt((r ** 2).mean()))
            out[name] = float(np.sqrt(((x - y) ** 2).mean())) / nrm if nrm > 0 else 0.0
        return out

    errs = rel_errors(em)
    floor = rel_errors(mx)
    qc_max = float((ck[C.RHOQ2][sl] / rho_k).max())
    for s in (ka, em, mx):
        s.close()
    cells = nx * ny * nz
    out = {
        "metric": "cell-updates/s per RK3 step (FP64 dycore, learned microphysics)",
        "value": cells / res["mlp"], "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.sc_steps, "ms_per_step": 1e3 * res["mlp"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+bf16", "data": "synthetic (seeded)",
        "config": {"workload": "supercell-type deep convection %dx%dx%d, dx=%g m, dz=%g m, Weisman-Klemp sounding "
                               "(qv <= 14 g/kg), quarter-circle hodograph (5 m/s), +3 K warm bubble, Kessler with "
                               "rain, dt=2 s, MRI ratio 6" % (nx, ny, nz, cfg.dx, cfg.dz),
                   "cells": cells, "train_seeds": list(range(a.sc_seeds)), "eval_seed": 8,
                   "steps_per_run": a.sc_steps},
        "kessler": {"value": cells / res["kessler"], "ms_per_step": 1e3 * res["kessler"]},
        "mlp_cells": n_mlp, "fallback_cells": n_fb,
        "heldout_mse_scaled": {"untrained": mse0, "trained": mse1},
        "fit_s": t_fit, "adam_steps": a.sc_fit, "harvest_s": t_harvest,
        "relative_l2_vs_kessler_after_run": errs,
        "relative_l2_mixed_precision_kessler_vs_fp64_kessler": floor,
        "qc_max_kessler": qc_max,
    }
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the x-y domain decomposition is a design note, not code)", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.workload == "ensemble":
        run_ensemble(a)
        return
    if a.workload == "supercell":
        run_supercell(a)
        return
    import numpy as np
    import torch
    from erf_amd import config as C, problems as P
    from erf_amd.lib import Sim, lib, check

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    ray = None
    if a.workload == "thermal":
        nx, ny, nz = (a.nx, a.ny, a.nz) if a.nx != 256 or a.nz != 128 else (64, 64, 64)
        cfg, f = P.rising_thermal(nx, ny, nz, seed=0)
        dt = 1.0
    elif a.workload == "moist_terrain":
        nx, ny, nz = (a.nx, a.ny, a.nz) if a.nx != 256 or a.nz != 128 else (512, 512, 96)
        print("building the moist terrain case %dx%dx%d" % (nx, ny, nz), file=sys.stderr, flush=True)
        cfg, f = P.moist_terrain(nx, ny, nz, seed=0)
        ray = P.rayleigh_profiles_uniform(cfg, f=f)
        dt = 1.0
    else:
        cfg, f = P.cbl(a.nx, a.ny, a.nz, seed=0)
        dt = 0.5
    if a.precision == "mixed":
        cfg.precision = C.PREC_MIXED

    def make_sim(c):
        s_ = Sim(c)
        s_.load(f)
        if ray is not None:
            s_.set_rayleigh(*ray)
        return s_

    sim = make_sim(cfg)
    print("state uploaded", file=sys.stderr, flush=True)
    stream = torch.cuda.ExternalStream(sim.stream)

    if a.graph:
        sim.set_graph(True)
    if not a.no_autotune and a.warmup > 0:
        sim.autotune(True)       # the warm-up steps time the column kernel's variants
    sim.step(0.0, dt, a.warmup)
    t = a.warmup * dt
    col_variant, col_ms = sim.column_variant()
    extra = 0
    while col_variant < 0 and extra < 5:     # short runs: finish timing the variants before the timed steps
        sim.step(t, dt, 1)
        t += dt
        extra += 1
        col_variant, col_ms = sim.column_variant()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    if a.workload == "cbl" and not a.graph:
        check(lib().erfx_sim_timer_reset(sim.h))
    with torch.cuda.stream(stream):
        e0.record(stream)
    wall0 = time.perf_counter()
    sim.step_async(t, dt, a.steps)
    with torch.cuda.stream(stream):
        e1.record(stream)
    sim.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = e0.elapsed_time(e1)

    cells = cfg.nx * cfg.ny * cfg.nz
    nsub = 1 + cfg.mri_ratio // 2 + cfg.mri_ratio
    value = cells * a.steps / (ms * 1e-3)

    # dominant kernel: per-launch timing collected by the library with HIP events
    kname = ctypes.create_string_buffer(128)
    kms = ctypes.c_double()
    kcnt = ctypes.c_long()
    kbytes = ctypes.c_double()
    check(lib().erfx_sim_timer_read(sim.h, kname, 128, ctypes.byref(kms), ctypes.byref(kcnt), ctypes.byref(kbytes)))
    avg_ms = kms.value / max(kcnt.value, 1)
    achieved = kbytes.value / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0

    # sanity: the state stayed physical
    rho = sim.download(C.F_CONS)[0]
    ng = sim.L.ng
    rv = rho[ng:ng + cfg.nz, ng:ng + cfg.ny, ng:ng + cfg.nx]
    assert np.all(np.isfinite(rv)) and rv.min() > 0.05, "state blew up"
    errors = None
    if a.precision == "mixed":
        # the same steps in FP64 from the same state: relative L2 / Linf of the difference
        c64 = C.Config.from_buffer_copy(cfg)
        c64.precision = C.PREC_FP64
        ref = make_sim(c64)
        ref.step(0.0, dt, a.warmup + extra + a.steps)
        errors = {}
        fields = [(C.F_CONS, C.RHO, "rho"), (C.F_CONS, C.RHOTHETA, "rhotheta"), (C.F_XVEL, 0, "u"),
                  (C.F_YVEL, 0, "v"), (C.F_ZVEL, 0, "w")]
        if cfg.moisture_type:
            fields += [(C.F_CONS, C.RHOQ1, "qv"), (C.F_CONS, C.RHOQ2, "qc"), (C.F_CONS, C.RHOQ3, "qr")]
        got = {}
        for fid, comp, name in fields:
            if fid not in got:
                got[fid] = (sim.download(fid), ref.download(fid))
            x = got[fid][0][comp, ng:ng + cfg.nz, ng:ng + cfg.ny, ng:ng + cfg.nx]
            y = got[fid][1][comp, ng:ng + cfg.nz, ng:ng + cfg.ny, ng:ng + cfg.nx]
            r = y - y.mean() if name in ("rho", "rhotheta") else y
            nrm = float(np.sqrt((r ** 2).mean()))
            errors[name] = {"rel_l2": float(np.sqrt(((x - y) ** 2).mean())) / nrm if nrm > 0 else 0.0,
                            "rel_linf": float(np.abs(x - y).max() / np.abs(r).max()) if np.abs(r).max() > 0 else 0.0}
        ref.close()
    sim.close()

    if a.workload != "cbl" or a.graph or a.precision == "mixed":
        wl = {"thermal": "dry rising thermal %dx%dx%d, C2 advection, dt=1 s, MRI ratio %d (%d substeps/step)",
              "moist_terrain": "moist flow over a 500 m Gaussian hill %dx%dx%d, dx=200 m, levels stretched "
                               "20-400 m to a 15 km top, WENO5 scalars, U3 momentum, Kessler, Rayleigh above "
                               "12 km, dt=1 s, MRI ratio %d (%d substeps/step)",
              "cbl": "dry CBL LES %dx%dx%d, U5 advection, Smagorinsky, Coriolis+geostrophic, dt=0.5 s, "
                     "MRI ratio %d (%d substeps/step)"}[a.workload]
        out = {
            "metric": "cell-updates/s per RK3 step (%s GPU path)" % ("mixed" if a.precision == "mixed" else "FP64"),
            "value": value, "unit": "cell-updates/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms / a.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64+f32" if a.precision == "mixed" else "f64",
            "data": "synthetic (seeded)",
            "config": {"workload": wl % (cfg.nx, cfg.ny, cfg.nz, cfg.mri_ratio, nsub),
                       "cells": cells, "precision": a.precision, "graph": bool(a.graph)},
            "wall_s": wall,
        }
        if errors is not None:
            out["error_vs_fp64"] = errors
        print(json.dumps(out))
        return

    out = {
        "metric": "cell-updates/s per RK3 step (FP64 GPU path)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": "dry CBL LES %dx%dx%d, U5 advection, Smagorinsky, Coriolis+geostrophic, "
                               "dt=0.5 s, MRI ratio %d (%d substeps/step)" % (cfg.nx, cfg.ny, cfg.nz,
                                                                             cfg.mri_ratio, nsub),
                   "cells": cells, "precision": "fp64"},
        "column_variant": {"variant": col_variant, "best_ms": col_ms},
        "roofline": {"bound": "hbm", "kernel": kname.value.decode(), "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kname.value.decode(), kbytes.value),
                     "algorithmic_bytes": kbytes.value, "avg_launch_ms": avg_ms, "launches": kcnt.value},
        "wall_s": wall,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cfg.nx, cfg.ny, cfg.nz, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
