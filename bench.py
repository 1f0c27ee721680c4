#!/usr/bin/env python3
"""Benchmark of the MI355X dycore step (one GPU).

Default workload (the headline): moist flow over terrain with Kessler warm
rain on the terrain-following path, 512 x 512 x 96 cells (dx = dy = 200 m,
levels stretched from 20 m to 400 m up to a 15 km top, 500 m Gaussian hill),
WENO5 scalars, 3rd-order upwind momentum, Rayleigh damping above 12 km,
dt = 0.5 s (the recipe's 1 s is unstable on the 20 m surface cells, in the
oracle as on the GPU), MRI ratio 6 (1 + 3 + 6 acoustic substeps per step).  Synthetic
seeded initial state.  A step is one full RK3 step (slow RHS, acoustic
substeps, boundary fills, microphysics) with the state resident in HBM.

The FP64 path and the mixed-precision path are timed on the same state, and
the mixed path's error against FP64 is measured after --check-steps steps
(100 by default) from the same initial state: relative L2 and Linf of rho,
rho*theta, u, v, w, qv, qc and qr.  One JSON line carries cell-updates/s of
both paths, the fraction of the HBM peak each reaches on the algorithmic bytes
of a step, the roofline of the dominant kernel (HIP events on the library's
stream around each launch) and the CPU oracle timed on a bounded sample.

Other workloads (--workload): cbl, thermal, ensemble, supercell.
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s per RK3 step, mixed vs FP64 GPU at stated L2 error; % HBM BW"


def measured_traffic(kernel, workload, bytes_per_launch, variant=None):
    """HBM bytes per launch of the timed kernel from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement (profiles/<kernel>_<workload>_traffic.json),
    if it was taken on this kernel and workload (same algorithmic byte count)
    and, where the file names one, on the same column-kernel variant as the
    timed run; otherwise None."""
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.json" % (kernel.split(" ")[0], workload))
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if abs(t.get("algorithmic_bytes_per_launch", -1) - bytes_per_launch) > 0.5:
        return None
    if t.get("column_variant") is not None and variant is not None and t["column_variant"] != variant:
        return None
    return t["traffic_bytes_per_launch"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--workload", choices=["moist_terrain", "cbl", "thermal", "ensemble", "supercell"],
                    default="moist_terrain",
                    help="moist_terrain (default): the headline, moist flow over a hill on the terrain-following "
                         "path, FP64 and mixed precision; cbl: the dry CBL LES 256x256x128; thermal: the 64^3 dry "
                         "rising thermal; ensemble: dry CBL members of uneven box sizes advanced as one batch; "
                         "supercell: Weisman-Klemp deep convection, the learned Kessler stand-in trained on the "
                         "plain method's runs and run against plain Kessler")
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default=None,
                    help="paths to time (default: both for moist_terrain, fp64 otherwise).  With mixed or both the "
                         "mixed path's error against FP64 is measured after --check-steps steps")
    ap.add_argument("--check-steps", type=int, default=100,
                    help="steps from the initial state after which the mixed path is compared with FP64 "
                         "(0: no comparison)")
    ap.add_argument("--sc-seeds", type=int, default=2, help="supercell: training runs (seeds 0..n-1)")
    ap.add_argument("--sc-steps", type=int, default=600, help="supercell: steps of each training and evaluation run")
    ap.add_argument("--sc-every", type=int, default=2, help="supercell: rows harvested every n-th step")
    ap.add_argument("--sc-rows", type=int, default=4_000_000, help="supercell: row cap per training run")
    ap.add_argument("--sc-fit", type=int, default=3000, help="supercell: Adam steps")
    ap.add_argument("--sc-per-step", type=int, default=0,
                    help="supercell: rows drawn at random from each harvested step (0: all, up to --sc-rows)")
    ap.add_argument("--members", type=int, default=48, help="ensemble members (--workload ensemble)")
    ap.add_argument("--ens-streams", type=int, default=2, help="ensemble: member groups, one stream each")
    ap.add_argument("--ens-box", type=int, nargs=3, default=None, metavar=("NX", "NY", "NZ"),
                    help="ensemble: every member of this one box size instead of the five drawn sizes")
    ap.add_argument("--unbatched-steps", type=int, default=2,
                    help="ensemble: steps timed with every member launched on its own, for comparison")
    ap.add_argument("--graph", action="store_true", help="replay the steps from a captured hipGraph")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default acoustic column kernel instead of timing its variants in the warm-up "
                         "(flat grids; the terrain column kernel has one form)")
    ap.add_argument("--cpu-seconds", type=float, default=6.0, help="CPU oracle: seconds per timed run (3 runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB (3 launches), the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counter readings")
    a = ap.parse_args()
    if a.precision is None:
        a.precision = "both" if a.workload == "moist_terrain" else "fp64"
    return a


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


# ---- workloads -----------------------------------------------------------------
def make_case(workload, nx=None, ny=None, nz=None, seed=0):
    """(cfg, fields, rayleigh profiles or None, dt, description) of a single-box workload."""
    from erf_amd import problems as P
    if workload == "moist_terrain":
        n = (nx or 512, ny or 512, nz or 96)
        cfg, f = P.moist_terrain(*n, seed=seed)
        ray = P.rayleigh_profiles_uniform(cfg, f=f)
        desc = ("moist flow over a 500 m Gaussian hill %dx%dx%d, dx=200 m, levels stretched 20-400 m to a 15 km top, "
                "WENO5 scalars, U3 momentum, Kessler, Rayleigh above 12 km, dt=0.5 s (1 s is unstable on the 20 m "
                "surface cells), MRI ratio %d (%d substeps/step)")
        return cfg, f, ray, P.MOIST_TERRAIN_DT, desc
    if workload == "thermal":
        n = (nx or 64, ny or 64, nz or 64)
        cfg, f = P.rising_thermal(*n, seed=seed)
        desc = "dry rising thermal %dx%dx%d, C2 advection, dt=1 s, MRI ratio %d (%d substeps/step)"
        return cfg, f, None, 1.0, desc
    n = (nx or 256, ny or 256, nz or 128)
    cfg, f = P.cbl(*n, seed=seed)
    desc = ("dry CBL LES %dx%dx%d, U5 advection, Smagorinsky, Coriolis+geostrophic, dt=0.5 s, "
            "MRI ratio %d (%d substeps/step)")
    return cfg, f, None, 0.5, desc


def n_substeps(cfg):
    return 1 + cfg.mri_ratio // 2 + cfg.mri_ratio


def alg_bytes_per_cell(cfg):
    """Algorithmic HBM bytes per cell and RK3 step: each distinct 3-D FP64 field
    read or written once per sweep (slow RHS per stage, acoustic substep, slow
    post per stage, once per step), intermediates, velocities and 1-D data not
    counted: 8 B x (3 F_pre + N_sub F_sub + 3 F_post + F_once)."""
    from erf_amd import config as C
    moist = cfg.moisture_type != C.MOIST_NONE
    smag = cfg.les_type == C.LES_SMAGORINSKY
    terr = cfg.use_terrain != 0
    f_pre = 10 + 2 * moist + smag + 4 * terr
    f_sub = 28 + 2 * moist + 4 * terr
    n_s = 4 if moist else 1
    f_post = 3 * n_s + 4 + smag + terr
    f_once = 5 * smag + 9 * moist
    return 8 * (3 * f_pre + n_substeps(cfg) * f_sub + 3 * f_post + f_once)


def time_leg(sim, dt, steps, warmup, use_timer, autotune, graph):
    """Warm-up, then `steps` steps timed with HIP events on the sim's stream.
    Returns (ms per step, timer reading or None)."""
    import torch
    from erf_amd.lib import lib, check
    stream = torch.cuda.ExternalStream(sim.stream)
    if graph:
        sim.set_graph(True)
    if autotune and warmup > 0:
        sim.autotune(True)       # the warm-up steps time the column kernel's variants
    sim.step(0.0, dt, warmup)
    t = warmup * dt
    variant, _ = sim.column_variant()
    extra = 0
    while variant < 0 and extra < 5:     # short runs: finish timing the variants before the timed steps
        sim.step(t, dt, 1)
        t += dt
        extra += 1
        variant, _ = sim.column_variant()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    if use_timer:
        check(lib().erfx_sim_timer_reset(sim.h))
    with torch.cuda.stream(stream):
        e0.record(stream)
    sim.step_async(t, dt, steps)
    with torch.cuda.stream(stream):
        e1.record(stream)
    sim.sync()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    timer = None
    if use_timer:
        kname = ctypes.create_string_buffer(128)
        kms, kcnt, kbytes = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
        check(lib().erfx_sim_timer_read(sim.h, kname, 128, ctypes.byref(kms), ctypes.byref(kcnt),
                                        ctypes.byref(kbytes)))
        timer = {"kernel": kname.value.decode(), "total_ms": kms.value, "launches": kcnt.value,
                 "bytes_per_launch": kbytes.value}
        check(lib().erfx_sim_timer_stop(sim.h))
    return ms, timer, variant


def check_state(sim, cfg):
    from erf_amd import config as C
    import numpy as np
    ng = sim.L.ng
    rho = sim.download(C.F_CONS)[0][ng:ng + cfg.nz, ng:ng + cfg.ny, ng:ng + cfg.nx]
    assert np.all(np.isfinite(rho)) and rho.min() > 0.05, "state blew up"


def errors_vs(sim, ref, cfg):
    """Relative L2 and Linf of the difference sim - ref over the valid cells:
    rho and rho*theta relative to their deviation from the mean, velocities and
    mixing ratios relative to the field itself."""
    import numpy as np
    from erf_amd import config as C
    ng = sim.L.ng
    fields = [(C.F_CONS, C.RHO, "rho"), (C.F_CONS, C.RHOTHETA, "rhotheta"), (C.F_XVEL, 0, "u"),
              (C.F_YVEL, 0, "v"), (C.F_ZVEL, 0, "w")]
    if cfg.moisture_type:
        fields += [(C.F_CONS, C.RHOQ1, "qv"), (C.F_CONS, C.RHOQ2, "qc"), (C.F_CONS, C.RHOQ3, "qr")]
    got = {}
    out = {}
    sl = (slice(ng, ng + cfg.nz), slice(ng, ng + cfg.ny), slice(ng, ng + cfg.nx))
    for fid, comp, name in fields:
        if fid not in got:
            got[fid] = (sim.download(fid), ref.download(fid))
        x = got[fid][0][comp][sl]
        y = got[fid][1][comp][sl]
        if fid == C.F_CONS and comp >= C.RHOQ1:      # mixing ratios q = rho q / rho
            x = x / got[fid][0][C.RHO][sl]
            y = y / got[fid][1][C.RHO][sl]
        r = y - y.mean() if name in ("rho", "rhotheta") else y
        nrm = float(np.sqrt((r ** 2).mean()))
        amax = float(np.abs(r).max())
        d = x - y
        out[name] = {"rel_l2": float(np.sqrt((d ** 2).mean())) / nrm if nrm > 0 else float(np.abs(d).max()),
                     "rel_linf": float(np.abs(d).max()) / amax if amax > 0 else float(np.abs(d).max())}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, nz, budget_s):
    """The CPU oracle (test infrastructure: C++ restatement of the reference,
    OpenMP over k-planes, g++ -O3 -march=x86-64-v3 -ffp-contract=off) on a bounded
    sample of the same workload: a 64 x 64 x nz block of the same case, as many
    steps as fit in ~budget_s, three timed runs, the median reported."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    cfg, f, ray, dt, _ = make_case(workload, 64, 64, nz, seed=0)
    orc = Oracle(cfg)
    orc.load(f)
    if ray is not None:
        orc.set_rayleigh(*ray)
    orc.step(0.0, dt, 1)                      # warm-up
    t = dt
    rates = []
    steps_per_run = []
    for _ in range(3):
        t0 = time.perf_counter()
        n = 0
        while True:
            orc.step(t, dt, 1)
            t += dt
            n += 1
            if time.perf_counter() - t0 > budget_s:
                break
        el = time.perf_counter() - t0
        rates.append(cfg.nx * cfg.ny * cfg.nz * n / el)
        steps_per_run.append(n)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": statistics.median(rates), "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "runs": rates,
            "sample": "median of 3 runs of %s steps of the same case on a 64x64x%d block (the oracle: C++ "
                      "restatement of the reference, OpenMP, -O3 -march=x86-64-v3)"
                      % ("/".join(str(s) for s in steps_per_run), nz)}


def run_single(a):
    """A single-box workload: the FP64 and / or mixed-precision path timed on the
    same state, the mixed path's error against FP64 after --check-steps steps."""
    import torch
    from erf_amd import config as C
    from erf_amd.lib import Sim

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    print("building the %s case" % a.workload, file=sys.stderr, flush=True)
    cfg, f, ray, dt, desc = make_case(a.workload, a.nx, a.ny, a.nz, seed=0)
    legs = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    cells = cfg.nx * cfg.ny * cfg.nz
    b_cell = alg_bytes_per_cell(cfg)

    def make_sim(prec):
        c = C.Config.from_buffer_copy(cfg)
        c.precision = C.PREC_MIXED if prec == "mixed" else C.PREC_FP64
        s_ = Sim(c)
        s_.load(f)
        if ray is not None:
            s_.set_rayleigh(*ray)
        return s_

    sims = {}
    res = {}
    timer = None
    variant = None
    for leg in legs:
        sims[leg] = make_sim(leg)
        print("%s state uploaded" % leg, file=sys.stderr, flush=True)
        ms, tm, var = time_leg(sims[leg], dt, a.steps, a.warmup, use_timer=(leg == legs[0] and not a.graph),
                               autotune=not a.no_autotune, graph=a.graph)
        check_state(sims[leg], cfg)
        if tm is not None:
            timer = tm
        if leg == legs[0]:
            variant = var
        step_gbs = b_cell * cells / (ms * 1e-3) / 1e9
        res[leg] = {"value": cells / (ms * 1e-3), "ms_per_step": ms,
                    "step_hbm": {"bytes_per_step": b_cell * cells, "bytes_per_cell": b_cell,
                                 "achieved_GBs": step_gbs, "frac": step_gbs / HBM_PEAK_GBS}}
        print("%s: %.2f ms/step" % (leg, ms), file=sys.stderr, flush=True)

    errors = None
    if "mixed" in legs and a.check_steps > 0:
        # both paths from the same initial state, check_steps steps each
        ref = sims.get("fp64") or make_sim("fp64")
        sims["fp64"] = ref
        for leg in ("fp64", "mixed"):
            sims[leg].load(f)
            sims[leg].step(0.0, dt, a.check_steps)
            print("%s: %d checking steps done" % (leg, a.check_steps), file=sys.stderr, flush=True)
        errors = {"steps": a.check_steps, "fields": errors_vs(sims["mixed"], ref, cfg)}
    for s_ in sims.values():
        s_.close()

    head = res[legs[-1]]
    precision = "mixed vs fp64" if len(legs) == 2 else legs[0]
    out = {
        "metric": METRIC if a.workload == "moist_terrain" else
        "cell-updates/s per RK3 step (%s GPU path)" % ("mixed" if legs[-1] == "mixed" else "FP64"),
        "value": head["value"], "unit": "cell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if legs[-1] == "mixed" else "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": desc % (cfg.nx, cfg.ny, cfg.nz, cfg.mri_ratio, n_substeps(cfg)),
                   "cells": cells, "precision": precision, "graph": bool(a.graph),
                   "value_is": "the %s path" % legs[-1]},
    }
    for leg in legs:
        out[leg] = res[leg]
    if len(legs) == 2:
        out["mixed_over_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    if errors is not None:
        out["error_vs_fp64"] = errors
    if timer is not None and timer["launches"] > 0:
        avg_ms = timer["total_ms"] / timer["launches"]
        achieved = timer["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": timer["kernel"], "path": legs[0], "achieved": achieved,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                           "traffic": measured_traffic(timer["kernel"], a.workload, timer["bytes_per_launch"],
                                                       None if cfg.use_terrain else variant),
                           "algorithmic_bytes": timer["bytes_per_launch"], "avg_launch_ms": avg_ms,
                           "launches": timer["launches"]}
    if variant is not None and not cfg.use_terrain:
        out["column_variant"] = variant
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.workload, cfg.nz, a.cpu_seconds)
    print(json.dumps(out))


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
        warm = max(warm, C.COL_VARIANTS)
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
            if variant >= 0:          # the tuned variant (with --no-autotune: each sim's default)
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
    nx, ny, nz = a.nx or 256, a.ny or 256, a.nz or 80
    dt = 2.0
    case = lambda seed: P.supercell(nx, ny, nz, seed=seed)
    w0 = time.perf_counter()
    log = lambda msg: print(msg, file=sys.stderr, flush=True)
    per = a.sc_per_step or None
    rows = E.harvest(seeds=tuple(range(a.sc_seeds)), nsteps=a.sc_steps, dt=dt, case=case, every=a.sc_every,
                     max_rows=a.sc_rows, per_step=per, log=log)
    held = E.harvest(seeds=(8,), nsteps=a.sc_steps, dt=dt, case=case, every=4 * a.sc_every,
                     max_rows=min(a.sc_rows, 1_000_000), per_step=per, log=log)
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
    log("fit: %d Adam steps on %d rows (%.1f s)" % (a.sc_fit, rows.shape[0], t_fit))
    mse1 = E.mse(net, sc, held)
    n_rows, n_held = rows.shape[0], held.shape[0]
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
        w0 = time.perf_counter()
        s.step(0.0, dt, a.sc_steps)
        torch.cuda.synchronize()
        log("held-out run (seed 8) with %s: %d steps (%.0f s)" % (name, a.sc_steps, time.perf_counter() - w0))
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
        "train_rows": n_rows, "heldout_rows": n_held, "rows_per_harvested_step": a.sc_per_step or None,
        "relative_l2_vs_kessler_after_run": errs,
        "relative_l2_mixed_precision_kessler_vs_fp64_kessler": floor,
        "qc_max_kessler": qc_max,
    }
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: erfx has no multi-GPU path "
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
    run_single(a)


if __name__ == "__main__":
    main()
