"""GPU parity per hook: every C-ABI hook (include/erfx.h) against the oracle's
restatement of the same ERF function, from identical inputs.

The inputs are a mid-run state: the oracle advances one step (so every
integrator buffer holds realistic data), then every buffer is copied into the
sim, and the hook runs on both.  With the EOS on the shared reproducible pow
and no FMA contraction on either side, the bar is bit-for-bit identity of
every buffer, ghost cells included (index and ghost-cell fills exact).
"""
import ctypes

import numpy as np
import pytest

from erf_amd import config as C, problems as P
from helpers import ALL_IDS
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _setup(kind):
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import Sim
    if kind == "thermal":
        cfg, f = P.rising_thermal(24, 20, 16, seed=3)
        dt = 1.0
    elif kind == "cbl":
        cfg, f = P.cbl(20, 18, 16, seed=5)
        dt = 0.5
    elif kind.startswith("moist"):   # Kessler, buoyancy type moist1..moist4
        cfg, f = P.moist_bubble(24, 20, 16, seed=13, buoyancy_type=int(kind[-1]))
        dt = 1.0
    elif kind == "norain":
        cfg, f = P.moist_bubble(24, 20, 16, seed=17, norain=True)
        dt = 1.0
    elif kind.startswith("inflow"):   # ERF's Inflow on x-lo, Outflow on x-hi
        cfg, f = P.inflow_outflow(24, 16, 16, seed=19, moist=(kind == "inflow_moist"))
        if kind == "inflow_diff":     # with Constant molecular diffusion
            cfg.molec_diff_type = C.MOLEC_CONSTANT
            cfg.dynamic_viscosity, cfg.alpha_T, cfg.alpha_C = 60.0, 40.0, 50.0
        dt = 1.0
    elif kind == "noslip":   # no-slip walls, the top moving; Constant diffusion
        cfg, f = P.rising_thermal(24, 16, 16, seed=29, adv=C.ADV_UPWIND3)
        cfg.bc[2] = cfg.bc[5] = C.BC_NO_SLIP_WALL
        cfg.wall_velocity[5][0], cfg.wall_velocity[5][1] = 3.0, -1.5
        cfg.molec_diff_type = C.MOLEC_CONSTANT
        cfg.dynamic_viscosity, cfg.alpha_T = 40.0, 20.0
        dt = 1.0
    elif kind.startswith("hill"):   # terrain-following grid over a Gaussian hill
        adv = C.ADV_CENTERED2 if kind == "hill_c2" else C.ADV_UPWIND3
        cfg, f = P.hill_flow(24, 20, 16, seed=7, adv=adv, moist=(kind == "hill_moist"))
        if kind == "hill_diff":       # ConstantAlpha molecular diffusion on the terrain grid
            cfg.molec_diff_type = C.MOLEC_CONSTANT_ALPHA
            cfg.rho0_trans, cfg.dynamic_viscosity, cfg.alpha_T, cfg.alpha_C = 1.0, 60.0, 45.0, 30.0
        elif kind == "hill_noslip":   # Constant diffusion, no-slip walls, the top moving
            cfg.bc[2] = cfg.bc[5] = C.BC_NO_SLIP_WALL
            cfg.wall_velocity[5][0], cfg.wall_velocity[5][1] = 3.0, -1.5
            cfg.molec_diff_type = C.MOLEC_CONSTANT
            cfg.dynamic_viscosity, cfg.alpha_T, cfg.alpha_C = 40.0, 20.0, 10.0
        dt = 1.0
    else:  # walls + molecular diffusion
        cfg, f = P.rising_thermal(24, 8, 16, seed=11, adv=C.ADV_UPWIND3)
        cfg.bc[0] = C.BC_SYMMETRY
        cfg.bc[3] = C.BC_OUTFLOW
        cfg.molec_diff_type = C.MOLEC_CONSTANT_ALPHA
        cfg.dynamic_viscosity = 75.0
        cfg.alpha_T = 75.0
        dt = 1.0
    orc = Oracle(cfg)
    orc.load(f)
    orc.step(0.0, dt, 1)
    sim = Sim(cfg)
    ids = list(ALL_IDS) + ([C.D_EDDY] if cfg.les_type != C.LES_NONE else [])
    ids += [C.D_RAIN_ACCUM] if cfg.moisture_type != C.MOIST_NONE else []
    ids += TERRAIN_IDS if cfg.use_terrain else []
    for fid in ids:
        sim.upload(fid, orc.get(fid))
    return cfg, orc, sim, ids, dt


def _same(sim, orc, ids):
    """ids whose buffers differ anywhere (NaN == NaN), with the count of points."""
    bad = {}
    for fid in ids:
        a = sim.download(fid)
        b = orc.get(fid)
        eq = (a == b) | (np.isnan(a) & np.isnan(b))
        if not eq.all():
            d = np.abs(np.where(eq, 0.0, a - b))
            bad[fid] = (int((~eq).sum()), float(np.nanmax(d)))
    return bad


def _st(sim, s):
    return ctypes.byref(sim.state(s))


TERRAIN_IDS = [C.F_Z_ND, C.D_DETJ, C.D_AX, C.D_AY, C.D_AZ, C.D_Z_CC]
HILLS = ["hill", "hill_c2", "hill_moist", "hill_diff", "hill_noslip"]
KINDS = ["thermal", "cbl", "walls", "moist1", "moist2", "norain", "inflow", "inflow_moist", "inflow_diff",
         "noslip"] + HILLS
MOIST_BUOYANCY = ["moist1", "moist2", "moist3", "moist4"]


@pytest.mark.parametrize("kind", KINDS)
def test_cons_to_prim(kind):
    from erf_amd.lib import lib, check
    cfg, orc, sim, ids, dt = _setup(kind)
    ng = sim.L.ng
    sid = C.state_id(C.S_NEW, 0)
    orc.call("orc_cons_to_prim", sid, ng)
    check(lib().erfx_cons_to_prim(sim.ctx, sim.ptr(sid), sim.ptr(C.D_PRIM), sim.ptr(C.D_PI_STAGE), ng))
    sim.sync()
    assert _same(sim, orc, ids) == {}


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("which,fast_only,synced", [(C.S_SUM, 1, 0), (C.S_NEW, 0, 0), (C.S_OLD, 0, 1)])
def test_apply_bcs(kind, which, fast_only, synced):
    from erf_amd.lib import lib, check
    cfg, orc, sim, ids, dt = _setup(kind)
    ngc, ngv = (1, 1) if fast_only else (sim.L.ng, sim.L.ng - 1)
    orc.call("orc_apply_bcs", which, ngc, ngv, fast_only, synced)
    check(lib().erfx_apply_bcs(sim.ctx, _st(sim, which), sim.ptr(C.F_XVEL), sim.ptr(C.F_YVEL),
                               sim.ptr(C.F_ZVEL), ngc, ngv, fast_only, synced))
    sim.sync()
    assert _same(sim, orc, ids) == {}


@pytest.mark.parametrize("kind", KINDS)
def test_make_fast_coeffs(kind):
    from erf_amd.lib import lib, check
    cfg, orc, sim, ids, dt = _setup(kind)
    dtau = dt / cfg.mri_ratio
    orc.call("orc_make_fast_coeffs", dtau, 0.1)
    check(lib().erfx_make_fast_coeffs(sim.ctx, sim.ptr(C.D_FAST_COEFFS), _st(sim, C.S_NEW), sim.ptr(C.D_PRIM),
                                      sim.ptr(C.D_PI_STAGE), dtau, 0.1))
    sim.sync()
    assert _same(sim, orc, ids) == {}


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("step,nrk", [(0, 0), (0, 2), (3, 2)])
def test_fast_rhs_N(kind, step, nrk):
    from erf_amd.lib import lib, check
    cfg, orc, sim, ids, dt = _setup(kind)
    dtau = dt / cfg.mri_ratio
    nsub = cfg.mri_ratio
    beta_s = 0.1 if cfg.implicit_substepping else -1.0
    T = "T" if cfg.use_terrain else "N"       # erf_fast_rhs_T on a terrain grid
    orc.call("orc_fast_rhs_" + T, step, nrk, dtau, beta_s, 1.0 / nsub)
    prev = C.S_OLD if step == 0 else C.S_SUM
    check(getattr(lib(), "erfx_fast_rhs_" + T)(sim.ctx, step, nrk, _st(sim, C.F_SLOW), _st(sim, prev), _st(sim, C.S_NEW),
                                sim.ptr(C.D_PRIM), sim.ptr(C.D_PI_STAGE), sim.ptr(C.D_FAST_COEFFS),
                                _st(sim, C.S_SUM), _st(sim, C.S_SCRATCH), dtau, beta_s, 1.0 / nsub))
    sim.sync()
    assert _same(sim, orc, ids) == {}


@pytest.mark.parametrize("kind", ["cbl"])
def test_turbulent_viscosity(kind):
    from erf_amd.lib import lib, check
    cfg, orc, sim, ids, dt = _setup(kind)
    orc.call("orc_turbulent_viscosity")
    check(lib().erfx_turbulent_viscosity(sim.ctx, sim.ptr(C.state_id(C.S_OLD, 0)), sim.ptr(C.D_XVEL_OLD),
                                         sim.ptr(C.D_YVEL_OLD), sim.ptr(C.D_ZVEL_OLD), sim.ptr(C.D_EDDY), None))
    sim.sync()
    assert _same(sim, orc, ids) == {}


@pytest.mark.parametrize("kind", KINDS + ["moist3", "moist4"])
@pytest.mark.parametrize("nrk", [0, 2])
def test_slow_rhs_pre(kind, nrk):
    from erf_amd.lib import lib, check
    cfg, orc, sim, ids, dt = _setup(kind)
    slow_dt = dt / 3.0 if nrk == 0 else dt
    orc.call("orc_slow_rhs_pre", nrk, slow_dt)
    eddy = sim.ptr(C.D_EDDY) if cfg.les_type != C.LES_NONE else None
    check(lib().erfx_slow_rhs_pre(sim.ctx, nrk, slow_dt, _st(sim, C.F_SLOW), _st(sim, C.S_NEW), sim.ptr(C.D_PRIM),
                                  _st(sim, C.S_SCRATCH), sim.ptr(C.F_XVEL), sim.ptr(C.F_YVEL), sim.ptr(C.F_ZVEL),
                                  eddy, None))
    sim.sync()
    assert _same(sim, orc, ids) == {}


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("nrk", [0, 2])
def test_slow_rhs_post(kind, nrk):
    from erf_amd.lib import lib, check
    cfg, orc, sim, ids, dt = _setup(kind)
    slow_dt = dt / 3.0 if nrk == 0 else dt
    orc.call("orc_slow_rhs_post", nrk, slow_dt)
    eddy = sim.ptr(C.D_EDDY) if cfg.les_type != C.LES_NONE else None
    check(lib().erfx_slow_rhs_post(sim.ctx, nrk, slow_dt, _st(sim, C.F_SLOW), _st(sim, C.S_OLD), _st(sim, C.S_NEW),
                                   _st(sim, C.S_SUM), sim.ptr(C.D_PRIM), _st(sim, C.S_SCRATCH), sim.ptr(C.F_XVEL),
                                   sim.ptr(C.F_YVEL), eddy))
    sim.sync()
    assert _same(sim, orc, ids) == {}


@pytest.mark.parametrize("kind", ["moist1", "norain", "hill_moist"])
@pytest.mark.parametrize("dt", [1.0, 3.0])
def test_kessler(kind, dt):
    from erf_amd.lib import lib, check
    cfg, orc, sim, ids, _ = _setup(kind)
    orc.kessler(dt)
    check(lib().erfx_kessler_advance(sim.ctx, sim.ptr(C.F_CONS), dt, sim.ptr(C.D_RAIN_ACCUM)))
    assert _same(sim, orc, ids) == {}


@pytest.mark.parametrize("h0", [0.0, 500.0])
def test_terrain_metrics(h0):
    """make_J / make_areas / make_zcc from uploaded node heights (the whole
    layout, padding included) and the dz_min of the time-step estimate."""
    import torch
    assert torch.cuda.is_available()
    from erf_amd.lib import Sim
    cfg, f = P.hill_flow(24, 20, 16, h0=h0)
    orc = Oracle(cfg)
    orc.load(f)
    sim = Sim(cfg)
    sim.upload(C.F_Z_ND, f[C.F_Z_ND])
    assert _same(sim, orc, TERRAIN_IDS) == {}


def test_fast_rhs_N_rejects_terrain():
    from erf_amd.lib import lib
    cfg, orc, sim, ids, dt = _setup("hill")
    rc = lib().erfx_fast_rhs_N(sim.ctx, 0, 0, _st(sim, C.F_SLOW), _st(sim, C.S_OLD), _st(sim, C.S_NEW),
                               sim.ptr(C.D_PRIM), sim.ptr(C.D_PI_STAGE), sim.ptr(C.D_FAST_COEFFS),
                               _st(sim, C.S_SUM), _st(sim, C.S_SCRATCH), 1.0 / 6, 0.1, 1.0)
    assert rc != 0
