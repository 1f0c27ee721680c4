"""The oracle against six more of the reference's regression gold files
(Tests/ERFGoldFiles/<case>, fcompare rtol = atol = 2e-10,
Tests/CMakeLists.txt:3-4).  Each case restates its input deck
(Tests/test_files/<case>/<case>.i) as an erfx_config plus the problem's
initial state (oracle/init.cpp) and ERF's time loop with its dt control
(orc_run: ComputeDt / estTimeStep before every step).

  ScalarDiffusionSine, ScalarDiffusionGaussian  Constant diffusion of a scalar (fixed dt and ratio)
  ScalarAdvectionUniformU                      C2 scalar advection, uniform wind, CFL dt
  ScalarAdvectionShearedU                      U3 advection in a log-profile wind, CFL dt
  TaylorGreenAdvecting(Diffusing)              momentum advection (and Constant diffusion), fixed dt, CFL ratio

The fixtures are the full gold arrays (tests/golden/make_gold_fixtures.py).
"""
import ctypes
import math
import os

import numpy as np
import pytest

from erf_amd import config as C
from helpers import valid
from oracle_lib import Oracle, lib as olib

HERE = os.path.dirname(os.path.abspath(__file__))
TWO_PI = 6.2831853071795864


def _cfg(n, extent, adv=C.ADV_UPWIND3, scal=C.ADV_UPWIND3):
    cfg = C.default_config(n[0], n[1], n[2], extent[0] / n[0], extent[1] / n[1], extent[2] / n[2])
    cfg.bc[0] = cfg.bc[3] = C.BC_PERIODIC
    cfg.bc[1] = cfg.bc[4] = C.BC_PERIODIC
    cfg.bc[2] = cfg.bc[5] = C.BC_SLIP_WALL
    cfg.use_gravity = 0
    cfg.dycore_horiz_adv = cfg.dycore_vert_adv = adv
    cfg.dryscal_horiz_adv = cfg.dryscal_vert_adv = scal
    cfg.c_p = 1004.5
    return cfg


# ScalarAdvDiff problem parameters (Exec/RegTests/ScalarAdvDiff/ERF_prob.H:11-27 defaults)
def _sad_params(**kw):
    p = dict(rho_0=0.0, T_0=1.0, A_0=1.0, B_0=0.0, u_0=0.0, v_0=0.0, w_0=0.0, rad_0=0.0, z0=0.1, zRef=80.0,
             uRef=0.0, xc_frac=0.5, yc_frac=0.5, zc_frac=0.5, prob_type=-1)
    p.update(kw)
    keys = ["rho_0", "T_0", "A_0", "B_0", "u_0", "v_0", "w_0", "rad_0", "z0", "zRef", "uRef", "xc_frac", "yc_frac",
            "zc_frac", "prob_type"]
    return np.array([float(p[k]) for k in keys])


def _case(name):
    """config, initialiser, run arguments (cfl, fixed_dt, fixed_fast_dt, fixed_mri_ratio) and steps."""
    if name == "ScalarDiffusionSine":
        cfg = _cfg((16, 16, 4), (TWO_PI, TWO_PI, TWO_PI))
        cfg.molec_diff_type = C.MOLEC_CONSTANT
        cfg.alpha_T, cfg.alpha_C, cfg.rho0_trans, cfg.dynamic_viscosity = 0.0, 1.0, 1.0, 0.0
        init = ("sad", _sad_params(rho_0=1.0, A_0=0.0, B_0=1.0, u_0=0.0, v_0=0.0, uRef=0.0, prob_type=10))
        return cfg, init, (0.8, 1e-2, -1.0, 4), 20
    if name == "ScalarDiffusionGaussian":
        cfg = _cfg((16, 16, 16), (1.0, 1.0, 1.0))
        cfg.molec_diff_type = C.MOLEC_CONSTANT
        cfg.alpha_T, cfg.alpha_C, cfg.rho0_trans, cfg.dynamic_viscosity = 0.0, 1.0, 1.0, 0.0
        init = ("sad", _sad_params(rho_0=1.0, A_0=1.0, B_0=0.0, u_0=0.0, v_0=0.0, uRef=0.0, prob_type=10))
        return cfg, init, (0.8, 1e-4, -1.0, 2), 20
    if name == "ScalarAdvectionUniformU":
        cfg = _cfg((64, 64, 4), (1.0, 1.0, 1.0), scal=C.ADV_CENTERED2)
        init = ("sad", _sad_params(rho_0=1.0, T_0=1.0, A_0=1.0, u_0=10.0, v_0=5.0, rad_0=0.125, uRef=0.0,
                                   prob_type=11))
        return cfg, init, (0.9, -1.0, -1.0, 0), 20
    if name == "ScalarAdvectionShearedU":
        cfg = _cfg((64, 4, 64), (8.0, 8.0, 8.0))
        init = ("sad", _sad_params(rho_0=1.0, T_0=1.0, A_0=1.0, u_0=0.0, v_0=0.0, rad_0=0.25, z0=0.1, zRef=80.0,
                                   uRef=8.0, prob_type=12))
        return cfg, init, (0.9, -1.0, -1.0, 0), 80
    if name in ("TaylorGreenAdvecting", "TaylorGreenAdvectingDiffusing"):
        ext = 6.283185307179586476925
        cfg = _cfg((16, 16, 16), (ext, ext, ext))
        if name.endswith("Diffusing"):
            cfg.molec_diff_type = C.MOLEC_CONSTANT
            cfg.dynamic_viscosity = 6.25e-4
        cfg.alpha_T = cfg.alpha_C = 0.0
        # the deck's "erf.mri_fixed_dt_ratio" is not a key ERF reads: the ratio comes from estTimeStep
        return cfg, ("tg", (1.0, 1.0, 1.0)), (0.8, 0.16, -1.0, 0), 10
    raise KeyError(name)


CASES = ["ScalarDiffusionSine", "ScalarDiffusionGaussian", "ScalarAdvectionUniformU", "ScalarAdvectionShearedU",
         "TaylorGreenAdvecting", "TaylorGreenAdvectingDiffusing"]


def run_case(name):
    cfg, init, run, nsteps = _case(name)
    orc = Oracle(cfg)
    l = olib()
    if init[0] == "sad":
        l.orc_init_scalar_advdiff.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        params = np.ascontiguousarray(init[1])
        orc._chk(l.orc_init_scalar_advdiff(orc.h, params.ctypes.data))
    else:
        l.orc_init_taylor_green.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3
        orc._chk(l.orc_init_taylor_green(orc.h, *init[1]))
    l.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                          ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                          ctypes.POINTER(ctypes.c_double)]
    t_end = ctypes.c_double()
    cfl, fdt, ffdt, fratio = run
    orc._chk(l.orc_run(orc.h, 0.0, nsteps, cfl, fdt, ffdt, fratio, 1.1, 1.0, 1.e100, ctypes.byref(t_end)))
    return cfg, orc, t_end.value


def plot_vars(cfg, orc):
    L = orc.L
    cons = valid(orc.get(C.F_CONS), cfg, L, C.F_CONS)
    u = valid(orc.get(C.F_XVEL), cfg, L, C.F_XVEL)[0]
    v = valid(orc.get(C.F_YVEL), cfg, L, C.F_YVEL)[0]
    w = valid(orc.get(C.F_ZVEL), cfg, L, C.F_ZVEL)[0]
    rho, rt = cons[C.RHO], cons[C.RHOTHETA]
    p = C.p_0 * np.power(C.R_d * rt * (1.0 / C.p_0), C.Gamma)
    return {
        "density": rho, "rhoadv_0": cons[C.RHOSCALAR], "scalar": cons[C.RHOSCALAR] / rho,
        "x_velocity": 0.5 * (u[:, :, :-1] + u[:, :, 1:]),
        "y_velocity": 0.5 * (v[:, :-1, :] + v[:, 1:, :]),
        "z_velocity": 0.5 * (w[:-1] + w[1:]),
        "theta": rt / rho, "pressure": p, "temp": p / (C.R_d * rho),
    }


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_gold(name):
    z = np.load(os.path.join(HERE, "golden", name + ".npz"))
    cfg, orc, t_end = run_case(name)
    assert abs(t_end - float(z["time"])) <= 1e-12 * max(1.0, abs(float(z["time"]))), (t_end, float(z["time"]))
    mine = plot_vars(cfg, orc)
    report = {}
    for nm, g in zip(list(z["names"]), z["data"]):
        a = mine[nm]
        aerr = float(np.abs(a - g).max())
        scale = float(np.abs(g).max())
        report[nm] = (aerr, aerr / scale if scale > 0 else aerr)
    bad = {k: v for k, v in report.items() if v[0] > 2e-10 and v[1] > 2e-10}
    assert not bad, "beyond fcompare tolerance: %s\nall: %s" % (bad, report)
