"""CPU checks of the oracle's moisture physics (Kessler warm rain and the moist
dycore terms), by what the mathematics and physics fix rather than by
restating the formulas:

* saturation vapour pressure: the Flatau et al. (1992) fits against two
  independent published formulas (Bolton 1980; Murphy & Koop 2005), the
  derivative fit against a finite difference of the pressure fit, and
  q_sat against its closed form eps e / (p - e);
* moist temperature: T (p0/p)^(R_d/c_p) = theta, the Poisson relation the
  moist equation of state reduces to;
* Kessler: per-cell conservation of total water without rain, the latent
  heating balance d(theta) = -(theta/T)(L/c_p) d(q_v), the column water budget
  with sedimentation (what leaves the column is the surface rain), no negative
  species, and rain falling (mass moves down);
* moist dycore: with zero moisture the moist step equals the dry step up to
  the rounding of Kessler's theta round trip.

Tolerances are stated per check.
"""
import math

import numpy as np
import pytest

from erf_amd import config as C, problems as P
from helpers import valid
from oracle_lib import Oracle, saturation, temperature

T0 = 273.16


def test_esatw_against_published_formulas():
    T = np.linspace(T0 - 40.0, T0 + 40.0, 81)
    es, des, _, _ = saturation(T, np.full_like(T, 1000.0))
    tc = T - 273.15
    bolton = 6.112 * np.exp(17.67 * tc / (tc + 243.5))
    # Murphy & Koop (2005) eq. 10, liquid water, Pa -> mbar
    lnp = (54.842763 - 6763.22 / T - 4.210 * np.log(T) + 0.000367 * T +
           np.tanh(0.0415 * (T - 218.8)) * (53.878 - 1331.22 / T - 9.44523 * np.log(T) + 0.014025 * T))
    mk = np.exp(lnp) / 100.0
    warm = T >= T0 - 10.0
    # the fit's own error grows towards -40 C; a wrong digit in any of the nine
    # coefficients gives errors of tens of percent at +-40 C
    assert np.max(np.abs(es / mk - 1.0)) < 1e-2
    assert np.max(np.abs(es[warm] / mk[warm] - 1.0)) < 2e-3
    assert np.max(np.abs(es[warm] / bolton[warm] - 1.0)) < 3e-3
    # a wrong coefficient (e.g. one digit or sign of a1..a8) breaks these bounds
    assert abs(es[40] - 6.11239921) < 1e-12          # dtt = 0 is the leading coefficient


def test_dtesatw_is_the_derivative_of_esatw():
    T = np.linspace(T0 - 40.0, T0 + 40.0, 41)
    h = 1e-3
    ep, _, _, _ = saturation(T + h, np.full_like(T, 1000.0))
    em, _, _, _ = saturation(T - h, np.full_like(T, 1000.0))
    _, des, _, _ = saturation(T, np.full_like(T, 1000.0))
    fd = (ep - em) / (2 * h)
    assert np.max(np.abs(des / fd - 1.0)) < 3e-3


def test_qsatw_closed_form_and_cap():
    T = np.array([250.0, 280.0, 300.0, 320.0])
    p = np.array([500.0, 900.0, 1000.0, 1013.25])
    es, _, qs, dqs = saturation(T, p)
    eps = C.R_d / C.R_v
    np.testing.assert_allclose(qs, eps * es / (p - es), rtol=1e-15)
    _, des, _, _ = saturation(T, p)
    np.testing.assert_allclose(dqs, eps * des / p, rtol=1e-15)
    # max(e, p - e): at p < 2e the denominator is e, so q_sat = eps
    es2, _, qs2, _ = saturation(np.array([340.0]), np.array([100.0]))
    assert es2[0] > 50.0 and qs2[0] == eps


def test_out_of_range_temperature_is_an_error():
    with pytest.raises(RuntimeError):
        saturation(np.array([150.0]), np.array([100.0]))


def test_moist_temperature_obeys_poisson():
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.3, 1.2, 200)
    theta = rng.uniform(280.0, 340.0, 200)
    qv = rng.uniform(0.0, 0.02, 200)
    rt = rho * theta
    T = temperature(rho, rt, qv)
    p = C.p_0 * np.power(C.R_d * rt * (1.0 + C.R_v / C.R_d * qv) / C.p_0, C.Gamma)
    np.testing.assert_allclose(T * np.power(C.p_0 / p, C.R_d / C.Cp_d), theta, rtol=2e-14)


def _kessler_case(nx=6, ny=5, nz=12, seed=0, qp_amp=5e-4, qc_amp=2e-3, rain=True):
    cfg, f = P.moist_bubble(nx, ny, nz, seed=seed, norain=not rain)
    L = C.layout_for(cfg)
    ng = L.ng
    rng = np.random.default_rng(seed)
    cons = f[C.F_CONS]
    sl = (slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    rho = cons[C.RHO][sl]
    cons[C.RHOQ2][sl] = rho * rng.uniform(0.0, qc_amp, rho.shape)
    cons[C.RHOQ3][sl] = rho * rng.uniform(0.0, qp_amp, rho.shape) if rain else 0.0
    return cfg, f, L, sl


def _q(cons, sl):
    rho = cons[C.RHO][sl]
    return rho, cons[C.RHOTHETA][sl] / rho, cons[C.RHOQ1][sl] / rho, cons[C.RHOQ2][sl] / rho, cons[C.RHOQ3][sl] / rho


def test_kessler_norain_conserves_water_and_balances_latent_heat():
    cfg, f, L, sl = _kessler_case(rain=False, qc_amp=8e-4)
    orc = Oracle(cfg)
    orc.load(f)
    before = _q(orc.get(C.F_CONS), sl)
    T_before = temperature(before[0], before[0] * before[1], before[2])
    orc.kessler(1.0)
    after = _q(orc.get(C.F_CONS), sl)
    rho = before[0]
    np.testing.assert_array_equal(after[0], rho)
    tw0 = before[2] + before[3]
    tw1 = after[2] + after[3]
    assert np.max(np.abs(tw1 - tw0)) < 1e-17
    dqv = after[2] - before[2]
    assert np.count_nonzero(np.abs(dqv) > 1e-9) > 10           # condensation and evaporation happened
    dtheta = after[1] - before[1]
    expect = -(before[1] / T_before) * (2.501e6 / cfg.c_p) * dqv
    assert np.max(np.abs(dtheta - expect)) < 1e-12 * np.max(np.abs(expect)) + 1e-12


def _saturate(cfg, f, sl, ratio):
    """Set qv = ratio * q_sat(T, p) in the valid cells (fixed point in qv, since
    T and p depend on it)."""
    cons = f[C.F_CONS]
    rho, rt = cons[C.RHO][sl], cons[C.RHOTHETA][sl]
    qv = cons[C.RHOQ1][sl] / rho
    for _ in range(6):
        T = temperature(rho, rt, qv)
        p = C.p_0 * np.power(C.R_d * rt * (1.0 + C.R_v / C.R_d * qv) / C.p_0, C.Gamma) / 100.0
        _, _, qs, _ = saturation(T.ravel(), p.ravel())
        qv = ratio * qs.reshape(T.shape)
    cons[C.RHOQ1][sl] = rho * qv


def test_kessler_column_budget_equals_surface_rain():
    """Column water changes only by the sedimentation fluxes through the bottom
    and top faces.  The bottom one is the surface rain the model reports; the
    top one is measured the same way on a twin state whose lowest layer holds
    the top layer's rho and rho*q_r (the one-sided face value at k = 0 and at
    k = nz is the same function of the adjacent cell).  ERF's Kessler can take
    cloud water twice in one cell (evaporation and conversion to rain are each
    limited by q_c alone) and clips q_c at 0, which creates water, so the state
    is supersaturated with rain everywhere: nothing is clipped."""
    cfg, f, L, sl = _kessler_case(nx=5, ny=4, nz=14, seed=2)
    ng, nz = L.ng, cfg.nz
    _saturate(cfg, f, sl, 1.03)
    cons = f[C.F_CONS]
    cons[C.RHOQ3][sl] += cons[C.RHO][sl] * 1e-4          # rain in every cell
    twin = {k: v.copy() for k, v in f.items()}
    tc = twin[C.F_CONS]
    top, bot = ng + nz - 1, ng
    theta_bot = cons[C.RHOTHETA][bot] / cons[C.RHO][bot]
    tc[C.RHO][bot] = cons[C.RHO][top]
    tc[C.RHOQ3][bot] = cons[C.RHOQ3][top]
    tc[C.RHOTHETA][bot] = tc[C.RHO][bot] * theta_bot
    dt = 2.0
    orc = Oracle(cfg)
    orc.load(f)
    before = _q(orc.get(C.F_CONS), sl)
    orc.kessler(dt)
    after = _q(orc.get(C.F_CONS), sl)
    rain = orc.get(C.D_RAIN_ACCUM)[ng:ng + cfg.ny, ng:ng + cfg.nx]
    o2 = Oracle(cfg)
    o2.load(twin)
    o2.kessler(dt)
    influx = o2.get(C.D_RAIN_ACCUM)[ng:ng + cfg.ny, ng:ng + cfg.nx]
    for q in after[2:]:
        assert np.all(q > 0)
    rho = before[0]
    w0 = (rho * (before[2] + before[3] + before[4])).sum(axis=0) * cfg.dz
    w1 = (rho * (after[2] + after[3] + after[4])).sum(axis=0) * cfg.dz
    assert np.all(rain > 0) and np.all(influx > 0)
    # kg m^-2 lost by the column = mm of surface rain - mm entering at the top
    np.testing.assert_allclose(w0 - w1, rain - influx, rtol=0, atol=1e-11)
    # rain moves down: the rain mass-weighted height drops
    z = (np.arange(nz) + 0.5)[:, None, None]
    h0 = (rho * before[4] * z).sum() / (rho * before[4]).sum()
    h1 = (rho * after[4] * z).sum() / (rho * after[4]).sum()
    assert h1 < h0


def test_kessler_clips_negative_species():
    cfg, f, L, sl = _kessler_case(seed=4)
    rng = np.random.default_rng(4)
    cons = f[C.F_CONS]
    for comp in (C.RHOQ1, C.RHOQ2, C.RHOQ3):
        cons[comp][sl] -= cons[C.RHO][sl] * rng.uniform(0.0, 3e-3, cons[C.RHO][sl].shape)
    orc = Oracle(cfg)
    orc.load(f)
    orc.kessler(1.0)
    after = _q(orc.get(C.F_CONS), sl)
    for q in after[2:]:
        assert np.all(q >= 0.0)


def _dry_twin(cfg_moist, f_moist):
    cfg = C.default_config(cfg_moist.nx, cfg_moist.ny, cfg_moist.nz, cfg_moist.dx, cfg_moist.dy, cfg_moist.dz)
    for name, _ in C.Config._fields_:
        if name not in ("ncons", "moisture_type", "buoyancy_type"):
            setattr(cfg, name, getattr(cfg_moist, name))
    f = dict(f_moist)
    f[C.F_CONS] = f_moist[C.F_CONS][:C.NCONS_DRY].copy()
    return cfg, f


def test_moist_step_with_zero_moisture_matches_dry_step():
    cfg, f = P.moist_bubble(12, 10, 12, seed=6, buoyancy_type=1)
    for comp in (C.RHOQ1, C.RHOQ2, C.RHOQ3):
        f[C.F_CONS][comp][:] = 0.0
    cfg_d, f_d = _dry_twin(cfg, f)
    om = Oracle(cfg)
    om.load(f)
    od = Oracle(cfg_d)
    od.load(f_d)
    om.step(0.0, 1.0, 2)
    od.step(0.0, 1.0, 2)
    for fid in (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL):
        a = valid(om.get(fid), cfg, om.L, fid)
        b = valid(od.get(fid), cfg_d, od.L, fid)
        n = b.shape[0]
        # Kessler writes rho * (rho theta / rho): rounding only
        np.testing.assert_allclose(a[:n], b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
        if fid == C.F_CONS:
            assert np.all(a[C.RHOQ1:] == 0.0)


@pytest.mark.parametrize("btype", [1, 2, 3, 4])
def test_moist_bubble_steps_stay_physical(btype):
    cfg, f = P.moist_bubble(12, 10, 12, seed=8, buoyancy_type=btype)
    orc = Oracle(cfg)
    orc.load(f)
    orc.step(0.0, 1.0, 3)
    c = valid(orc.get(C.F_CONS), cfg, orc.L, C.F_CONS)
    w = valid(orc.get(C.F_ZVEL), cfg, orc.L, C.F_ZVEL)
    assert np.all(np.isfinite(c)) and np.all(np.isfinite(w))
    assert np.all(c[C.RHOQ1:] >= 0.0)
    assert np.max(np.abs(w)) < 5.0
    # the microphysics acted: rain fell out at the surface
    rain = orc.get(C.D_RAIN_ACCUM)
    assert rain.max() > 0.0
