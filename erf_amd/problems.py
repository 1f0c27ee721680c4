"""Synthetic initial states for the benchmark and test cases.

Each setup returns (cfg, fields) where `fields` maps a field id (F_CONS, F_XVEL,
..., F_PI0) to a float64 array shaped (ncomp, nz_tot, ny_tot, sy) in the layout
of include/erfx.h (index [n, k+ng, j+ng, i+ng]).  Inputs are generated with a
seeded numpy Generator; ERF's own random numbers cannot be reproduced.

The base state follows ERF's discrete hydrostatic balance:
  * density by Newton iteration per level (HSEutils::init_isentropic_hse,
    Source/Utils/ERF_HSE_utils.H:88-160; generalised to a theta(z) profile),
  * pressure, Exner function and ghost layers by ERF::erf_enforce_hse
    (Source/Initialization/ERF_init1d.cpp:153-313).
"""
import math
import numpy as np

from . import config as C


def _p_of_rt(rt):
    """getPgivenRTh (ERF_EOS.H:89-93), dry."""
    return C.p_0 * np.power(C.R_d * rt * 1.0 * (1.0 / C.p_0), C.Gamma)


def _rt_of_p(p):
    """getRhoThetagivenP (ERF_EOS.H:144-149), dry."""
    return np.power(p * math.pow(C.p_0, C.Gamma - 1), 1.0 / C.Gamma) * (1.0 / C.R_d) / 1.0


def _exner_of_p(p, rdOcp):
    """getExnergivenP (ERF_EOS.H:124-128)."""
    return np.power(p * (1.0 / C.p_0), rdOcp)


def hse_density(nz, dz, theta_cc, g=C.CONST_GRAV):
    """Column density in discrete HSE for cell-centre potential temperatures
    theta_cc[0..nz-1] (Newton iteration as in init_isentropic_hse with the
    surface pressure p_0 at z = 0).  Returns r[0..nz] (r[nz] = r[nz-1])."""
    max_iter, tol = 10, 1.e-8
    r = np.zeros(nz + 1)
    p = np.zeros(nz + 1)
    half_dz = 0.5 * dz
    th = theta_cc[0]
    r[0] = C.p_0 / (C.R_d * th)
    for _ in range(max_iter):
        p_hse = C.p_0 - half_dz * r[0] * g
        p_eos = float(_p_of_rt(r[0] * th))
        A = p_hse - p_eos
        dpdr = C.Gamma * C.p_0 * math.pow(C.R_d * th * (1.0 / C.p_0), C.Gamma) * math.pow(r[0], C.Gamma - 1.0)
        drho = A / (dpdr + half_dz * g)
        r[0] = r[0] + drho
        p[0] = float(_p_of_rt(r[0] * th))
        if abs(drho) < tol:
            break
    for k in range(1, nz):
        th = theta_cc[k]
        r[k] = r[k - 1]
        for _ in range(max_iter):
            r_avg = 0.5 * (r[k - 1] + r[k])
            p_hse = p[k - 1] - dz * r_avg * g
            p_eos = float(_p_of_rt(r[k] * th))
            A = p_hse - p_eos
            dpdr = C.Gamma * C.p_0 * math.pow(C.R_d * th * (1.0 / C.p_0), C.Gamma) * math.pow(r[k], C.Gamma - 1.0)
            drho = A / (dpdr + dz * g)
            r[k] = r[k] + drho
            p[k] = float(_p_of_rt(r[k] * th))
            if abs(drho) < tol * r[k - 1]:
                break
    r[nz] = r[nz - 1]
    return r


def enforce_hse(r, dz, rdOcp, g=C.CONST_GRAV):
    """erf_enforce_hse on a flat column: p, pi for k = -1..nz from r[-1..nz]
    (index offset 1: arrays have length nz+2)."""
    nz = len(r) - 2
    p = np.zeros(nz + 2)
    hz = 0.5 * dz
    p[1] = C.p_0 - hz * r[1] * g
    p[0] = C.p_0 + hz * r[1] * g
    for k in range(2, nz + 2):
        dens_interp = 0.5 * (r[k] + r[k - 1])
        p[k] = p[k - 1] - dz * dens_interp * g
    pi = _exner_of_p(p, rdOcp)
    return p, pi


def empty_fields(cfg):
    L = C.layout_for(cfg)
    shp = L.shape3()
    f = {}
    for fid in (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL, C.F_R0, C.F_P0, C.F_PI0):
        f[fid] = np.zeros((C.ncomp_of(cfg, fid),) + shp)
    if cfg.use_terrain:   # flat until a problem sets the real node heights
        zeta = cfg.prob_lo[2] + (np.arange(shp[0]) - L.ng) * cfg.dz
        f[C.F_Z_ND] = np.broadcast_to(zeta[:, None, None], shp)[None].copy()
    return L, f


def _base_state(cfg, L, f, theta_cc):
    """Fill r0, p0, pi0 (cells, one ghost layer all round, then copied into every
    lateral ghost column as FillBoundary would on a periodic box)."""
    g = C.CONST_GRAV if cfg.use_gravity else 0.0
    rdOcp = C.R_d / cfg.c_p
    nz, ng = cfg.nz, L.ng
    r = hse_density(nz, cfg.dz, theta_cc, g)
    rr = np.concatenate([[r[0]], r])               # k = -1 (Neumann) .. nz
    p, pi = enforce_hse(rr, cfg.dz, rdOcp, g)
    for fid, col in ((C.F_R0, rr), (C.F_P0, p), (C.F_PI0, pi)):
        a = f[fid][0]
        a[:] = 0.0
        a[ng - 1:ng + nz + 1, :, :] = col[:, None, None]
    return rr, p, pi


def _cell_centres(cfg, L):
    ng = L.ng
    kz = (np.arange(L.shape3()[0]) - ng + 0.5) * cfg.dz + cfg.prob_lo[2]
    jy = (np.arange(L.shape3()[1]) - ng + 0.5) * cfg.dy + cfg.prob_lo[1]
    ix = (np.arange(L.shape3()[2]) - ng + 0.5) * cfg.dx + cfg.prob_lo[0]
    return ix, jy, kz


def _fill_cons_from_theta(cfg, L, f, theta3):
    """rho*theta from the base-state pressure; rho = rho*theta / theta (pressure
    is unchanged by a theta perturbation: p depends on rho*theta only)."""
    ng, nz = L.ng, cfg.nz
    p0 = f[C.F_P0][0]
    rt = _rt_of_p(p0)
    cons = f[C.F_CONS]
    sl = (slice(ng, ng + nz),)
    cons[C.RHOTHETA][sl] = rt[sl]
    cons[C.RHO][sl] = rt[sl] / theta3[sl]


def rising_thermal(nx=64, ny=64, nz=64, seed=0, adv=C.ADV_CENTERED2, scal_adv=None):
    """Dry rising thermal: 20x20x10 km, periodic x/y, slip walls, theta = 300 K
    isentropic, +2 K cos^2 bubble of radius 2 km at (10, 10, 2) km, uniform
    noise of +-0.01 m/s on u, v, w; centred 2nd-order advection, no diffusion,
    dt = 1 s, MRI ratio 6.  `scal_adv` picks another scheme (e.g. WENO) for the
    passive scalars."""
    cfg = C.default_config(nx, ny, nz, 20000.0 / nx, 20000.0 / ny, 10000.0 / nz)
    cfg.dycore_horiz_adv = cfg.dycore_vert_adv = adv
    sa = adv if scal_adv is None else scal_adv
    cfg.dryscal_horiz_adv = cfg.dryscal_vert_adv = sa
    cfg.mri_ratio = 6
    L, f = empty_fields(cfg)
    _base_state(cfg, L, f, np.full(nz, 300.0))
    ix, jy, kz = _cell_centres(cfg, L)
    Lr = np.sqrt(((ix[None, None, :] - 10000.0) / 2000.0) ** 2 +
                 ((jy[None, :, None] - 10000.0) / 2000.0) ** 2 +
                 ((kz[:, None, None] - 2000.0) / 2000.0) ** 2)
    theta = 300.0 + np.where(Lr <= 1.0, 2.0 * (np.cos(math.pi * Lr) + 1.0) / 2.0, 0.0)
    _fill_cons_from_theta(cfg, L, f, theta)
    rng = np.random.default_rng(seed)
    _noise_vel(cfg, L, f, rng, 0.01, (0.0, 0.0, 0.0))
    return cfg, f


def _esat_bolton(T):
    """Saturation vapour pressure over water [mbar], Bolton (1980), for
    initial humidity only."""
    tc = T - 273.15
    return 6.112 * np.exp(17.67 * tc / (tc + 243.5))


def moist_bubble(nx=24, ny=20, nz=16, seed=0, buoyancy_type=None, norain=False, adv=C.ADV_UPWIND3):
    """Moist rising thermal with Kessler microphysics on the rising-thermal
    domain (20x20x10 km, periodic x/y, slip walls, isentropic 300 K base state):
    relative humidity 90% below 2 km decaying with a 1 km e-folding scale above;
    inside the +2 K cos^2 bubble the air is 5% supersaturated and carries up
    to 2 g/kg of cloud water and 0.5 g/kg of rain, so condensation,
    autoconversion, accretion, rain evaporation and sedimentation all act in
    the first step.  rho*theta keeps the base-state pressure with the moist
    equation of state (getRhoThetagivenP with qv, ERF_EOS.H:144-149).
    Dycore and dry scalars 3rd-order upwind, moist scalars WENO3 (ERF's
    defaults), dt = 1 s, MRI ratio 6."""
    cfg = C.default_config(nx, ny, nz, 20000.0 / nx, 20000.0 / ny, 10000.0 / nz)
    cfg.dycore_horiz_adv = cfg.dycore_vert_adv = adv
    cfg.dryscal_horiz_adv = cfg.dryscal_vert_adv = adv
    cfg.mri_ratio = 6
    C.set_moisture(cfg, C.MOIST_KESSLER_NORAIN if norain else C.MOIST_KESSLER, buoyancy_type)
    L, f = empty_fields(cfg)
    _base_state(cfg, L, f, np.full(nz, 300.0))
    ix, jy, kz = _cell_centres(cfg, L)
    Lr = np.sqrt(((ix[None, None, :] - 10000.0) / 2000.0) ** 2 +
                 ((jy[None, :, None] - 10000.0) / 2000.0) ** 2 +
                 ((kz[:, None, None] - 2000.0) / 2000.0) ** 2)
    shape = np.where(Lr <= 1.0, (np.cos(math.pi * Lr) + 1.0) / 2.0, 0.0)
    theta = 300.0 + 2.0 * shape
    rh_env = np.where(kz <= 2000.0, 0.9, 0.9 * np.exp(-(kz - 2000.0) / 1000.0))[:, None, None]
    rh = rh_env + (1.05 - rh_env) * shape
    p = f[C.F_P0][0]
    rdOcp = C.R_d / cfg.c_p
    T = theta * np.power(np.maximum(p, 1.0) / C.p_0, rdOcp)
    es = _esat_bolton(T)
    pmb = np.maximum(p, 1.0) / 100.0
    qv = rh * (C.R_d / C.R_v) * es / (pmb - es)
    qc = 2.0e-3 * shape
    qp = 5.0e-4 * shape if not norain else 0.0 * shape
    ng = L.ng
    sl = (slice(ng, ng + nz),)
    rt = _rt_of_p(p) / (1.0 + C.R_v / C.R_d * qv)
    cons = f[C.F_CONS]
    cons[C.RHOTHETA][sl] = rt[sl]
    cons[C.RHO][sl] = rt[sl] / theta[sl]
    for comp, q in ((C.RHOQ1, qv), (C.RHOQ2, qc), (C.RHOQ3, qp)):
        cons[comp][sl] = (cons[C.RHO] * q)[sl]
    rng = np.random.default_rng(seed)
    _noise_vel(cfg, L, f, rng, 0.01, (0.0, 0.0, 0.0))
    return cfg, f


def wk_sounding(z, theta_0=300.0, theta_tr=343.0, z_tr=12000.0, T_tr=213.0, cp=C.Cp_d, g=C.CONST_GRAV):
    """Weisman-Klemp (1982) analytic profiles at heights z: potential
    temperature theta_0 + (theta_tr - theta_0) (z/z_tr)^(5/4) below the
    tropopause, theta_tr exp(g (z - z_tr) / (c_p T_tr)) above; relative
    humidity 1 - 0.75 (z/z_tr)^(5/4) below, 0.25 above."""
    z = np.asarray(z, dtype=np.float64)
    s = np.power(np.clip(z / z_tr, 0.0, None), 1.25)
    theta = np.where(z <= z_tr, theta_0 + (theta_tr - theta_0) * s,
                     theta_tr * np.exp(g / (cp * T_tr) * (z - z_tr)))
    rh = np.where(z <= z_tr, 1.0 - 0.75 * s, 0.25)
    return theta, rh


def moist_hse(zc, dz, theta, rh, qv_max=0.014, rdOcp=C.R_d / C.Cp_d, g=C.CONST_GRAV, iters=40):
    """Discrete hydrostatic column of a moist sounding on flat levels: the total
    density r = rho_d (1 + qv) at cell centres with p_0 at z = 0,
    p[0] = p_0 - dz/2 r[0] g, p[k] = p[k-1] - dz (r[k] + r[k-1])/2 g (the
    recurrence of erf_enforce_hse), where rho_d theta (1 + R_v/R_d qv) gives
    p through the moist EOS and qv = min(rh qsat(T, p), qv_max).  Each level is
    solved by fixed-point iteration.  Returns (r, p, qv, rho_d)."""
    nz = len(zc)
    r = np.zeros(nz)
    p = np.zeros(nz)
    qv = np.zeros(nz)
    rd = np.zeros(nz)

    def state(pk, th, h):
        T = th * (pk / C.p_0) ** rdOcp
        es = float(_esat_bolton(T))
        q = min(h * (C.R_d / C.R_v) * es / (pk / 100.0 - es), qv_max)
        rho_d = float(_rt_of_p(pk)) / (1.0 + C.R_v / C.R_d * q) / th
        return rho_d * (1.0 + q), q, rho_d

    for k in range(nz):
        pk = C.p_0 if k == 0 else p[k - 1]
        for _ in range(iters):
            rk, q, rho_d = state(pk, theta[k], rh[k])
            pk = C.p_0 - 0.5 * dz * rk * g if k == 0 else p[k - 1] - dz * 0.5 * (rk + r[k - 1]) * g
        r[k], qv[k], rd[k] = state(pk, theta[k], rh[k])
        p[k] = pk
    return r, p, qv, rd


def supercell(nx=256, ny=256, nz=80, seed=0, lx=256000.0, ly=256000.0, lz=20000.0, radius=5.0, u_top=15.0,
              bubble=3.0, adv=C.ADV_UPWIND3):
    """Supercell-type deep convection: Weisman-Klemp sounding (wk_sounding,
    qv capped at 14 g/kg) in discrete moist hydrostatic balance (moist_hse),
    a quarter-circle hodograph of radius `radius` m/s over the lowest 2 km
    (u = R (1 - cos a), v = R sin a, a = pi/2 z/2 km), then u rising linearly
    to u_top at 6 km and constant above, v = R; a +bubble K cos^2 warm
    ellipsoid of radii 10 km x 10 km x 1.5 km centred 1.5 km above the domain
    centre (at the environment's pressure and vapour).  Default domain
    256 x 256 x 20 km with dx = dy = 1 km, dz = 250 m at 256x256x80; periodic
    laterally, slip walls.  Kessler with rain (buoyancy type 2), dycore U3,
    moist scalars WENO3, dt = 2 s, MRI ratio 6.  The seed draws +-0.1 K theta
    noise below 3 km and +-0.01 m/s velocity noise."""
    cfg = C.default_config(nx, ny, nz, lx / nx, ly / ny, lz / nz)
    cfg.dycore_horiz_adv = cfg.dycore_vert_adv = adv
    cfg.dryscal_horiz_adv = cfg.dryscal_vert_adv = adv
    cfg.mri_ratio = 6
    C.set_moisture(cfg, C.MOIST_KESSLER)
    L, f = empty_fields(cfg)
    ng = L.ng
    g = C.CONST_GRAV
    rdOcp = C.R_d / cfg.c_p
    zc = (np.arange(nz) + 0.5) * cfg.dz
    th_col, rh_col = wk_sounding(zc)
    r, _, qv_col, _ = moist_hse(zc, cfg.dz, th_col, rh_col, rdOcp=rdOcp, g=g)
    rr = np.concatenate([[r[0]], r, [r[-1]]])          # k = -1 .. nz
    p, pi = enforce_hse(rr, cfg.dz, rdOcp, g)
    for fid, col in ((C.F_R0, rr), (C.F_P0, p), (C.F_PI0, pi)):
        a = f[fid][0]
        a[:] = 0.0
        a[ng - 1:ng + nz + 1, :, :] = col[:, None, None]
    ix, jy, kz = _cell_centres(cfg, L)
    beta = np.sqrt(((ix[None, None, :] - 0.5 * lx) / 10000.0) ** 2 +
                   ((jy[None, :, None] - 0.5 * ly) / 10000.0) ** 2 +
                   ((kz[:, None, None] - 1500.0) / 1500.0) ** 2)
    dth = np.where(beta <= 1.0, bubble * np.cos(0.5 * math.pi * beta) ** 2, 0.0)
    rng = np.random.default_rng(seed)
    shp = L.shape3()
    noise = rng.uniform(-0.1, 0.1, shp) * (kz < 3000.0)[:, None, None]
    sl = (slice(ng, ng + nz),)
    theta = np.zeros(shp)
    theta[sl] = th_col[:, None, None] + dth[sl] + noise[sl]
    qv = np.zeros(shp)
    qv[sl] = qv_col[:, None, None]
    p3 = f[C.F_P0][0]
    cons = f[C.F_CONS]
    rt = _rt_of_p(p3) / (1.0 + C.R_v / C.R_d * qv)
    cons[C.RHOTHETA][sl] = rt[sl]
    cons[C.RHO][sl] = rt[sl] / theta[sl]
    cons[C.RHOQ1][sl] = (cons[C.RHO] * qv)[sl]
    _noise_vel(cfg, L, f, rng, 0.01, (0.0, 0.0, 0.0))
    a = 0.5 * math.pi * np.clip(zc / 2000.0, 0.0, 1.0)
    u = np.where(zc <= 2000.0, radius * (1.0 - np.cos(a)),
                 radius + (u_top - radius) * np.clip((zc - 2000.0) / 4000.0, 0.0, 1.0))
    v = np.where(zc <= 2000.0, radius * np.sin(a), radius)
    f[C.F_XVEL][0][ng:ng + nz] += u[:, None, None]
    f[C.F_YVEL][0][ng:ng + nz] += v[:, None, None]
    return cfg, f


def z_levels_uniform(nz, dz, zsurf=0.0):
    """init_zlevels without stretching (ERF_TerrainMetrics.cpp:27-33): k dz."""
    return zsurf + np.arange(nz + 1) * dz


def z_levels_stretched(nz, dz0, ratio, zsurf=0.0, dz_max=None):
    """init_zlevels with grid_stretching_ratio (ERF_TerrainMetrics.cpp:34-44):
    z_0 = zsurf, z_k = z_{k-1} + dz, dz *= ratio after each level; `dz_max`
    caps the spacing (an explicit terrain_z_levels list of that shape)."""
    z = np.zeros(nz + 1)
    z[0] = zsurf
    dz = dz0
    for k in range(1, nz + 1):
        z[k] = z[k - 1] + dz
        dz *= ratio
        if dz_max is not None and dz > dz_max:
            dz = dz_max
    return z


def terrain_nodes(cfg, L, height, z_levels=None):
    """Node heights z_phys_nd of the BTF terrain-following grid
    (init_which_terrain_grid case 0, ERF_TerrainMetrics.cpp:256-336; Klemp 2011):
    z = ((h - z_lev[0]) z_top + (z_top - h) z_lev[k]) / (z_top - z_lev[0]) on
    nodes k = 0..nz, with h the surface height `height(x, y)` at the node's
    column; z(-1) = 2 z(0) - z(1).  Lateral ghost nodes are periodic images on
    periodic axes and copies of the edge column otherwise (the reference's
    clamped index); the vertical ghost nodes the reference leaves unset
    continue the columns linearly (nothing reads them).  `z_levels` defaults
    to k dz."""
    ng = L.ng
    nzt, nyt, nxt = L.shape3()
    nz = cfg.nz
    zl = z_levels_uniform(nz, cfg.dz, cfg.prob_lo[2]) if z_levels is None else np.asarray(z_levels, float)
    assert zl.size == nz + 1
    ii = np.arange(nxt) - ng
    jj = np.arange(nyt) - ng
    ii = np.mod(ii, cfg.nx) if cfg.bc[0] == C.BC_PERIODIC else np.clip(ii, 0, cfg.nx)
    jj = np.mod(jj, cfg.ny) if cfg.bc[1] == C.BC_PERIODIC else np.clip(jj, 0, cfg.ny)
    x = cfg.prob_lo[0] + ii * cfg.dx
    y = cfg.prob_lo[1] + jj * cfg.dy
    h = height(x[None, :], y[:, None]) * np.ones((nyt, nxt))   # (ny_t, nx_t)
    z_top, z_lev_sfc = zl[-1], zl[0]
    z = np.zeros((nzt, nyt, nxt))
    for k in range(nz + 1):
        z[ng + k] = ((h - z_lev_sfc) * z_top + (z_top - h) * zl[k]) / (z_top - z_lev_sfc)
    z[ng - 1] = 2.0 * z[ng] - z[ng + 1]
    for kk in range(ng - 2, -1, -1):
        z[kk] = 2.0 * z[kk + 1] - z[kk + 2]
    for kk in range(ng + nz + 1, nzt):
        z[kk] = 2.0 * z[kk - 1] - z[kk - 2]
    return z[None]


def hse_terrain(zc, zs, ps, theta, g=C.CONST_GRAV, rdOcp=C.R_d / C.Cp_d):
    """Discrete hydrostatic balance on terrain-following columns, all columns
    at once.  zc: cell-centre heights (k = -1 .. nz, first axis), zs, ps:
    surface height and pressure of each column, theta: potential temperature
    at zc.  Density from p(k) - p(k-1) = -g (rho(k) + rho(k-1)) / 2
    (zc(k) - zc(k-1)) with the equation of state p = p(rho theta) (Newton per
    level, as hse_density); the k = -1 layer copies rho(0) and extends p
    hydrostatically.  Returns rho, p (the hydrostatic recursion) and Exner."""
    nk = zc.shape[0]
    r = np.zeros_like(zc)
    p = np.zeros_like(zc)

    def newton(k, p_ref, r_ref, w_ref, w_self):
        # solve p_ref - g (w_ref r_ref + w_self rho) = p(rho theta) for rho
        th = theta[k]
        rk = np.full_like(p_ref, 1.0) if r_ref is None else r_ref.copy()
        rr = 0.0 if r_ref is None else r_ref
        for _ in range(15):   # quadratic convergence: ~6 iterations reach rounding
            p_hse = p_ref - g * (w_ref * rr + w_self * rk)
            A = p_hse - _p_of_rt(rk * th)
            dpdr = C.Gamma * C.p_0 * np.power(C.R_d * th / C.p_0, C.Gamma) * np.power(rk, C.Gamma - 1.0)
            rk = rk + A / (dpdr + g * w_self)
        return rk, p_ref - g * (w_ref * rr + w_self * rk)

    # k = 0: from the surface pressure, rho(0) over the half cell below the centre
    r[1], p[1] = newton(1, ps, None, 0.0, zc[1] - zs)
    for k in range(2, nk):
        hdz = 0.5 * (zc[k] - zc[k - 1])
        r[k], p[k] = newton(k, p[k - 1], r[k - 1], hdz, hdz)
    r[0] = r[1]
    p[0] = p[1] + g * r[1] * (zc[1] - zc[0])
    return r, p, _exner_of_p(p, rdOcp)


def hill_flow(nx=24, ny=16, nz=16, seed=0, h0=500.0, a=2000.0, u0=10.0, v0=0.0, moist=False,
              adv=C.ADV_UPWIND3, lx=20000.0, ly=16000.0, noise=0.01, bubble=(0.3, 0.5), n2=0.0,
              ztop=10000.0, z_levels=None, theta_noise=0.0, scal_adv=None):
    """Flow over a Gaussian hill on a terrain-following grid: lx x ly x 10 km
    computational box, periodic laterally, slip walls; hill of height h0 and
    half-width a at the domain centre; isentropic theta = 300 K base state at
    the cells' true heights; uniform (u0, v0) plus seeded +-noise m/s and a
    +1 K bubble at 3 km (at fractions `bubble` of the box, None for none).
    n2 > 0 gives the initial state a stable stratification of
    Brunt-Vaisala frequency squared n2 in hydrostatic balance (the base state
    stays isentropic).  With `moist`, Kessler and a humid layer.  The
    computational box is ztop high (dz = ztop / nz); `z_levels` (nz+1 node
    levels, default k dz) stretch the grid; theta_noise adds seeded uniform
    +-theta_noise K; `scal_adv` picks the scalar scheme (default: adv)."""
    cfg = C.default_config(nx, ny, nz, lx / nx, ly / ny, ztop / nz)
    cfg.use_terrain = 1
    cfg.dycore_horiz_adv = cfg.dycore_vert_adv = adv
    sa = adv if scal_adv is None else scal_adv
    cfg.dryscal_horiz_adv = cfg.dryscal_vert_adv = sa
    cfg.mri_ratio = 6
    if moist:
        C.set_moisture(cfg, C.MOIST_KESSLER, 1)
        cfg.moistscal_horiz_adv = cfg.moistscal_vert_adv = sa if scal_adv is not None else cfg.moistscal_horiz_adv
    L, f = empty_fields(cfg)
    xc, yc = 0.5 * lx, 0.5 * ly
    zn = terrain_nodes(cfg, L, lambda x, y: h0 * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / a ** 2), z_levels)
    f[C.F_Z_ND] = zn
    z = zn[0]
    ng = L.ng
    # cell-centre and surface heights of every column (k = -1 .. nz)
    zc = np.zeros_like(z)
    zc[:-1, :-1, :-1] = 0.125 * (z[:-1, :-1, :-1] + z[:-1, :-1, 1:] + z[:-1, 1:, :-1] + z[:-1, 1:, 1:] +
                                 z[1:, :-1, :-1] + z[1:, :-1, 1:] + z[1:, 1:, :-1] + z[1:, 1:, 1:])
    zc[-1] = zc[-2]; zc[:, -1] = zc[:, -2]; zc[:, :, -1] = zc[:, :, -2]
    zs = 0.25 * (z[ng, :-1, :-1] + z[ng, :-1, 1:] + z[ng, 1:, :-1] + z[ng, 1:, 1:])
    zs = np.pad(zs, ((0, 1), (0, 1)), mode="edge")
    kk = slice(ng - 1, ng + nz + 1)
    g = C.CONST_GRAV
    rdOcp = C.R_d / cfg.c_p
    # surface pressures of the continuous profiles (isentropic; constant n2)
    ps_isen = C.p_0 * np.power(1.0 - g * zs / (cfg.c_p * 300.0), 1.0 / rdOcp)
    rho0, p0, pi0 = hse_terrain(zc[kk], zs, ps_isen, np.full_like(zc[kk], 300.0), g, rdOcp)
    for fid, a3 in ((C.F_R0, rho0), (C.F_P0, p0), (C.F_PI0, pi0)):
        f[fid][0][kk] = a3
    ix, jy, kz = _cell_centres(cfg, L)
    theta = np.full_like(zc, 300.0)
    pr = f[C.F_P0][0]
    if n2 > 0:   # theta = 300 exp(n2 z / g) in discrete balance of its own
        theta = 300.0 * np.exp(n2 * zc / g)
        pr = np.zeros_like(zc)
        pis = 1.0 + g * g / (cfg.c_p * 300.0 * n2) * (np.exp(-n2 * zs / g) - 1.0)
        pr[kk] = hse_terrain(zc[kk], zs, C.p_0 * np.power(pis, 1.0 / rdOcp), theta[kk], g, rdOcp)[1]
    rng = np.random.default_rng(seed)
    if theta_noise > 0:
        theta = theta + rng.uniform(-theta_noise, theta_noise, theta.shape)
    if bubble is not None:
        Lr = np.sqrt(((ix[None, None, :] - bubble[0] * lx) / 2000.0) ** 2 +
                     ((jy[None, :, None] - bubble[1] * ly) / 2000.0) ** 2 + ((zc - 3000.0) / 1500.0) ** 2)
        theta = theta + np.where(Lr <= 1.0, (np.cos(math.pi * Lr) + 1.0) / 2.0, 0.0)
    sl = (slice(ng, ng + nz),)
    rt = _rt_of_p(np.maximum(pr, 1.0))
    cons = f[C.F_CONS]
    if moist:
        T = theta * np.power(np.maximum(f[C.F_P0][0], 1.0) / C.p_0, rdOcp)
        es = _esat_bolton(T)
        pmb = np.maximum(f[C.F_P0][0], 1.0) / 100.0
        rh = np.where(zc <= 2000.0, 0.9, 0.9 * np.exp(-(zc - 2000.0) / 1000.0))
        qv = rh * (C.R_d / C.R_v) * es / (pmb - es)
        rt = rt / (1.0 + C.R_v / C.R_d * qv)
    cons[C.RHOTHETA][sl] = rt[sl]
    cons[C.RHO][sl] = rt[sl] / theta[sl]
    if moist:
        cons[C.RHOQ1][sl] = (cons[C.RHO] * qv)[sl]
    _noise_vel(cfg, L, f, rng, noise, (u0, v0, 0.0))
    return cfg, f


MOIST_TERRAIN_DT = 0.5


def moist_terrain(nx=512, ny=512, nz=96, seed=0, h0=500.0, a=2000.0):
    """Moist flow over terrain with Kessler on the terrain-following path:
    dx = dy = 200 m, levels stretched from 20 m at the surface to 400 m
    (geometric growth capped at 400 m) with the top at 15 km, Gaussian hill
    h0 = 500 m, a = 2 km at the centre; u = 10 m/s, N = 0.01 1/s, RH 90 %
    below 2 km decaying with a 1 km e-fold scale above, seeded +-0.1 K theta
    noise; WENO5 scalars (dry and moist), 3rd-order upwind momentum, Rayleigh
    damping above 12 km towards the initial wind, 6 substeps per RK stage.
    Laterally periodic (the inflow is the periodic image of the outflow).
    The nz+1 levels reach 15 km exactly: the growth ratio is solved for.

    Time step: 0.5 s (MOIST_TERRAIN_DT), not the 1 s of the case's recipe.  At
    1 s the moist state on these 20 m surface cells grows an oscillation that
    changes sign every step (w doubling per step from about step 14, q_v
    swinging by tens of percent between steps, no cloud yet) and the run
    leaves the saturation fit's range (below -85 C) within 20 steps, in the
    oracle exactly as on the GPU; 6 and 12 substeps, 32^2 and 64^2 columns, a
    flat bottom (h0 = 0), C2 momentum and no Rayleigh layer all fail the same
    way, the dry state on the same levels and the moist state on uniform
    150 m levels do not.  At 0.5 s the run stays bounded (|w| <= 2.3 m/s over
    120 steps).  The work of a step does not depend on dt."""
    lo, hi = 1.0, 1.2
    for _ in range(200):   # bisection on the growth ratio for the 15 km top
        r = 0.5 * (lo + hi)
        top = z_levels_stretched(nz, 20.0, r, dz_max=400.0)[-1]
        lo, hi = (r, hi) if top < 15000.0 else (lo, r)
    zl = z_levels_stretched(nz, 20.0, 0.5 * (lo + hi), dz_max=400.0)
    zl = zl * (15000.0 / zl[-1])
    cfg, f = hill_flow(nx, ny, nz, seed=seed, h0=h0, a=a, u0=10.0, moist=True, adv=C.ADV_UPWIND3,
                       lx=200.0 * nx, ly=200.0 * ny, noise=0.0, bubble=None, n2=1e-4, ztop=15000.0,
                       z_levels=zl, theta_noise=0.1, scal_adv=C.ADV_WENO5)
    cfg.rayleigh_damp_U = cfg.rayleigh_damp_V = cfg.rayleigh_damp_W = cfg.rayleigh_damp_T = 1
    cfg.rayleigh_zdamp = 3000.0
    cfg.rayleigh_dampcoef = 0.2
    return cfg, f


def rayleigh_profiles_uniform(cfg, u=10.0, v=0.0, theta=None, f=None):
    """Rayleigh reference profiles: the uniform wind, w = 0, and theta of the
    initial state's column-mean per level (k = 0..nz, the top copies nz-1)."""
    n1 = cfg.nz + 1
    if theta is None:
        L = C.layout_for(cfg)
        ng = L.ng
        c = f[C.F_CONS]
        th = (c[C.RHOTHETA] / np.where(c[C.RHO] > 0, c[C.RHO], 1.0))[ng:ng + cfg.nz, ng:ng + cfg.ny, ng:ng + cfg.nx]
        tm = th.mean(axis=(1, 2))
        theta = np.concatenate([tm, tm[-1:]])
    return np.full(n1, u), np.full(n1, v), np.zeros(n1), np.asarray(theta, float)


def _noise_vel(cfg, L, f, rng, amp, mean):
    ng, nx, ny, nz = L.ng, cfg.nx, cfg.ny, cfg.nz
    for fid, (ex, ey, ez), m in ((C.F_XVEL, (1, 0, 0), mean[0]), (C.F_YVEL, (0, 1, 0), mean[1]),
                                 (C.F_ZVEL, (0, 0, 1), mean[2])):
        a = f[fid][0]
        shape = (nz + ez, ny + ey, nx + ex)
        a[ng:ng + nz + ez, ng:ng + ny + ey, ng:ng + nx + ex] = m + rng.uniform(-amp, amp, shape)
    # periodic faces must agree on both sides (ERF OverrideSync, ERF.cpp:1350)
    if cfg.bc[0] == C.BC_PERIODIC:
        f[C.F_XVEL][0][:, :, ng + nx] = f[C.F_XVEL][0][:, :, ng]
    if cfg.bc[1] == C.BC_PERIODIC:
        f[C.F_YVEL][0][:, ng + ny, :] = f[C.F_YVEL][0][:, ng, :]
    # no-penetration walls in z
    f[C.F_ZVEL][0][ng, :, :] = 0.0
    f[C.F_ZVEL][0][ng + nz, :, :] = 0.0


def cbl(nx=256, ny=256, nz=128, seed=0, ug0=10.0):
    """Dry convective boundary layer LES: dx = dy = 20 m, dz = 10 m, periodic
    laterally, no-slip bottom with a surface heat flux of 0.1 K m/s, slip top;
    theta = 300 K below 500 m then +0.003 K/m; uniform theta noise of +-0.25 K
    in the lowest 200 m; geostrophic wind (ug0, 0) m/s, ug0 = 10 by default,
    with f = 1e-4 s^-1;
    5th-order upwind advection, Smagorinsky Cs = 0.1, dt = 0.5 s.

    The MRI ratio is 16 (1 + 8 + 16 acoustic substeps per step), not 4: at
    dx = dy = 20 m a ratio of 4 gives dtau = 0.125 s and a horizontal acoustic
    Courant number c*dtau*sqrt(dx^-2 + dy^-2) ~ 3, and the forward-backward
    substep of erf_fast_rhs_N diverges within one step (in the oracle as on
    the GPU); 16 is the smallest even ratio that stays stable."""
    cfg = C.default_config(nx, ny, nz, 20.0, 20.0, 10.0)
    cfg.bc[2] = C.BC_NO_SLIP_WALL
    cfg.bc[5] = C.BC_SLIP_WALL
    cfg.dycore_horiz_adv = cfg.dycore_vert_adv = C.ADV_UPWIND5
    cfg.dryscal_horiz_adv = cfg.dryscal_vert_adv = C.ADV_UPWIND5
    cfg.les_type = C.LES_SMAGORINSKY
    cfg.Cs = 0.1
    cfg.mri_ratio = 16
    cfg.use_coriolis = 1
    # rotational_time_period = 4 pi / f, latitude 90 deg (ERF_DataStruct.H:475-516)
    rot_period = 4.0 * math.pi / 1.0e-4
    cfg.coriolis_factor = 2.0 * 2.0 * math.pi / rot_period
    lat = 90.0 * (math.pi / 180.)
    cfg.sinphi = math.sin(lat)
    cfg.cosphi = math.cos(lat)
    ug = (float(ug0), 0.0, 0.0)
    cfg.abl_geo_forcing[0] = -cfg.coriolis_factor * (ug[1] * cfg.sinphi - ug[2] * cfg.cosphi)
    cfg.abl_geo_forcing[1] = cfg.coriolis_factor * ug[0] * cfg.sinphi
    cfg.abl_geo_forcing[2] = -cfg.coriolis_factor * ug[0] * cfg.cosphi
    L, f = empty_fields(cfg)
    zc = (np.arange(nz) + 0.5) * cfg.dz
    th_col = np.where(zc < 500.0, 300.0, 300.0 + 0.003 * (zc - 500.0))
    rr, _, _ = _base_state(cfg, L, f, th_col)
    # surface flux of rho*theta [kg m^-2 s^-1 K] = rho_sfc * 0.1 K m/s
    cfg.use_surface_heat_flux = 1
    cfg.surface_heat_flux = rr[1] * 0.1
    ix, jy, kz = _cell_centres(cfg, L)
    rng = np.random.default_rng(seed)
    theta = np.zeros(L.shape3())
    ng = L.ng
    theta[ng:ng + nz] = th_col[:, None, None]
    noise = rng.uniform(-0.25, 0.25, (nz, L.shape3()[1], L.shape3()[2]))
    mask = (zc < 200.0)[:, None, None]
    theta[ng:ng + nz] += np.where(mask, noise, 0.0)
    theta[:ng] = 300.0
    theta[ng + nz:] = 300.0
    _fill_cons_from_theta(cfg, L, f, theta)
    _noise_vel(cfg, L, f, rng, 0.0, ug)
    f[C.F_XVEL][0][ng:ng + nz, ng:ng + ny, ng:ng + nx + 1] = ug[0]
    return cfg, f


ENSEMBLE_SIZES = ((48, 48, 48), (64, 64, 96), (96, 96, 64), (128, 128, 80), (160, 160, 80))


def inflow_outflow(nx=24, ny=16, nz=16, seed=0, u0=10.0, moist=False, adv=C.ADV_UPWIND3):
    """Flow through the rising-thermal domain (20 x 20 x 10 km): ERF's Inflow
    on the x-lo face (velocity (u0, 0, 0), density 1.1 kg/m^3, theta 300 K,
    scalar 0.5 and, with moisture, qv 5 g/kg, qc 0), Outflow on x-hi, periodic
    in y, slip walls; the thermal's +2 K bubble (or the moist bubble) carried
    downstream by u0 with +-0.01 m/s noise; dt = 1 s, MRI ratio 6."""
    if moist:
        cfg, f = moist_bubble(nx, ny, nz, seed=seed, adv=adv)
    else:
        cfg, f = rising_thermal(nx, ny, nz, seed=seed, adv=adv)
    cfg.bc[0] = C.BC_INFLOW
    cfg.bc[3] = C.BC_OUTFLOW
    cfg.inflow_velocity[0][0] = u0
    cfg.inflow_density[0] = 1.1
    cfg.inflow_theta[0] = 300.0
    cfg.inflow_scalar[0] = 0.5
    if moist:
        cfg.inflow_qv[0] = 5.0e-3
    L = C.layout_for(cfg)
    ng = L.ng
    f[C.F_XVEL][0][ng:ng + nz, ng:ng + ny, ng:ng + nx + 1] += u0
    return cfg, f


def cbl_ensemble(n=48, seed=0, sizes=ENSEMBLE_SIZES):
    """Members of the batched ensemble: independent dry CBL boxes (cbl), each
    with its own noise seed, a geostrophic speed drawn uniformly from
    [5, 15] m/s and a box size drawn uniformly from `sizes`.  Yields
    (cfg, fields) one member at a time."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(sizes), n)
    speed = rng.uniform(5.0, 15.0, n)
    for m in range(n):
        nx, ny, nz = sizes[int(pick[m])]
        yield cbl(nx, ny, nz, seed=1000 * seed + m + 1, ug0=float(speed[m]))


def small_case(kind, nx, ny, nz, seed=0, **kw):
    """Reduced-size versions of the cases for tests (same physics)."""
    if kind == "thermal":
        return rising_thermal(nx, ny, nz, seed)
    if kind == "cbl":
        return cbl(nx, ny, nz, seed)
    raise ValueError(kind)
