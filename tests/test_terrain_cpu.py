"""Terrain-following grid in the oracle (use_terrain): metrics, the terrain
forms of the slow and fast right-hand sides, and the wall condition.

What pins them, since no golden vectors for terrain exist in the reference:
- the metrics of a Gaussian hill against their closed forms;
- a flat grid stretched by 2 in the vertical (z = 2 zeta, so detJ = ax = ay =
  2) run in terrain form against the same problem on a plain grid with 2 dz:
  every terrain factor must cancel, to rounding;
- conservation of the integrals of detJ rho and detJ rho theta over the hill;
- mirror symmetry in y and the x <-> y transpose over a symmetric hill;
- the horizontal pressure gradient of a horizontally homogeneous atmosphere
  at rest over the hill, against the same gradient without the metric term;
- the wall value of w against u dh/dx (the kinematic condition).
"""
import copy

import numpy as np
import pytest

from erf_amd import config as C, problems as P
from oracle_lib import Oracle


def _valid(cfg, L, a, ex=0, ey=0, ez=0):
    ng = L.ng
    return a[..., ng:ng + cfg.nz + ez, ng:ng + cfg.ny + ey, ng:ng + cfg.nx + ex]


def test_metrics_closed_form():
    h0, a = 500.0, 2000.0
    cfg, f = P.hill_flow(20, 16, 10, h0=h0, a=a)
    o = Oracle(cfg)
    o.load(f)
    L = C.layout_for(cfg)
    ng, nx, ny, nz = L.ng, cfg.nx, cfg.ny, cfg.nz
    ztop = nz * cfg.dz
    x = (np.arange(nx + 1)) * cfg.dx
    y = (np.arange(ny + 1)) * cfg.dy
    h = h0 * np.exp(-((x[None, :] - 10000.0) ** 2 + (y[:, None] - 8000.0) ** 2) / a ** 2)   # nodes (j, i)
    # z = zeta + h (1 - zeta / ztop) is linear in zeta: detJ = 1 - mean4(h) / ztop
    h4 = 0.25 * (h[:-1, :-1] + h[:-1, 1:] + h[1:, :-1] + h[1:, 1:])
    J = _valid(cfg, L, o.get(C.D_DETJ)[0])
    assert np.abs(J - (1.0 - h4 / ztop)[None]).max() < 1e-13
    # areas: the same on the two nodes of each face
    ax = _valid(cfg, L, o.get(C.D_AX)[0], ex=1)
    assert np.abs(ax - (1.0 - 0.5 * (h[:-1, :] + h[1:, :]) / ztop)[None]).max() < 1e-13
    ay = _valid(cfg, L, o.get(C.D_AY)[0], ey=1)
    assert np.abs(ay - (1.0 - 0.5 * (h[:, :-1] + h[:, 1:]) / ztop)[None]).max() < 1e-13
    assert np.all(_valid(cfg, L, o.get(C.D_AZ)[0], ez=1) == 1.0)
    zeta_c = (np.arange(nz) + 0.5) * cfg.dz
    zc = zeta_c[:, None, None] + h4[None] * (1.0 - zeta_c[:, None, None] / ztop)
    assert np.abs(_valid(cfg, L, o.get(C.D_Z_CC)[0]) - zc).max() < 1e-9
    # a column of cells spans the height between surface and lid
    assert np.abs((J * cfg.dz).sum(axis=0) - (ztop - h4)).max() < 1e-9


@pytest.mark.parametrize("adv,rayleigh", [(C.ADV_CENTERED2, False), (C.ADV_UPWIND3, False), (C.ADV_WENO5, False),
                                          (C.ADV_UPWIND3, True)])
def test_flat_stretched_equals_flat(adv, rayleigh):
    """z = 2 zeta in terrain form is the plain grid with 2 dz (with a Rayleigh
    damping layer: its heights come from z_cc and the nodes on the terrain grid)."""
    cfgA, fA = P.rising_thermal(16, 12, 16, adv=adv)
    fA[C.F_XVEL][0] += 5.0
    fA[C.F_YVEL][0] -= 3.0
    if rayleigh:
        cfgA.rayleigh_damp_U = cfgA.rayleigh_damp_V = cfgA.rayleigh_damp_W = cfgA.rayleigh_damp_T = 1
        cfgA.rayleigh_zdamp = 4000.0
        cfgA.rayleigh_dampcoef = 0.05
    cfgB = copy.deepcopy(cfgA)
    cfgB.dz = cfgA.dz / 2
    cfgB.use_terrain = 1
    L = C.layout_for(cfgB)
    assert L.nelem == C.layout_for(cfgA).nelem
    fB = dict(fA)
    zeta = cfgA.prob_lo[2] + (np.arange(L.shape3()[0]) - L.ng) * cfgA.dz
    fB[C.F_Z_ND] = np.broadcast_to(zeta[:, None, None], L.shape3())[None].copy()
    oA, oB = Oracle(cfgA), Oracle(cfgB)
    oA.load(fA)
    oB.load(fB)
    if rayleigh:
        n1 = cfgA.nz + 1
        prof = (np.full(n1, 4.0), np.full(n1, -2.0), np.zeros(n1), np.full(n1, 300.4))
        oA.set_rayleigh(*prof)
        oB.set_rayleigh(*prof)
    assert np.all(_valid(cfgB, L, oB.get(C.D_DETJ)[0]) == 2.0)
    oA.step(0.0, 1.0, 4)
    oB.step(0.0, 1.0, 4)
    for fid in (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL):
        a, b = oA.get(fid), oB.get(fid)
        for c in range(a.shape[0]):
            change = np.abs(a[c] - fA[fid][c]).max()
            assert np.abs(a[c] - b[c]).max() <= 1e-9 * change + 1e-300, (fid, c)


def _integrals(o, cfg, L):
    c = o.get(C.F_CONS)
    J = _valid(cfg, L, o.get(C.D_DETJ)[0])
    return [float((_valid(cfg, L, c[n]) * J).sum()) for n in range(c.shape[0])]


@pytest.mark.parametrize("moist", [False, True])
def test_conservation_over_hill(moist):
    cfg, f = P.hill_flow(24, 16, 16, moist=moist)
    o = Oracle(cfg)
    o.load(f)
    L = C.layout_for(cfg)
    m0 = _integrals(o, cfg, L)
    o.step(0.0, 1.0, 8)
    m1 = _integrals(o, cfg, L)
    assert abs(m1[C.RHO] - m0[C.RHO]) <= 1e-13 * abs(m0[C.RHO])
    if not moist:   # latent heating changes rho theta
        assert abs(m1[C.RHOTHETA] - m0[C.RHOTHETA]) <= 1e-13 * abs(m0[C.RHOTHETA])
    else:
        assert o.get(C.D_RAIN_ACCUM).min() >= 0.0
    for fid in (C.F_CONS, C.F_XVEL, C.F_YVEL, C.F_ZVEL):
        assert np.all(np.isfinite(o.get(fid)))
    w = o.get(C.F_ZVEL)
    assert 0.5 < np.abs(w).max() < 10.0       # the hill lifts the flow, boundedly


def test_mirror_symmetry_in_y():
    """Hill, bubble and flow symmetric about the mid-plane in y: u, w, rho,
    rho theta stay symmetric and v antisymmetric."""
    cfg, f = P.hill_flow(20, 16, 12, noise=0.0)
    o = Oracle(cfg)
    o.load(f)
    o.step(0.0, 1.0, 5)
    L = C.layout_for(cfg)
    c = _valid(cfg, L, o.get(C.F_CONS))
    for n in (C.RHO, C.RHOTHETA):
        assert np.abs(c[n] - c[n][:, ::-1]).max() <= 1e-13 * np.abs(c[n]).max()
    u = _valid(cfg, L, o.get(C.F_XVEL)[0], ex=1)
    v = _valid(cfg, L, o.get(C.F_YVEL)[0], ey=1)
    w = _valid(cfg, L, o.get(C.F_ZVEL)[0], ez=1)
    assert np.abs(v).max() > 0.1
    assert np.abs(u - u[:, ::-1]).max() <= 1e-12 * np.abs(u).max()
    assert np.abs(v + v[:, ::-1]).max() <= 1e-12 * np.abs(v).max()
    assert np.abs(w - w[:, ::-1]).max() <= 1e-12 * np.abs(w).max()


def test_transpose_xy():
    """Flow along x over a round hill is the transpose of flow along y."""
    kw = dict(noise=0.0, lx=16000.0, ly=16000.0)
    cfgA, fA = P.hill_flow(16, 16, 12, u0=8.0, v0=0.0, bubble=(0.3, 0.5), **kw)
    cfgB, fB = P.hill_flow(16, 16, 12, u0=0.0, v0=8.0, bubble=(0.5, 0.3), **kw)
    oA, oB = Oracle(cfgA), Oracle(cfgB)
    oA.load(fA)
    oB.load(fB)
    oA.step(0.0, 1.0, 5)
    oB.step(0.0, 1.0, 5)
    L = C.layout_for(cfgA)
    cA, cB = _valid(cfgA, L, oA.get(C.F_CONS)), _valid(cfgA, L, oB.get(C.F_CONS))
    for n in (C.RHO, C.RHOTHETA):
        assert np.abs(cA[n] - cB[n].transpose(0, 2, 1)).max() <= 1e-13 * np.abs(cA[n]).max()
    uA = _valid(cfgA, L, oA.get(C.F_XVEL)[0], ex=1)
    vB = _valid(cfgA, L, oB.get(C.F_YVEL)[0], ey=1)
    assert np.abs(uA - vB.transpose(0, 2, 1)).max() <= 1e-12 * np.abs(uA).max()
    wA = _valid(cfgA, L, oA.get(C.F_ZVEL)[0], ez=1)
    wB = _valid(cfgA, L, oB.get(C.F_ZVEL)[0], ez=1)
    assert np.abs(wA - wB.transpose(0, 2, 1)).max() <= 1e-12 * np.abs(wA).max()


def test_pressure_gradient_metric_term():
    """A stratified atmosphere at rest, horizontally homogeneous and in
    discrete hydrostatic balance in every column: along the sloping grid
    surfaces p' varies strongly, and the metric term must cancel that."""
    cfg, f = P.hill_flow(32, 32, 20, u0=0.0, noise=0.0, bubble=None, n2=1e-4, h0=800.0,
                         lx=16000.0, ly=16000.0)
    o = Oracle(cfg)
    o.load(f)
    L = C.layout_for(cfg)
    ng = L.ng
    o.set(C.state_id(C.S_NEW, C.C_CONS), f[C.F_CONS])
    for cc in (C.C_XMOM, C.C_YMOM, C.C_ZMOM):
        o.set(C.state_id(C.S_NEW, cc), np.zeros_like(f[C.F_XVEL]))
    o.call("orc_apply_bcs", C.S_NEW, ng, ng - 1, 0, 1)
    o.call("orc_cons_to_prim", C.state_id(C.S_NEW, C.C_CONS), ng)
    o.call("orc_slow_rhs_pre", 0, 1.0)
    rur = _valid(cfg, L, o.get(C.state_id(C.F_SLOW, C.C_XMOM))[0], ex=1)
    rvr = _valid(cfg, L, o.get(C.state_id(C.F_SLOW, C.C_YMOM))[0], ey=1)
    c = o.get(C.state_id(C.S_NEW, C.C_CONS))
    pp = _valid(cfg, L, P._p_of_rt(c[C.RHOTHETA]) - f[C.F_P0][0], ex=1, ey=1)
    naive_x = np.abs(np.diff(pp[:, :-1, :], axis=2)).max() / cfg.dx
    naive_y = np.abs(np.diff(pp[:, :, :-1], axis=1)).max() / cfg.dy
    assert naive_x > 0.04 and naive_y > 0.04          # Pa/m along the grid surfaces
    # interior levels: centred vertical differences; k = 0 and nz-1 take
    # one-sided ones (ERF_slow_rhs_pre.cpp:606-613), first order
    assert np.abs(rur[1:-1]).max() < 0.01 * naive_x
    assert np.abs(rvr[1:-1]).max() < 0.01 * naive_y
    assert np.abs(rur).max() < 0.2 * naive_x
    assert np.abs(rvr).max() < 0.2 * naive_y
    # and the state stays near rest through full steps
    o2 = Oracle(cfg)
    o2.load(f)
    o2.step(0.0, 1.0, 5)
    assert np.abs(o2.get(C.F_XVEL)).max() < 0.05
    assert np.abs(o2.get(C.F_ZVEL)).max() < 0.05


def test_wall_kinematic_condition():
    """w at the ground follows the terrain: w = u dh/dx there."""
    h0, a = 500.0, 2000.0
    cfg, f = P.hill_flow(48, 8, 16, noise=0.0, bubble=None, h0=h0, a=a, lx=20000.0, ly=4000.0)
    o = Oracle(cfg)
    o.load(f)
    o.step(0.0, 1.0, 1)
    L = C.layout_for(cfg)
    ng, nx = L.ng, cfg.nx
    w = o.get(C.F_ZVEL)[0]
    u = o.get(C.F_XVEL)[0]
    j = ng + 2
    x = (np.arange(nx) + 0.5) * cfg.dx
    yj = (j - ng + 0.5) * cfg.dy
    dhdx = -2 * (x - 10000.0) / a ** 2 * h0 * np.exp(-((x - 10000.0) ** 2 + (yj - 2000.0) ** 2) / a ** 2)
    ub = 0.5 * (u[ng, j, ng:ng + nx] + u[ng, j, ng + 1:ng + nx + 1])
    ua = 0.5 * (u[ng + 1, j, ng:ng + nx] + u[ng + 1, j, ng + 1:ng + nx + 1])
    us = 1.5 * ub - 0.5 * ua                 # u extrapolated to the ground
    w0 = w[ng, j, ng:ng + nx]
    assert np.abs(w0).max() > 1.0
    assert np.abs(w0 - us * dhdx).max() < 0.05 * np.abs(w0).max()


def test_terrain_with_les_rejected():
    cfg, _ = P.hill_flow(8, 8, 8)
    cfg.les_type = C.LES_SMAGORINSKY
    with pytest.raises(RuntimeError):
        Oracle(cfg)


@pytest.mark.parametrize("molec", [C.MOLEC_CONSTANT, C.MOLEC_CONSTANT_ALPHA])
def test_terrain_diffusion_on_flat_levels_is_the_flat_diffusion(molec):
    """Over flat ground with uniform levels every slope of the terrain grid
    vanishes and detJ = 1, so DiffusionSrcForState_T, ComputeStrain_T /
    ComputeStress_T and DiffusionSrcForMom_T reduce to their flat forms (which
    the DensityCurrent, ScalarDiffusion and Couette gold files pin): the
    terrain step must give the flat step's fields to rounding.  A dropped
    metric factor or a slope term with the wrong sign would not vanish here only
    if it multiplied detJ or h_zeta, which is what DensityCurrent_detJ2 pins."""
    cfg, f = P.hill_flow(12, 10, 10, seed=5, h0=0.0, noise=0.5)
    cfg.molec_diff_type = molec
    cfg.dynamic_viscosity = 20.0
    cfg.rho0_trans = 1.0
    cfg.alpha_T = 30.0
    cfg.alpha_C = 10.0
    orc_t = Oracle(cfg)
    orc_t.load(f)
    flat = C.Config.from_buffer_copy(cfg)
    flat.use_terrain = 0
    orc_f = Oracle(flat)
    orc_f.load({k: v for k, v in f.items() if k != C.F_Z_ND})
    orc_t.step(0.0, 1.0, 2)
    orc_f.step(0.0, 1.0, 2)
    L = C.layout_for(cfg)
    assert L.nelem == C.layout_for(flat).nelem
    ext = {C.F_CONS: (0, 0, 0), C.F_XVEL: (1, 0, 0), C.F_YVEL: (0, 1, 0), C.F_ZVEL: (0, 0, 1)}
    for fid, (ex, ey, ez) in ext.items():
        a = _valid(cfg, L, orc_t.get(fid), ex, ey, ez)
        b = _valid(cfg, L, orc_f.get(fid), ex, ey, ez)
        init = _valid(cfg, L, f[fid], ex, ey, ez)
        for c in range(a.shape[0]):
            if fid == C.F_CONS and c in (C.RHOKE, C.RHOQKE):
                continue
            change = np.abs(b[c] - init[c]).max()     # 0 for the scalar, which starts at 0
            assert np.abs(a[c] - b[c]).max() <= 1e-10 * change, (fid, c)


def test_stretched_levels_and_btf_grid():
    """init_zlevels with a stretching ratio (capped) and the BTF grid: the
    stretched column over flat ground is the level list itself, and over the
    hill the lid stays flat at z_top."""
    zl = P.z_levels_stretched(96, 20.0, 1.035883433721025, dz_max=400.0)
    dz = np.diff(zl)
    assert dz[0] == 20.0 and dz.max() == 400.0 and abs(zl[-1] - 15000.0) < 1e-6
    assert np.all(np.diff(dz) >= 0.0)
    cfg, f = P.moist_terrain(24, 20, 16, a=500.0)     # a narrow hill: flat corners
    L = C.layout_for(cfg)
    ng, nz = L.ng, cfg.nz
    z = f[C.F_Z_ND][0]
    assert np.allclose(z[ng + nz], 15000.0, rtol=0, atol=1e-9)          # flat lid
    corner = z[ng:ng + nz + 1, ng, ng]                                   # far from the hill: h ~ 0
    assert corner[0] < 1e-3 and np.all(np.diff(corner) > 0)
    assert z[ng, ng + cfg.ny // 2, ng + cfg.nx // 2] > 400.0             # hill top near 500 m
    assert np.all(z[ng - 1] == 2.0 * z[ng] - z[ng + 1])                 # k = -1 extrapolated


def test_moist_terrain_case_runs():
    """The moist terrain case (stretched levels, WENO5 scalars, Kessler,
    Rayleigh layer) at a reduced size: finite, mass of dry air conserved,
    water species non-negative, rain reaching the ground is non-negative."""
    cfg, f = P.moist_terrain(24, 20, 16)
    o = Oracle(cfg)
    o.load(f)
    o.set_rayleigh(*P.rayleigh_profiles_uniform(cfg, f=f))
    L = C.layout_for(cfg)
    m0 = _integrals(o, cfg, L)
    o.step(0.0, 1.0, 3)
    m1 = _integrals(o, cfg, L)
    assert abs(m1[C.RHO] - m0[C.RHO]) <= 1e-13 * abs(m0[C.RHO])
    c = _valid(cfg, L, o.get(C.F_CONS))
    assert np.all(np.isfinite(c))
    for n in (C.RHOQ1, C.RHOQ2, C.RHOQ3):
        assert c[n].min() >= 0.0
    assert o.get(C.D_RAIN_ACCUM).min() >= 0.0
