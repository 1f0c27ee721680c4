"""Run configuration and field layout, mirroring include/erfx.h.

`Config` is a ctypes mirror of `erfx_config`; `default_config()` restates ERF's
defaults for the fields the dycore step reads (Source/DataStructs/ERF_DataStruct.H,
ERF_AdvStruct.H:281-296, ERF_DiffStruct.H, ERF_TurbStruct.H:170-179).
"""
import ctypes
import math

# ---- enums (include/erfx.h) -------------------------------------------------
BC_PERIODIC, BC_SLIP_WALL, BC_NO_SLIP_WALL, BC_SYMMETRY, BC_OUTFLOW, BC_INFLOW = range(6)
ADV_CENTERED2, ADV_UPWIND3, ADV_CENTERED4, ADV_UPWIND5, ADV_CENTERED6 = range(5)
# WENO family (scalars only): ERF_Interpolation_WENO.H, ERF_Interpolation_WENO_Z.H
ADV_WENO3, ADV_WENOZ3, ADV_WENOMZQ3, ADV_WENO5, ADV_WENOZ5 = range(5, 10)
MOLEC_NONE, MOLEC_CONSTANT, MOLEC_CONSTANT_ALPHA = range(3)
LES_NONE, LES_SMAGORINSKY = range(2)
RHO, RHOTHETA, RHOKE, RHOQKE, RHOSCALAR = range(5)
RHOQ1, RHOQ2, RHOQ3 = 5, 6, 7          # vapour, cloud water, rain (Kessler)
NCONS_DRY, NCONS_MOIST = 5, 8
COL_VARIANTS = 10   # ERFX_COL_VARIANTS: acoustic column kernel variants (9 = the tile form)
MOIST_NONE, MOIST_KESSLER, MOIST_KESSLER_NORAIN = range(3)

F_CONS, F_XVEL, F_YVEL, F_ZVEL, F_R0, F_P0, F_PI0 = range(7)
F_Z_ND = 7          # node heights of a terrain-following grid
PREC_FP64, PREC_MIXED = 0, 1
D_DETJ, D_AX, D_AY, D_AZ, D_Z_CC = 27, 28, 29, 30, 31
D_PRIM, D_PI_STAGE, D_FAST_COEFFS, D_EDDY = 16, 17, 18, 19
D_XVEL_OLD, D_YVEL_OLD, D_ZVEL_OLD = 20, 21, 22
D_XMOM_SRC, D_YMOM_SRC, D_ZMOM_SRC = 23, 24, 25
D_RAIN_ACCUM = 26
D_STATE_BASE = 32
S_OLD, S_NEW, S_SUM, S_SCRATCH, F_SLOW = range(5)
C_CONS, C_XMOM, C_YMOM, C_ZMOM = range(4)


def state_id(s, c):
    """Diagnostic id of component c (cons/xmom/ymom/zmom) of integrator state s."""
    return D_STATE_BASE + 4 * s + c


NCOMP = {F_CONS: None, D_PRIM: -1, D_FAST_COEFFS: 5, D_EDDY: 13}

# ---- constants (Source/ERF_Constants.H) -------------------------------------
R_d = 287.0
R_v = 461.505
Cp_d = 1004.5
p_0 = 1.0e5
Gamma = 1.4
CONST_GRAV = 9.81


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("prob_lo", ctypes.c_double * 3),
        ("bc", ctypes.c_int * 6),
        ("ncons", ctypes.c_int),
        ("dycore_horiz_adv", ctypes.c_int), ("dycore_vert_adv", ctypes.c_int),
        ("dryscal_horiz_adv", ctypes.c_int), ("dryscal_vert_adv", ctypes.c_int),
        ("dycore_horiz_upw_frac", ctypes.c_double), ("dycore_vert_upw_frac", ctypes.c_double),
        ("dryscal_horiz_upw_frac", ctypes.c_double), ("dryscal_vert_upw_frac", ctypes.c_double),
        ("molec_diff_type", ctypes.c_int),
        ("dynamic_viscosity", ctypes.c_double), ("rho0_trans", ctypes.c_double),
        ("alpha_T", ctypes.c_double), ("alpha_C", ctypes.c_double),
        ("les_type", ctypes.c_int),
        ("Cs", ctypes.c_double), ("Pr_t", ctypes.c_double), ("Sc_t", ctypes.c_double),
        ("use_surface_heat_flux", ctypes.c_int), ("surface_heat_flux", ctypes.c_double),
        ("use_gravity", ctypes.c_int), ("c_p", ctypes.c_double),
        ("use_coriolis", ctypes.c_int),
        ("coriolis_factor", ctypes.c_double), ("sinphi", ctypes.c_double), ("cosphi", ctypes.c_double),
        ("abl_geo_forcing", ctypes.c_double * 3),
        ("mri_ratio", ctypes.c_int),
        ("force_stage1_single_substep", ctypes.c_int),
        ("implicit_substepping", ctypes.c_int),
        ("rayleigh_damp_U", ctypes.c_int), ("rayleigh_damp_V", ctypes.c_int),
        ("rayleigh_damp_W", ctypes.c_int), ("rayleigh_damp_T", ctypes.c_int),
        ("rayleigh_dampcoef", ctypes.c_double), ("rayleigh_zdamp", ctypes.c_double),
        ("moisture_type", ctypes.c_int), ("buoyancy_type", ctypes.c_int),
        ("moistscal_horiz_adv", ctypes.c_int), ("moistscal_vert_adv", ctypes.c_int),
        ("moistscal_horiz_upw_frac", ctypes.c_double), ("moistscal_vert_upw_frac", ctypes.c_double),
        ("use_terrain", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("inflow_velocity", (ctypes.c_double * 3) * 6),
        ("inflow_density", ctypes.c_double * 6), ("inflow_theta", ctypes.c_double * 6),
        ("inflow_scalar", ctypes.c_double * 6), ("inflow_qv", ctypes.c_double * 6),
        ("inflow_qc", ctypes.c_double * 6),
        ("prob_hi", ctypes.c_double * 3),
        ("no_substepping", ctypes.c_int),
        ("wall_velocity", (ctypes.c_double * 3) * 6),
        ("abl_pressure_grad", ctypes.c_double * 3),
    ]


class DtControl(ctypes.Structure):
    """erfx_dt_control: ERF's ComputeDt / estTimeStep settings (ERF.cpp:27-30 defaults)."""
    _fields_ = [("cfl", ctypes.c_double), ("fixed_dt", ctypes.c_double), ("fixed_fast_dt", ctypes.c_double),
                ("fixed_mri_ratio", ctypes.c_int), ("change_max", ctypes.c_double),
                ("init_shrink", ctypes.c_double), ("stop_time", ctypes.c_double)]


def dt_control(cfl=0.8, fixed_dt=-1.0, fixed_fast_dt=-1.0, fixed_mri_ratio=0, change_max=1.1, init_shrink=1.0,
               stop_time=1.e100):
    return DtControl(cfl, fixed_dt, fixed_fast_dt, fixed_mri_ratio, change_max, init_shrink, stop_time)


class Layout(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("ng", ctypes.c_int),
                ("sy", ctypes.c_long), ("sz", ctypes.c_long), ("nelem", ctypes.c_long)]

    def shape3(self):
        """(k, j, i) extents of one component including ghosts and padding."""
        nzt = self.nelem // self.sz
        return (nzt, self.sz // self.sy, self.sy)


def default_config(nx, ny, nz, dx, dy, dz):
    c = Config()
    c.nx, c.ny, c.nz = nx, ny, nz
    c.dx, c.dy, c.dz = dx, dy, dz
    for d in range(3):
        c.prob_lo[d] = 0.0
    for o in range(6):
        c.bc[o] = BC_PERIODIC
    c.bc[2] = BC_SLIP_WALL
    c.bc[5] = BC_SLIP_WALL
    c.ncons = 5
    c.dycore_horiz_adv = c.dycore_vert_adv = ADV_UPWIND3
    c.dryscal_horiz_adv = c.dryscal_vert_adv = ADV_UPWIND3
    c.dycore_horiz_upw_frac = c.dycore_vert_upw_frac = 1.0
    c.dryscal_horiz_upw_frac = c.dryscal_vert_upw_frac = 1.0
    c.molec_diff_type = MOLEC_NONE
    c.dynamic_viscosity = 0.0
    c.rho0_trans = 1.0
    c.alpha_T = c.alpha_C = 0.0
    c.les_type = LES_NONE
    c.Cs = 0.0
    c.Pr_t = 1.0 / 3.0
    c.Sc_t = 1.0
    c.use_surface_heat_flux = 0
    c.surface_heat_flux = 0.0
    c.use_gravity = 1
    c.c_p = Cp_d
    c.use_coriolis = 0
    c.coriolis_factor = 2.0 * 2.0 * math.pi / 86400.0
    c.sinphi = 1.0
    c.cosphi = 0.0
    for d in range(3):
        c.abl_geo_forcing[d] = 0.0
    c.mri_ratio = 6
    c.force_stage1_single_substep = 1
    c.implicit_substepping = 1
    c.rayleigh_damp_U = c.rayleigh_damp_V = c.rayleigh_damp_W = c.rayleigh_damp_T = 0
    c.rayleigh_dampcoef = 0.2
    c.rayleigh_zdamp = 500.0
    c.moisture_type = MOIST_NONE
    c.buoyancy_type = 1
    c.moistscal_horiz_adv = c.moistscal_vert_adv = ADV_WENO3
    c.moistscal_horiz_upw_frac = c.moistscal_vert_upw_frac = 1.0
    c.use_terrain = 0
    c.precision = PREC_FP64
    return c


def set_moisture(c, moisture_type=MOIST_KESSLER, buoyancy_type=None):
    """Switch a config to a Kessler model: 8 cons comps and ERF's moist default
    buoyancy (type 2 with rain, type 1 without; ERF_DataStruct.H:126-137)."""
    c.moisture_type = moisture_type
    c.ncons = NCONS_MOIST
    if buoyancy_type is None:
        buoyancy_type = 2 if moisture_type == MOIST_KESSLER else 1
    c.buoyancy_type = buoyancy_type
    return c


def ghost_cells(cfg):
    """ERF::ComputeGhostCells (Source/ERF.H:1081-1125) + 1 for cons.  As in ERF,
    the moist vertical type counts only for WENO5/5Z."""
    wide = (ADV_UPWIND5, ADV_CENTERED6, ADV_WENO5, ADV_WENOZ5)
    g = 3 if any(t in wide for t in (cfg.dycore_horiz_adv, cfg.dycore_vert_adv,
                                      cfg.dryscal_horiz_adv, cfg.dryscal_vert_adv,
                                      cfg.moistscal_horiz_adv)) else 2
    if cfg.moistscal_vert_adv in (ADV_WENO5, ADV_WENOZ5):
        g = 3
    return g + 1


def layout_for(cfg):
    ng = ghost_cells(cfg)
    L = Layout()
    L.nx, L.ny, L.nz, L.ng = cfg.nx, cfg.ny, cfg.nz, ng
    row = cfg.nx + 1 + 2 * ng
    L.sy = (row + 15) // 16 * 16
    L.sz = L.sy * (cfg.ny + 1 + 2 * ng)
    L.nelem = L.sz * (cfg.nz + 1 + 2 * ng)
    return L


def ncomp_of(cfg, fid):
    if fid == F_CONS or (fid >= D_STATE_BASE and (fid - D_STATE_BASE) % 4 == 0):
        return cfg.ncons
    if fid == D_PRIM:
        return cfg.ncons - 1
    if fid == D_FAST_COEFFS:
        return 5
    if fid == D_EDDY:
        return 13
    return 1


def field_size(cfg, fid):
    """Doubles a host buffer for field fid holds."""
    L = layout_for(cfg)
    if fid == D_RAIN_ACCUM:
        return L.sz
    return ncomp_of(cfg, fid) * L.nelem
