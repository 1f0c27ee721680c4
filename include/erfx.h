/*
 * erfx.h -- C ABI of the MI355X-native compressible dycore step.
 *
 * This is the drop-in boundary for ERF's multirate (MRI) RK3 dycore step.
 * ERF wires its step through std::function hooks on
 * MRISplitIntegrator<Vector<MultiFab>> (Source/TimeIntegration/ERF_MRI.H:137-180);
 * each hook below replaces one of those callees with plain pointers and sizes.
 *
 * Data model
 * ----------
 * One box per context.  Every field of the box (cell-centred or face-centred)
 * is stored with the SAME layout `erfx_layout`: i fastest, then j, then k, then
 * component, with `ng` ghost cells on every side and room for the extra face
 * in each direction:
 *
 *     elem(i,j,k,n) = p[(i+ng) + (j+ng)*sy + (k+ng)*sz + n*nelem]
 *     valid cells  i in [0,nx-1]  (x-faces: [0,nx]), likewise j and k.
 *
 * This is AMReX's Array4 order (ERF_IndexDefines.H, component-major) with one
 * common stride set so fused kernels can address every field with one index.
 * A "state" is ERF's T = Vector<MultiFab>{cons, xmom, ymom, zmom}.
 *
 * Conventions
 * -----------
 *  - All pointers passed to hook functions are device pointers, laid out as
 *    `erfx_layout` reports.  The caller owns them.
 *  - Every call is stream-ordered on the context's stream and returns an int
 *    status (0 = ok, < 0 = error); erfx_last_error() gives the message.  No
 *    C++ exception crosses the ABI.  A context is not re-entrant.
 *  - Arithmetic is IEEE FP64 with no contraction, matching the reference's
 *    operation order (ERF computes in amrex::Real = double).
 */
#ifndef ERFX_H
#define ERFX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    ERFX_OK              = 0,
    ERFX_ERR_ARG         = -1,   /* bad argument (null pointer, bad enum, bad size) */
    ERFX_ERR_LAYOUT      = -2,   /* layout does not match what the kernels assume */
    ERFX_ERR_HIP         = -3,   /* a HIP runtime call failed                     */
    ERFX_ERR_UNSUPPORTED = -4,   /* configuration not implemented                 */
    ERFX_ERR_STATE       = -5,   /* physically invalid state (e.g. rho*theta <= 0) */
    ERFX_ERR_NODEVICE    = -6    /* no usable GPU                                  */
};

/* ---- boundary types per domain face (ERF_BC, Source/ERF_IndexDefines.H) -- */
enum {
    ERFX_BC_PERIODIC     = 0,
    ERFX_BC_SLIP_WALL    = 1,
    ERFX_BC_NO_SLIP_WALL = 2,
    ERFX_BC_SYMMETRY     = 3,
    ERFX_BC_OUTFLOW      = 4,
    /* ERF's "Inflow" (ERF_init_bcs.cpp:100-178, 345-361, 544-580): velocity
     * ext_dir with the face's inflow_velocity; cons ext_dir with density,
     * density*theta, density*scalar, density*qv, density*qc (0 for the other
     * components), or foextrap for every cons component when density is 0.
     * Lateral faces only, with the other lateral direction periodic. */
    ERFX_BC_INFLOW       = 5
};

/* ---- advection schemes (AdvType, Source/DataStructs/ERF_AdvStruct.H) ----
 * The WENO family (ERF_Interpolation_WENO.H, ERF_Interpolation_WENO_Z.H) is
 * for scalars only; ERF applies the horizontal WENO type in z as well. */
enum {
    ERFX_ADV_CENTERED2 = 0,
    ERFX_ADV_UPWIND3   = 1,
    ERFX_ADV_CENTERED4 = 2,
    ERFX_ADV_UPWIND5   = 3,
    ERFX_ADV_CENTERED6 = 4,
    ERFX_ADV_WENO3     = 5,
    ERFX_ADV_WENOZ3    = 6,
    ERFX_ADV_WENOMZQ3  = 7,
    ERFX_ADV_WENO5     = 8,
    ERFX_ADV_WENOZ5    = 9
};

/* ---- diffusion (MolecDiffType / LESType, ERF_DiffStruct.H, ERF_TurbStruct.H) */
enum { ERFX_MOLEC_NONE = 0, ERFX_MOLEC_CONSTANT = 1, ERFX_MOLEC_CONSTANT_ALPHA = 2 };
enum { ERFX_LES_NONE = 0, ERFX_LES_SMAGORINSKY = 1 };

/* ---- conserved-variable components (ERF_IndexDefines.H:13-50) ---------- */
enum {
    ERFX_RHO = 0, ERFX_RHOTHETA = 1, ERFX_RHOKE = 2, ERFX_RHOQKE = 3,
    ERFX_RHOSCALAR = 4, ERFX_NCONS_DRY = 5,
    /* moisture species of the Kessler models: vapour, cloud water, rain */
    ERFX_RHOQ1 = 5, ERFX_RHOQ2 = 6, ERFX_RHOQ3 = 7, ERFX_NCONS_MOIST = 8
};

/* ---- moisture models (MoistureType, ERF_DataStruct.H) ------------------ */
enum { ERFX_MOIST_NONE = 0, ERFX_MOIST_KESSLER = 1, ERFX_MOIST_KESSLER_NORAIN = 2 };

/* ---- run configuration: the subset of SolverChoice the step reads ----- */
typedef struct erfx_config {
    int    nx, ny, nz;            /* valid cells                                     */
    double dx, dy, dz;            /* uniform spacing [m]                             */
    double prob_lo[3];            /* domain origin [m]                               */
    int    bc[6];                 /* xlo, ylo, zlo, xhi, yhi, zhi (AMReX Orientation) */
    int    ncons;                 /* 5 (dry) or 8 (Kessler: + RhoQ1..RhoQ3)          */

    /* advection (ERF_AdvStruct.H) */
    int    dycore_horiz_adv, dycore_vert_adv;
    int    dryscal_horiz_adv, dryscal_vert_adv;
    double dycore_horiz_upw_frac, dycore_vert_upw_frac;
    double dryscal_horiz_upw_frac, dryscal_vert_upw_frac;

    /* diffusion (ERF_DiffStruct.H / ERF_TurbStruct.H) */
    int    molec_diff_type;
    double dynamic_viscosity, rho0_trans, alpha_T, alpha_C;
    int    les_type;
    double Cs, Pr_t, Sc_t;
    /* prescribed surface flux of rho*theta at the bottom face, used in place of
     * the diffusive z-flux at k = 0 (the explicit-MOST branch of
     * DiffusionSrcForState_N, ERF_DiffusionSrcForState_N.cpp:300-301) */
    int    use_surface_heat_flux;
    double surface_heat_flux;     /* kg m^-2 s^-1 K */

    /* gravity / thermodynamics */
    int    use_gravity;
    double c_p;                   /* rdOcp = R_d / c_p                               */

    /* forcing (ERF_make_mom_sources.cpp) */
    int    use_coriolis;
    double coriolis_factor, sinphi, cosphi;
    double abl_geo_forcing[3];

    /* time integration (ERF_MRI.H, ERF_TI_fast_rhs_fun.H) */
    int    mri_ratio;             /* fixed_mri_dt_ratio, even, > 1                   */
    int    force_stage1_single_substep;
    int    implicit_substepping;  /* 1: beta_s = 0.1, 0: beta_s = -1 (explicit)      */

    /* Rayleigh damping layer below the domain top (ERF_DataStruct.H:567-573;
     * ERF_make_sources.cpp:203-225, ERF_make_mom_sources.cpp:247-296): rate
     * dampcoef * sin^2(pi/2 * zfrac), zfrac = 1 - (ztop - z) / zdamp, towards
     * the profiles given to erfx_set_rayleigh; ztop is the domain top
     * (ERF_init_rayleigh.cpp:17). */
    int    rayleigh_damp_U, rayleigh_damp_V, rayleigh_damp_W, rayleigh_damp_T;
    double rayleigh_dampcoef;     /* 1/s, default 0.2                                */
    double rayleigh_zdamp;        /* m, default 500                                  */

    /* moisture (ERF_DataStruct.H:100-140; Kessler, ERF_Kessler.H).  With
     * moisture the three species ride in cons comps 5..7; the dycore divides
     * the pressure gradients by 1 + qv + qc and evaluates p' with qv. */
    int    moisture_type;         /* ERFX_MOIST_*                                    */
    /* make_buoyancy expression (ERF_make_buoyancy.cpp:32-380): 1 = density
     * perturbation (the only dry form here); 2, 3, 4 = the moist forms with
     * plane averages.  ERF picks 2 for Kessler with rain, 1 otherwise. */
    int    buoyancy_type;
    int    moistscal_horiz_adv, moistscal_vert_adv;          /* WENO3 in ERF */
    double moistscal_horiz_upw_frac, moistscal_vert_upw_frac;

    /* terrain-following grid (use_terrain, static terrain; ERF_DataStruct.H):
     * the node heights come from erfx_set_terrain, the step uses
     * erf_fast_rhs_T's acoustic substep and the metric terms of the slow RHS.
     * dz is the computational spacing (detJ = height of a cell / dz). */
    int    use_terrain;

    /* ERFX_PREC_FP64 (default) or ERFX_PREC_MIXED: the mixed form evaluates
     * the face values of the advected scalars and of theta in FP32, as
     * perturbations about an FP64 reference taken at each face, and keeps
     * the flux products, their divergence, momentum advection and everything
     * else (acoustic solve, averaged mass fluxes, base state, sums) in FP64.
     * No counterpart in the reference, whose results the FP64 form reproduces. */
    int    precision;

    /* values of ERFX_BC_INFLOW faces, per orientation (xlo, ylo, zlo, xhi,
     * yhi, zhi): the inputs xlo.velocity, xlo.density, xlo.theta, xlo.scalar,
     * xlo.qv, xlo.qc of ERF_init_bcs.cpp:100-178 */
    double inflow_velocity[6][3];
    double inflow_density[6];
    double inflow_theta[6];
    double inflow_scalar[6];
    double inflow_qv[6];
    double inflow_qc[6];

    /* upper corner of the domain [m] (geometry.prob_hi).  The Rayleigh layer's
     * ztop is prob_hi[2] (ERF_init_rayleigh.cpp:17); when prob_hi[2] <=
     * prob_lo[2] (e.g. all zero) the top is taken as prob_lo[2] + nz * dz. */
    double prob_hi[3];

    /* substepping_type = None: each RK stage advances the fast variables with
     * the slow right-hand side alone, S_sum = S_old + dt_stage * F_slow, and
     * no acoustic substeps (ERF_MRI.H:255-320, ERF_TI_no_substep_fun.H:1-146);
     * mri_ratio is then not used. */
    int    no_substepping;

    /* velocity of a NoSlipWall face (the <face>.velocity input,
     * ERF_init_bcs.cpp:184-200): the ext_dir values of the three velocity
     * components on that face; the normal component is forced to 0 as ERF
     * does.  Per orientation (xlo, ylo, zlo, xhi, yhi, zhi). */
    double wall_velocity[6][3];

    /* constant driving pressure gradient of the "PressureGradient" ABL driver
     * (erf.abl_pressure_grad, ERF_DataStruct.H:268-270), subtracted from the
     * momentum right-hand sides (ERF_slow_rhs_pre.cpp:630, 679, 739) */
    double abl_pressure_grad[3];
} erfx_config;

enum { ERFX_PREC_FP64 = 0, ERFX_PREC_MIXED = 1 };

/* ---- layout shared by every field of a context ------------------------- */
typedef struct erfx_layout {
    int  nx, ny, nz, ng;
    long sy, sz;                  /* strides in elements                             */
    long nelem;                   /* elements per component (>= (nz+1+2ng)*sz)       */
} erfx_layout;

/* ---- a state T = {cons, xmom, ymom, zmom} (device pointers) ------------ */
typedef struct erfx_state {
    double* cons;                 /* ncons components, nelem apart                  */
    double* xmom;
    double* ymom;
    double* zmom;
} erfx_state;

typedef struct erfx_ctx erfx_ctx;

/* Thread-local text of the last error on this thread. */
const char* erfx_last_error(void);

/* Fill `cfg` with ERF's defaults for the fields the step reads
 * (ERF_DataStruct.H, ERF_AdvStruct.H:90-99,281-286, ERF_DiffStruct.H). */
void erfx_config_defaults(erfx_config* cfg);

/* Ghost width ERF would use for this configuration, plus one for cons
 * (ERF::ComputeGhostCells, ERF.H:1081-1125; ERF_make_new_arrays.cpp:112-113). */
int erfx_ghost_cells(const erfx_config* cfg);

/* Create a context: validates cfg, picks the layout, allocates the workspace
 * on the current device.  `stream` is a hipStream_t or NULL (default stream). */
int erfx_ctx_create(const erfx_config* cfg, void* stream, erfx_ctx** out);
int erfx_ctx_destroy(erfx_ctx* ctx);
int erfx_ctx_layout(const erfx_ctx* ctx, erfx_layout* out);

/* Base state {r0, p0, pi0} on cells with ghosts (BaseState, ERF_IndexDefines.H:63-71).
 * The pointers are kept and read by later calls. */
int erfx_set_base_state(erfx_ctx* ctx, const double* r0, const double* p0, const double* pi0);

/* Node heights z_phys_nd of a terrain-following grid (use_terrain; the
 * reference's z_phys_nd, ERF_TerrainMetrics.H:21 init_terrain_grid): a device
 * field of the layout holding every node, ghosts included (periodic images on
 * periodic axes), copied.  The metrics detJ, ax, ay, az and z_cc (make_J,
 * make_areas, make_zcc; ERF_Utils.H:27-44) are made from it on the context's
 * stream.  Until it is called the grid is flat (z = prob_lo[2] + k dz). */
int erfx_set_terrain(erfx_ctx* ctx, const double* z_nd);

/* Reference profiles of the Rayleigh damping (ERF's Rayleigh::ubar, vbar,
 * wbar, thetabar; ERF::initRayleigh, ERF_init_rayleigh.cpp:14-45): host arrays
 * of nz+1 values indexed by k (cell k for ubar, vbar and thetabar, z-face k for
 * wbar), copied.  Needed when any rayleigh_damp_* flag is set. */
int erfx_set_rayleigh(erfx_ctx* ctx, const double* ubar, const double* vbar, const double* wbar,
                      const double* thetabar);

/* cons_to_prim (ERF_TI_utils.H:4-45): prim(n) = cons(n+1)/rho on the box grown
 * by `ng`; pi_stage = Exner(rho*theta) on the box grown by 1. */
int erfx_cons_to_prim(erfx_ctx* ctx, const double* cons, double* prim, double* pi_stage, int ng);

/* apply_bcs (ERF_TI_utils.H:53-150): FillIntermediatePatch at level 0 on a state:
 * momentum -> velocity on the valid region, physical/periodic ghost fills of cons
 * and velocity, velocity -> momentum on valid + 1 ghost (ERF_FillIntermediatePatch.cpp:27-261). */
int erfx_apply_bcs(erfx_ctx* ctx, erfx_state* S, double* xvel, double* yvel, double* zvel,
                   int ng_cons, int ng_vel, int fast_only, int vel_and_mom_synced);

/* make_fast_coeffs (ERF_make_fast_coeffs.cpp:27-294, flat branch): writes the
 * five z-face coefficient fields A, 1/B, C, P, Q into `fast_coeffs` (5 comps). */
int erfx_make_fast_coeffs(erfx_ctx* ctx, double* fast_coeffs, const erfx_state* stage,
                          const double* prim, const double* pi_stage, double dtau, double beta_s);

/* erf_fast_rhs_N (ERF_fast_rhs_N.cpp:36-614; ERF_TI_fast_headers.H:22-41):
 * one acoustic substep on a flat grid.  For step > 0 the reference passes
 * S_prev == S_data (same memory); pass the same pointers here. */
int erfx_fast_rhs_N(erfx_ctx* ctx, int step, int nrk,
                    const erfx_state* slow_rhs, const erfx_state* prev,
                    const erfx_state* stage, const double* prim, const double* pi_stage,
                    const double* fast_coeffs, erfx_state* data, erfx_state* scratch,
                    double dtau, double beta_s, double facinv);

/* erf_fast_rhs_N with no device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under stream or graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as erfx_fast_rhs_N, with the configuration and the
 * stream (a hipStream_t, NULL for the default stream) in place of a context:
 * the substep's work fields live in a static device array of 64 MiB per GPU
 * (held until the process ends) instead of a context's workspace, looked up
 * once per device on the first call made with that device current.  Boxes
 * whose 5 work fields of the layout exceed it (about 96 x 96 x 96 cells and
 * up) are refused with ERFX_ERR_ARG before any device work.  Calls on one GPU
 * share the array: two calls that overlap on one GPU race.  Flat grids only
 * (a terrain grid returns ERFX_ERR_STATE). */
int erfx_fast_rhs_N_noalloc(const erfx_config* cfg, void* stream, int step, int nrk,
                            const erfx_state* slow_rhs, const erfx_state* prev,
                            const erfx_state* stage, const double* prim, const double* pi_stage,
                            const double* fast_coeffs, erfx_state* data, erfx_state* scratch,
                            double dtau, double beta_s, double facinv);

/* erf_fast_rhs_T (ERF_fast_rhs_T.cpp:39-704; ERF_TI_fast_headers.H:47-71):
 * one acoustic substep on a static terrain grid (use_terrain), with the
 * arguments of erfx_fast_rhs_N. */
int erfx_fast_rhs_T(erfx_ctx* ctx, int step, int nrk,
                    const erfx_state* slow_rhs, const erfx_state* prev,
                    const erfx_state* stage, const double* prim, const double* pi_stage,
                    const double* fast_coeffs, erfx_state* data, erfx_state* scratch,
                    double dtau, double beta_s, double facinv);

/* Strain + Smagorinsky eddy viscosity, once per step (ERF_advance_dycore.cpp:114-223;
 * ComputeStrain_N, ComputeTurbulentViscosityLES).  `eddy` holds 13 components
 * (EddyDiff::NumDiffs, ERF_IndexDefines.H:152-169), `hfx3` the vertical SGS heat flux. */
int erfx_turbulent_viscosity(erfx_ctx* ctx, const double* cons, const double* xvel,
                             const double* yvel, const double* zvel, double* eddy, double* hfx3);

/* erf_slow_rhs_pre (ERF_slow_rhs_pre.cpp:67-774) preceded by make_mom_sources
 * (ERF_make_mom_sources.cpp:40-450) as in ERF_TI_slow_rhs_fun.H:216-241:
 * slow RHS of rho, rho*theta and the momenta; seeds the averaged momenta in
 * scratch.  `eddy`/`hfx3` may be NULL when no LES model is configured. */
int erfx_slow_rhs_pre(erfx_ctx* ctx, int nrk, double slow_dt,
                      erfx_state* rhs, const erfx_state* data, const double* prim,
                      erfx_state* scratch, const double* xvel, const double* yvel,
                      const double* zvel, const double* eddy, double* hfx3);

/* erf_slow_rhs_post (ERF_slow_rhs_post.cpp:49-559): slow scalars with the
 * time-averaged momenta, then S_new <- S_data. */
int erfx_slow_rhs_post(erfx_ctx* ctx, int nrk, double slow_dt,
                       erfx_state* rhs, const erfx_state* old, erfx_state* snew,
                       erfx_state* data, const double* prim, const erfx_state* scratch,
                       const double* xvel, const double* yvel, const double* eddy);

/* Kessler warm rain, once per step after the dycore (ERF::advance_microphysics,
 * ERF_advance_microphysics.cpp:5-16): Copy_State_to_Micro
 * (ERF_Init_Kessler.cpp:64-99), Kessler::AdvanceKessler with the step's dt
 * (ERF_Kessler.cpp:11-250; sedimentation only for ERFX_MOIST_KESSLER),
 * Copy_Micro_to_State and its FillBoundary (ERF_Update_Kessler.cpp:14-40).
 * `rain_accum` (may be NULL) is the k = 0 plane of the layout,
 * elem(i,j) = p[(i+ng) + (j+ng)*sy], and accumulates surface rain in mm.
 * Synchronises the stream at return: ERF aborts when a temperature leaves the
 * saturation fit's range (below -85 C), reported here as ERFX_ERR_STATE. */
int erfx_kessler_advance(erfx_ctx* ctx, double* cons, double dt, double* rain_accum);

/* ---- hooks on Array4-shaped views ---------------------------------------- *
 * ERF holds each field of a box in its own FArrayBox: cons with
 * ComputeGhostCells + 1 ghost cells, the face-centred momenta and velocities
 * with ComputeGhostCells, each with its own strides
 * (Source/ERF_make_new_arrays.cpp:112-130), reached through amrex::Array4.  A
 * view describes one such field of one box: element (i, j, k, n) is
 *     p[(i - lo[0]) + (j - lo[1]) * js + (k - lo[2]) * ks + n * ns]
 * for lo <= (i, j, k) <= hi, with the indices of the valid box (cells 0 ..
 * n-1; faces 0 .. n along a nodal direction), so lo is minus the ghost width.
 * ixtype[d] = 1 where the field is nodal along d.  The _v hooks take views in
 * place of the common-layout pointers: a view already in the common layout
 * is used in place, any other is gathered into the context's staging field
 * of that layout and scattered back after the hook when the hook writes it
 * (views of the same memory share one staging field, as ERF's S_prev ==
 * S_data).  Each view must cover the valid box grown by the ghost width its
 * role needs (erfx_check_view), else the hook returns ERFX_ERR_ARG before any
 * device work.  Results are those of the pointer hooks, bit for bit. */
typedef struct erfx_view {
    double* p;                    /* element (lo[0], lo[1], lo[2], 0)               */
    int  lo[3], hi[3];            /* index box of the allocation                    */
    long js, ks, ns;              /* strides in elements (i stride 1)               */
    int  ncomp;
    int  ixtype[3];               /* 1 = nodal along that direction                  */
} erfx_view;
typedef struct erfx_state_view { erfx_view cons, xmom, ymom, zmom; } erfx_state_view;

/* what a view stands for: components, index type, ghost width needed
 * (ng = erfx_ghost_cells(cfg)):
 *   CONS ncons comps, cells, ng;  XMOM/YMOM/ZMOM and XVEL/YVEL/ZVEL 1 comp, nodal
 *   in x / y / z, ng - 1;  PRIM ncons - 1 comps, cells, ng;  PI 1 comp, cells, 1;
 *   FAST_COEFFS 5 comps, nodal in z, 0;  EDDY 13 comps, cells, 1 */
enum {
    ERFX_ROLE_CONS = 0, ERFX_ROLE_XMOM = 1, ERFX_ROLE_YMOM = 2, ERFX_ROLE_ZMOM = 3,
    ERFX_ROLE_XVEL = 4, ERFX_ROLE_YVEL = 5, ERFX_ROLE_ZVEL = 6,
    ERFX_ROLE_PRIM = 7, ERFX_ROLE_PI = 8, ERFX_ROLE_FAST_COEFFS = 9, ERFX_ROLE_EDDY = 10
};

/* Check a view against a role for boxes of this configuration (no device
 * needed): ERFX_OK, or ERFX_ERR_ARG with the reason in erfx_last_error(). */
int erfx_check_view(const erfx_config* cfg, int role, const erfx_view* v);

/* The hooks above on views (same meaning, same argument order) */
int erfx_cons_to_prim_v(erfx_ctx* ctx, const erfx_view* cons, erfx_view* prim, erfx_view* pi_stage, int ng);
int erfx_apply_bcs_v(erfx_ctx* ctx, erfx_state_view* S, erfx_view* xvel, erfx_view* yvel, erfx_view* zvel,
                     int ng_cons, int ng_vel, int fast_only, int vel_and_mom_synced);
int erfx_make_fast_coeffs_v(erfx_ctx* ctx, erfx_view* fast_coeffs, const erfx_state_view* stage,
                            const erfx_view* prim, const erfx_view* pi_stage, double dtau, double beta_s);
int erfx_fast_rhs_N_v(erfx_ctx* ctx, int step, int nrk, const erfx_state_view* slow_rhs,
                      const erfx_state_view* prev, const erfx_state_view* stage, const erfx_view* prim,
                      const erfx_view* pi_stage, const erfx_view* fast_coeffs, erfx_state_view* data,
                      erfx_state_view* scratch, double dtau, double beta_s, double facinv);
int erfx_fast_rhs_T_v(erfx_ctx* ctx, int step, int nrk, const erfx_state_view* slow_rhs,
                      const erfx_state_view* prev, const erfx_state_view* stage, const erfx_view* prim,
                      const erfx_view* pi_stage, const erfx_view* fast_coeffs, erfx_state_view* data,
                      erfx_state_view* scratch, double dtau, double beta_s, double facinv);
int erfx_turbulent_viscosity_v(erfx_ctx* ctx, const erfx_view* cons, const erfx_view* xvel,
                               const erfx_view* yvel, const erfx_view* zvel, erfx_view* eddy);
/* eddy may be NULL when no LES model is configured */
int erfx_slow_rhs_pre_v(erfx_ctx* ctx, int nrk, double slow_dt, erfx_state_view* rhs,
                        const erfx_state_view* data, const erfx_view* prim, erfx_state_view* scratch,
                        const erfx_view* xvel, const erfx_view* yvel, const erfx_view* zvel,
                        const erfx_view* eddy);
int erfx_slow_rhs_post_v(erfx_ctx* ctx, int nrk, double slow_dt, erfx_state_view* rhs,
                         const erfx_state_view* old, erfx_state_view* snew, erfx_state_view* data,
                         const erfx_view* prim, const erfx_state_view* scratch, const erfx_view* eddy);
/* without the surface-rain accumulator */
int erfx_kessler_advance_v(erfx_ctx* ctx, erfx_view* cons, double dt);

/* ---- learned stand-in for the Kessler cell tendencies ------------------- *
 * No counterpart in ERF: a multilayer perceptron 8 -> 128 -> 128 -> 4 with GELU
 * activations, BF16 weights and activations, FP32 accumulation, trained on the
 * GPU by the library's own forward, backward and Adam kernels on samples of
 * Kessler's own tendencies (erfx_kessler_samples).  Inputs per cell:
 * (T [K], p [mbar], rho, qv, qc, qr, w [m/s, cell centre], dt [s]); outputs:
 * the local tendencies of one step (dqv, dqc, dqr, dtheta), sedimentation
 * excluded.  Parameters (FP32 master copy, ERFX_MLP_NPARAM floats):
 * W1 [8][128], b1 [128], W2 [128][128], b2 [128], W3 [128][4], b3 [4],
 * row-major; a layer computes z_u = b_u + sum_k h_k W[k][u]. */
typedef struct erfx_mlp erfx_mlp;
enum { ERFX_MLP_NIN = 8, ERFX_MLP_NHID = 128, ERFX_MLP_NOUT = 4, ERFX_MLP_NPARAM = 18180 };

/* `capacity`: largest number of samples per training pass. */
int erfx_mlp_create(int capacity, erfx_mlp** out);
int erfx_mlp_destroy(erfx_mlp* m);
int erfx_mlp_set_params(erfx_mlp* m, const float* host_params);
int erfx_mlp_get_params(erfx_mlp* m, float* host_params);
/* Feature scaling: the network sees (x - x_mean) / x_scale and produces
 * y / y_scale; x_lo, x_hi bound the raw inputs it was trained on (outside,
 * Kessler itself is used).  8, 8, 8, 8 and 4 host doubles. */
int erfx_mlp_set_scaling(erfx_mlp* m, const double* x_mean, const double* x_scale, const double* x_lo,
                         const double* x_hi, const double* y_scale);
/* Network on device arrays of scaled values: x [n][8] -> y [n][4] (FP32). */
int erfx_mlp_forward(erfx_mlp* m, const float* x, float* y, int n);
/* Gradient of the mean squared error over n <= capacity samples (device x
 * [n][8], targets t [n][4], scaled) into grad_host (ERFX_MLP_NPARAM floats);
 * *loss = the mean squared error.  Synchronous. */
int erfx_mlp_gradient(erfx_mlp* m, const float* x, const float* t, int n, float* grad_host, double* loss);
/* One Adam step (beta 0.9, 0.999; eps 1e-8) on a minibatch; *loss as above,
 * before the update.  Synchronous when loss is non-NULL. */
int erfx_mlp_train_step(erfx_mlp* m, const float* x, const float* t, int n, double lr, double* loss);

/* Training rows from Kessler itself (rain model): for every cell whose local
 * processes are active (supersaturated, cloudy, or rainy in subsaturated air;
 * elsewhere Kessler's local tendencies are exactly 0), one row of 12 floats
 * (T, p, rho, qv, qc, qr, w, dt, dqv, dqc, dqr, dtheta) on the device, appended
 * in no fixed order; *count = active cells found (rows written: min(count,
 * max_rows)).  cons and zvel as before the step's microphysics. Synchronous. */
int erfx_kessler_samples(erfx_ctx* ctx, const double* cons, const double* zvel, double dt, float* rows,
                         long max_rows, long* count);
/* The microphysics step of erfx_kessler_advance (rain model) with the
 * network's tendencies for the active cells inside its training range;
 * inactive cells get Kessler's exact 0, active cells outside the range
 * Kessler's own update.  The network's (dqv, dqc, dqr) are projected in FP64
 * to conserve qv + qc + qr in each cell and keep every species >= 0; theta
 * follows the latent heat of the vapour change, as in Kessler
 * (ERF_Kessler.cpp:172-173); sedimentation is Kessler's flux form.
 * *n_mlp, *n_fallback (may be NULL) count the two kinds of active cells. */
int erfx_kessler_mlp_advance(erfx_ctx* ctx, erfx_mlp* m, double* cons, const double* zvel, double dt,
                             double* rain_accum, long* n_mlp, long* n_fallback);

/* ---- whole-step driver ------------------------------------------------- *
 * ERF::Advance + advance_dycore + MRISplitIntegrator::advance on one box
 * (ERF_Advance.cpp:20-233, ERF_advance_dycore.cpp:37-305, ERF_MRI.H:187-369).
 * The sim owns every buffer in HBM; the persistent state is {cons, xvel, yvel,
 * zvel} like ERF's vars_new. */
typedef struct erfx_sim erfx_sim;

enum {
    ERFX_F_CONS = 0, ERFX_F_XVEL = 1, ERFX_F_YVEL = 2, ERFX_F_ZVEL = 3,
    ERFX_F_R0 = 4, ERFX_F_P0 = 5, ERFX_F_PI0 = 6,
    ERFX_F_Z_ND = 7,   /* node heights z_phys_nd (use_terrain only); uploading them
                          recomputes detJ, ax, ay, az and z_cc */
    ERFX_F_NUM = 8
};

/* Diagnostic ids: the integrator's internal buffers, for hook-level testing.
 * State buffers are ERFX_D_STATE_BASE + 4*s + c with s = 0 S_old, 1 S_new,
 * 2 S_sum, 3 S_scratch, 4 F_slow (ERF_MRI.H:67-83) and c = 0 cons, 1 xmom,
 * 2 ymom, 3 zmom. */
enum {
    ERFX_D_PRIM = 16, ERFX_D_PI_STAGE = 17, ERFX_D_FAST_COEFFS = 18, ERFX_D_EDDY = 19,
    ERFX_D_XVEL_OLD = 20, ERFX_D_YVEL_OLD = 21, ERFX_D_ZVEL_OLD = 22,
    ERFX_D_XMOM_SRC = 23, ERFX_D_YMOM_SRC = 24, ERFX_D_ZMOM_SRC = 25,
    ERFX_D_RAIN_ACCUM = 26,       /* Kessler surface rain [mm], one k = 0 plane (sz doubles) */
    /* terrain metrics (use_terrain): make_J, make_areas, make_zcc
       (ERF_TerrainMetrics.cpp:583-715) */
    ERFX_D_DETJ = 27, ERFX_D_AX = 28, ERFX_D_AY = 29, ERFX_D_AZ = 30, ERFX_D_Z_CC = 31,
    ERFX_D_STATE_BASE = 32
};

/* Layout every context built from `cfg` uses (no device needed):
 * ng = erfx_ghost_cells(cfg), sy = roundup(nx+1+2ng, 16), sz = sy*(ny+1+2ng),
 * nelem = sz*(nz+1+2ng). */
int  erfx_layout_for(const erfx_config* cfg, erfx_layout* out);

int  erfx_sim_create(const erfx_config* cfg, erfx_sim** out);
/* device pointer of a sim buffer (ERFX_F_* or ERFX_D_*), NULL if unknown */
double* erfx_sim_field_ptr(erfx_sim* sim, int id);
/* the hook context the sim drives (same stream, same layout) */
erfx_ctx* erfx_sim_ctx(erfx_sim* sim);
int  erfx_sim_destroy(erfx_sim* sim);
int  erfx_sim_layout(const erfx_sim* sim, erfx_layout* out);
/* number of doubles a host buffer for field `id` must hold (ncomp * nelem) */
long erfx_sim_field_size(const erfx_sim* sim, int id);
int  erfx_sim_upload(erfx_sim* sim, int id, const double* host);
int  erfx_sim_download(erfx_sim* sim, int id, double* host);
/* Advance `nsteps` steps of size dt from time t0 (synchronous at return). */
int  erfx_sim_step(erfx_sim* sim, double t0, double dt, int nsteps);
/* Same, asynchronous on the sim's stream (for timing with events). */
int  erfx_sim_step_async(erfx_sim* sim, double t0, double dt, int nsteps);
int  erfx_sim_sync(erfx_sim* sim);
/* Replay the step sequence from a hipGraph (no counterpart in the reference;
 * for small boxes, where launch overhead dominates).  When on, erfx_sim_step*
 * captures two whole steps (the state buffers alternate, so two steps return
 * every pointer to its place) into a graph on first use, re-captures when dt
 * or the MRI ratio changes, or the stage time steps of the reference's
 * MRI loop ((t + dt/3) - t etc., which round differently as t grows) differ
 * from the captured ones, and replays it; an odd last step runs directly.
 * The results are those of the direct path.  Not used while the learned
 * microphysics, row sampling or the kernel timer is on. */
int  erfx_sim_set_graph(erfx_sim* sim, int on);
/* Variants of the acoustic column kernel of erf_fast_rhs_N: 0..8 one lane per
 * column (levels of loads in flight 2/3/4 x work-group width 64/128/256), 9
 * the tile form (64 columns per work-group of 16 waves, the sweep's
 * intermediates in LDS; boxes of up to 128 levels, variant 7 above that), the
 * default.  Every variant computes the same values. */
#define ERFX_COL_VARIANTS 10
/* Autotuner of the acoustic column kernel (no counterpart in the reference):
 * with `on`, the next 3 x ERFX_COL_VARIANTS launches of the erf_fast_rhs_N
 * column kernel cycle through the variants, each timed with HIP events, and
 * the fastest is kept for the rest of the run.  On a terrain grid
 * (use_terrain) the call does nothing (graph replay stays available). */
int  erfx_sim_autotune(erfx_sim* sim, int on);
/* The column variant in use (0..ERFX_COL_VARIANTS-1; -1 if still timing) and,
 * when ms is not NULL, the best time of each variant in ms
 * (ERFX_COL_VARIANTS values, 0 if not timed). */
int  erfx_sim_column_variant(erfx_sim* sim, int* variant, float* ms);
/* Fix the column variant (0..ERFX_COL_VARIANTS-1) and stop any tuning. */
int  erfx_sim_set_column_variant(erfx_sim* sim, int variant);
/* The sim's stream (a hipStream_t) for event timing. */
void* erfx_sim_stream(erfx_sim* sim);
/* Time-step control of ERF's time loop (ERF::ComputeDt / ERF::estTimeStep,
 * Source/TimeIntegration/ERF_ComputeTimestep.cpp:9-233; defaults in ERF.cpp:27-30):
 * fixed_dt <= 0 selects the CFL time step; the MRI ratio comes from
 * fixed_mri_ratio when > 0, from fixed_dt / fixed_fast_dt when both are > 0,
 * and from the CFL estimates otherwise (made even). */
typedef struct erfx_dt_control {
    double cfl;                   /* 0.8 in ERF                                      */
    double fixed_dt;              /* > 0: fixed slow time step                       */
    double fixed_fast_dt;         /* > 0 with fixed_dt: fixed acoustic time step     */
    int    fixed_mri_ratio;       /* > 0: fixed_mri_dt_ratio                         */
    double change_max;            /* 1.1 in ERF                                      */
    double init_shrink;           /* 1.0 in ERF                                      */
    double stop_time;             /* run ends there                                  */
} erfx_dt_control;

/* estTimeStep on the sim's current state: the slow time step and the MRI ratio
 * (an FP64 max-reduction on the GPU, one host read). */
int  erfx_sim_est_dt(erfx_sim* sim, const erfx_dt_control* dtc, double* dt, int* mri_ratio);
/* ERF::Evolve's loop: up to max_steps steps from t0, each with ComputeDt first
 * (limited to change_max times the previous step, init_shrink on the first,
 * cut to land on stop_time).  Synchronous; *t_end and *steps_done report. */
int  erfx_sim_run(erfx_sim* sim, double t0, int max_steps, const erfx_dt_control* dtc, double* t_end,
                  int* steps_done);

/* The volume-weighted sums of ERF::sum_integrated_quantities (volWgtSumMF,
 * ERF_WriteScalarProfiles.cpp:11-60, 378-410) over the valid cells: total mass,
 * rho*theta and RhoScalar.  Summed in double-double, so the result is the
 * exactly rounded sum of the per-cell terms, independent of reduction order. */
int  erfx_sim_integrals(erfx_sim* sim, double* mass, double* rhotheta, double* scalar);

/* Checkpoint and restart (the role of ERF::WriteCheckpointFile /
 * ReadCheckpointFile, Source/IO/ERF_Checkpoint.cpp:25-590): the sim's
 * persistent state -- cons, the three velocities, the base state, the node
 * heights of a terrain grid, the accumulated surface rain and the Rayleigh
 * profiles; every other buffer is made again from these at the start of a
 * step -- written to directory `dir` (created if absent) as a text Header and
 * one raw FP64 file, with the caller's time and step.  Reading it into a sim
 * made from the same configuration and stepping on gives the results of the
 * run that wrote it, bit for bit; a checkpoint of another grid is refused. */
int  erfx_sim_write_checkpoint(erfx_sim* sim, const char* dir, double time, int step);
int  erfx_sim_read_checkpoint(erfx_sim* sim, const char* dir, double* time, int* step);

/* Microphysics of the sim's steps: the network (erfx_kessler_mlp_advance) when
 * m is non-NULL, Kessler otherwise; counts of network and fallback cells
 * summed over the steps since it was set. */
int  erfx_sim_set_mlp(erfx_sim* sim, erfx_mlp* m);
int  erfx_sim_mlp_counts(erfx_sim* sim, long* n_mlp, long* n_fallback);
/* Harvest erfx_kessler_samples rows before every step's microphysics into a
 * device buffer of max_rows x 12 floats (NULL / 0 stops); the count keeps
 * growing past max_rows, rows beyond it are dropped. */
int  erfx_sim_set_sampling(erfx_sim* sim, float* rows, long max_rows);
int  erfx_sim_sample_count(erfx_sim* sim, long* count);

/* Per-launch timing of the step's dominant kernel with HIP events on the sim's
 * stream (the role of ERF's BL_PROFILE regions): reset, run steps, then read
 * the kernel name, summed milliseconds, launch count and algorithmic HBM bytes
 * per launch. */
int  erfx_sim_timer_reset(erfx_sim* sim);
int  erfx_sim_timer_read(erfx_sim* sim, char* name, int name_len, double* total_ms, long* launches,
                         double* bytes_per_launch);
/* Stop the per-launch timing; the readings stay available to timer_read. */
int  erfx_sim_timer_stop(erfx_sim* sim);

/* ---- batched ensembles (no counterpart in the reference) --------------------
 * ERF advances each box of a level with its own kernel launches (one
 * ParallelFor per box, MFIter loops in ERF_slow_rhs_pre.cpp:244,
 * ERF_fast_rhs_N.cpp:115-566, ...).  An ensemble advances n independent sims (boxes
 * of any sizes, or ensemble members) together: every kernel of the step runs
 * ONCE for all members of a group, its work-groups split into one range per
 * member.  The members are split into groups of about equal cell count (2 by
 * default), each group on a stream of its own so that the groups' kernels
 * overlap.
 * Members keep their own buffers, configs and scalars (seed, wind, dt-derived
 * constants); the members of a group must record the same sequence of kernels
 * (same scheme options, same MRI ratio; sizes may differ), else
 * erfx_ensemble_step returns ERFX_ERR_UNSUPPORTED.  Not with the learned microphysics or row sampling.
 * The results are those of stepping each member alone. */
typedef struct erfx_ensemble erfx_ensemble;
/* members stay owned by the caller and must outlive the ensemble */
int  erfx_ensemble_create(erfx_sim* const* members, int n, erfx_ensemble** out);
int  erfx_ensemble_destroy(erfx_ensemble* e);
/* nsteps steps of size dt from t0 for every member (synchronous at return).
 * The step's batched launch list is recorded on first use per (dt, stage time
 * steps, members' buffer parity, column variant) and reused. */
int  erfx_ensemble_step(erfx_ensemble* e, double t0, double dt, int nsteps);
/* number of member groups / streams (1..8; 2 by default) */
int  erfx_ensemble_set_streams(erfx_ensemble* e, int nstreams);
/* replay each recorded step as a hipGraph (one graph launch per group and step) */
int  erfx_ensemble_set_graph(erfx_ensemble* e, int on);
/* With `on`, the next ERFX_COL_VARIANTS steps each use one variant of the
 * acoustic column kernel for all members, timed with HIP events; the fastest
 * is kept.
 * Every variant computes the same values.  Does nothing when a member has a
 * terrain grid (the erf_fast_rhs_T column kernel has one form). */
int  erfx_ensemble_autotune(erfx_ensemble* e, int on);
int  erfx_ensemble_set_column_variant(erfx_ensemble* e, int variant);
/* the variant in use (-1 while timing) and, when ms is not NULL, the step
 * time in ms of each variant timed (ERFX_COL_VARIANTS values) */
int  erfx_ensemble_column_variant(erfx_ensemble* e, int* variant, float* ms);
/* kernel launches per step (batched, all groups) and the sum of the members' cells */
int  erfx_ensemble_info(erfx_ensemble* e, long* launches_per_step, long* cells);

#ifdef __cplusplus
}
#endif
#endif /* ERFX_H */
