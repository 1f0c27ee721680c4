// erfx_api.hip -- C ABI (include/erfx.h): configuration, hook contexts and the
// whole-step driver (ERF::Advance -> advance_dycore -> MRISplitIntegrator::advance).
#include "erfx_common.h"
#include "erfx_internal.h"
#include "erfx_mlp.h"

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <sys/stat.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace erfx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
thread_local Recorder* g_rec = nullptr;

int hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return ERFX_ERR_HIP;
}

// ERF::init_bcs (Source/Initialization/ERF_init_bcs.cpp:300-500): physical face
// types -> logical kinds per variable; plus derived constants.
void make_params(const erfx_config& c, const erfx_layout& lay, Params& P) {
    P.cfg = c;
    P.L = Lay{lay.nx, lay.ny, lay.nz, lay.ng, lay.sy, lay.sz, lay.nelem};
    P.ng_cons = lay.ng;
    P.ng_vel = lay.ng - 1;
    P.rdOcp = R_d / c.c_p;
    P.gravity = c.use_gravity ? CONST_GRAV : 0.0;
    P.dxInv[0] = 1.0 / c.dx; P.dxInv[1] = 1.0 / c.dy; P.dxInv[2] = 1.0 / c.dz;
    for (int d = 0; d < 3; ++d) P.periodic[d] = (c.bc[d] == ERFX_BC_PERIODIC);
    for (int ori = 0; ori < 6; ++ori) {
        int dir = ori % 3, t = c.bc[ori];
        for (int v = 0; v < 3; ++v) {
            P.vel_extdir[v][ori] = 0.0;
            int k = INT_DIR;
            switch (t) {
                case ERFX_BC_PERIODIC: k = INT_DIR; break;
                case ERFX_BC_SYMMETRY: k = (v == dir) ? REFLECT_ODD : REFLECT_EVEN; break;
                case ERFX_BC_OUTFLOW: k = (v == dir) ? NEUMANN_INT : FOEXTRAP; break;
                case ERFX_BC_NO_SLIP_WALL:   // <face>.velocity, normal component 0 (ERF_init_bcs.cpp:184-200)
                    k = EXT_DIR;
                    P.vel_extdir[v][ori] = (v == dir) ? 0.0 : c.wall_velocity[ori][v];
                    break;
                case ERFX_BC_SLIP_WALL: k = (v == dir) ? EXT_DIR : FOEXTRAP; break;
                case ERFX_BC_INFLOW: k = EXT_DIR; P.vel_extdir[v][ori] = c.inflow_velocity[ori][v]; break;
            }
            P.bck[v][ori] = k;
        }
        P.bck[3][ori] = (t == ERFX_BC_PERIODIC) ? INT_DIR : (t == ERFX_BC_SYMMETRY ? REFLECT_EVEN : FOEXTRAP);
        for (int n = 0; n < ERFX_NCONS_MOIST; ++n) P.cons_extdir[n][ori] = 0.0;
        if (t == ERFX_BC_INFLOW && c.inflow_density[ori] != 0.0) {
            const double r = c.inflow_density[ori];
            P.bck[3][ori] = EXT_DIR;
            P.cons_extdir[ERFX_RHO][ori] = r;
            P.cons_extdir[ERFX_RHOTHETA][ori] = r * c.inflow_theta[ori];
            P.cons_extdir[ERFX_RHOSCALAR][ori] = r * c.inflow_scalar[ori];
            if (c.moisture_type != ERFX_MOIST_NONE) {
                P.cons_extdir[ERFX_RHOQ1][ori] = r * c.inflow_qv[ori];
                P.cons_extdir[ERFX_RHOQ2][ori] = r * c.inflow_qc[ori];
            }
        }
    }
    P.consA = c.molec_diff_type == ERFX_MOLEC_CONSTANT_ALPHA;
    P.use_turb = c.les_type == ERFX_LES_SMAGORINSKY;
    P.use_diff = c.molec_diff_type != ERFX_MOLEC_NONE || c.les_type != ERFX_LES_NONE;
    P.moist = c.moisture_type != ERFX_MOIST_NONE;
    P.n_qstate = P.moist ? 3 : 0;
    P.terrain = c.use_terrain != 0;
    P.dz_min = c.dz;
}

int validate_config(const erfx_config& c) {
    if (c.nx < 4 || c.ny < 4 || c.nz < 4) { set_error("box must be at least 4 cells in each direction"); return ERFX_ERR_ARG; }
    if (!(c.dx > 0 && c.dy > 0 && c.dz > 0)) { set_error("cell sizes must be positive"); return ERFX_ERR_ARG; }
    if (c.moisture_type < ERFX_MOIST_NONE || c.moisture_type > ERFX_MOIST_KESSLER_NORAIN) {
        set_error("bad moisture_type"); return ERFX_ERR_ARG;
    }
    const int want_ncons = c.moisture_type == ERFX_MOIST_NONE ? ERFX_NCONS_DRY : ERFX_NCONS_MOIST;
    if (c.ncons != want_ncons) {
        set_error("ncons must be 5 without moisture and 8 with a Kessler model"); return ERFX_ERR_ARG;
    }
    if (c.moisture_type == ERFX_MOIST_NONE && c.buoyancy_type != 1) {
        set_error("only buoyancy_type 1 is implemented for the dry state"); return ERFX_ERR_UNSUPPORTED;
    }
    if (c.buoyancy_type < 1 || c.buoyancy_type > 4) { set_error("buoyancy_type must be 1, 2, 3 or 4"); return ERFX_ERR_ARG; }
    if (c.precision != ERFX_PREC_FP64 && c.precision != ERFX_PREC_MIXED) { set_error("bad precision"); return ERFX_ERR_ARG; }
    if (c.use_terrain && c.les_type != ERFX_LES_NONE) {
        set_error("an LES model on a terrain-following grid is not implemented"); return ERFX_ERR_UNSUPPORTED;
    }
    if (c.moisture_type == ERFX_MOIST_KESSLER_NORAIN && c.buoyancy_type != 1) {
        // asserted by ERF (ERF_make_buoyancy.cpp:171-175)
        set_error("Kessler_NoRain requires buoyancy_type 1"); return ERFX_ERR_ARG;
    }
    for (int o = 0; o < 6; ++o) {
        if (c.bc[o] < ERFX_BC_PERIODIC || c.bc[o] > ERFX_BC_INFLOW) { set_error("bad bc type"); return ERFX_ERR_ARG; }
        if (c.bc[o] == ERFX_BC_INFLOW) {
            const int d = o % 3;
            if (d == 2) { set_error("inflow on a vertical face is not supported"); return ERFX_ERR_UNSUPPORTED; }
            if (c.bc[1 - d] != ERFX_BC_PERIODIC) {
                set_error("inflow needs the other lateral direction periodic"); return ERFX_ERR_UNSUPPORTED;
            }
            if (c.les_type != ERFX_LES_NONE || c.use_terrain) {
                set_error("inflow with an LES model or terrain is not supported"); return ERFX_ERR_UNSUPPORTED;
            }
        }
        if ((c.bc[o] == ERFX_BC_PERIODIC) != (c.bc[o % 3 == o ? o + 3 : o - 3] == ERFX_BC_PERIODIC)) {
            set_error("periodic faces must come in pairs"); return ERFX_ERR_ARG;
        }
    }
    if (c.bc[2] == ERFX_BC_PERIODIC) { set_error("z cannot be periodic"); return ERFX_ERR_UNSUPPORTED; }
    if (c.bc[2] == ERFX_BC_SYMMETRY || c.bc[2] == ERFX_BC_OUTFLOW || c.bc[5] == ERFX_BC_SYMMETRY) {
        set_error("vertical faces must be slip or no-slip walls (or an outflow top)"); return ERFX_ERR_UNSUPPORTED;
    }
    int dyc[2] = {c.dycore_horiz_adv, c.dycore_vert_adv};
    for (int a : dyc)
        if (a < ERFX_ADV_CENTERED2 || a > ERFX_ADV_CENTERED6) {
            set_error("unsupported dycore advection scheme (WENO is for scalars only)"); return ERFX_ERR_UNSUPPORTED;
        }
    int scal[4] = {c.dryscal_horiz_adv, c.dryscal_vert_adv, c.moistscal_horiz_adv, c.moistscal_vert_adv};
    for (int a : scal)
        if (a < ERFX_ADV_CENTERED2 || a > ERFX_ADV_WENOZ5) { set_error("unsupported scalar advection scheme"); return ERFX_ERR_UNSUPPORTED; }
    if (c.molec_diff_type < 0 || c.molec_diff_type > 2 || c.les_type < 0 || c.les_type > 1) {
        set_error("bad diffusion choice"); return ERFX_ERR_ARG;
    }
    if (c.les_type == ERFX_LES_SMAGORINSKY && !(c.Cs > 0)) { set_error("Smagorinsky needs Cs > 0"); return ERFX_ERR_ARG; }
    if ((c.rayleigh_damp_U || c.rayleigh_damp_V || c.rayleigh_damp_W || c.rayleigh_damp_T) && !(c.rayleigh_zdamp > 0)) {
        set_error("Rayleigh damping needs rayleigh_zdamp > 0"); return ERFX_ERR_ARG;
    }
    if (c.no_substepping && c.use_terrain) {
        set_error("substepping_type None on a terrain grid is not implemented"); return ERFX_ERR_UNSUPPORTED;
    }
    if (c.no_substepping) return ERFX_OK;   // the ratio is not used (ERF_MRI.H:205-207)
    if (!(c.mri_ratio > 1 && c.mri_ratio % 2 == 0)) { set_error("mri_ratio must be even and > 1 (ERF_MRI.H:205-207)"); return ERFX_ERR_ARG; }
    if (!c.force_stage1_single_substep && c.mri_ratio % 3 != 0) {
        set_error("without the single first-stage substep the ratio must be a multiple of 6"); return ERFX_ERR_ARG;
    }
    return ERFX_OK;
}

}  // namespace erfx

using namespace erfx;

// ---- hook context ------------------------------------------------------------
struct erfx_ctx {
    Params P;
    hipStream_t stream = nullptr;
    const double* r0 = nullptr;
    const double* p0 = nullptr;
    const double* pi0 = nullptr;
    FastWork fw{};
    SlowWork sw{};
    double* ray = nullptr;    // Rayleigh profiles and per-level factors (RayleighTable layout)
    std::vector<double> ray_prof;   // host copy of the four profiles as given (4 x (nz+1))
    double* kfz = nullptr;    // Kessler sedimentation flux on z-faces (moist only)
    int* kbad = nullptr;      // device word: a Kessler temperature left the saturation fit's range
    // terrain-following grid: node heights and metrics (the Params pointers
    // refer to these) and the work fields of erf_fast_rhs_T
    double *z_nd = nullptr, *detJ = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr, *z_cc = nullptr;
    double *mrx = nullptr, *mry = nullptr;
    double* ray_pt = nullptr;   // per-point Rayleigh factors on a terrain grid (RPT_* blocks)
    TerrainWork tw{};
    // learned microphysics work space (allocated on first use): per-cell slot,
    // network rows, their cells, two device counters and their pinned host copy
    int* eslot = nullptr;
    float* ex = nullptr;
    float* ey = nullptr;
    long* erows = nullptr;
    unsigned long long* ecount = nullptr;
    unsigned long long* ecount_host = nullptr;
    std::vector<void*> allocs;
};

namespace erfx {
const Params& ctx_params(const erfx_ctx* c) { return c->P; }
hipStream_t ctx_stream(const erfx_ctx* c) { return c->stream; }
}  // namespace erfx

// ---- learned microphysics ------------------------------------------------------
struct erfx_mlp {
    Net net;
    std::vector<void*> allocs;
    MlpScaling sc;
    bool scaled = false;
};

// ---- whole-step driver ---------------------------------------------------------
struct erfx_sim {
    erfx_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double* cons[2] = {nullptr, nullptr};     // vars_old / vars_new, swapped each step
    double* vel[2][3] = {{nullptr}};
    int cur = 0;                              // index of the current (new) state
    double *r0 = nullptr, *p0 = nullptr, *pi0 = nullptr;
    erfx_state S[5];                          // S_old, S_new, S_sum, S_scratch, F_slow
    double *prim = nullptr, *pi_stage = nullptr, *fast_coeffs = nullptr, *eddy = nullptr;
    std::vector<void*> allocs;
    KTimer timer;
    double* red = nullptr;    // partial maxima of the time-step estimate (EST_BLOCKS x 2)
    double* red_host = nullptr;
    double* rain_accum = nullptr;   // Kessler surface rain, k = 0 plane (sz doubles)
    erfx_mlp* mlp = nullptr;        // learned microphysics in place of Kessler's local processes
    long mlp_rows = 0, mlp_fallback = 0;
    float* samples = nullptr;       // training rows harvested before each step's microphysics
    long samples_max = 0;
    unsigned long long* samples_count = nullptr;
    ColTuner* tuner = nullptr;      // column-kernel autotuner (erfx_sim_autotune)
    // two captured steps replayed as one graph (erfx_sim_set_graph)
    bool graph_on = false;
    hipGraphExec_t gexec = nullptr;
    double g_dt = 0.0;
    double g_sdt[6] = {};   // stage time steps of the two captured steps
    int g_ratio = 0, g_cur = -1;
};

static void tuner_free(erfx_sim* s) {
    if (!s->tuner) return;
    for (hipEvent_t& e : s->tuner->ev)
        if (e) (void)hipEventDestroy(e);
    delete s->tuner;
    s->tuner = nullptr;
    s->ctx->P.tuner = nullptr;
}

// after the last timed launch: keep the fastest variant
static int tuner_poll(erfx_sim* s) {
    ColTuner* t = s->tuner;
    if (!t || t->chosen >= 0 || t->next < COL_NV * ColTuner::SAMPLES) return ERFX_OK;
    ERFX_HIP(hipEventSynchronize(t->ev[2 * t->next - 1]));
    int best = 0;
    for (int v = 0; v < COL_NV; ++v) {
        float m = 1e30f;
        for (int r = 0; r < ColTuner::SAMPLES; ++r) {
            const int tag = r * COL_NV + v;
            float x = 0.f;
            ERFX_HIP(hipEventElapsedTime(&x, t->ev[2 * tag], t->ev[2 * tag + 1]));
            m = std::min(m, x);
        }
        t->ms[v] = m;
        if (m < t->ms[best]) best = v;
    }
    t->chosen = best;
    s->ctx->P.col_variant = best;
    s->ctx->P.tuner = nullptr;
    return ERFX_OK;
}

static void graph_free(erfx_sim* s) {
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
    s->g_cur = -1;
}

static void timer_free(KTimer& t) {
    for (int i = 0; i < t.cap; ++i) (void)hipEventDestroy(t.ev[i]);
    delete[] t.ev;
    t = KTimer();
}

extern "C" {

const char* erfx_last_error(void) { return g_last_error.c_str(); }

void erfx_config_defaults(erfx_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->nx = c->ny = c->nz = 16;
    c->dx = c->dy = c->dz = 100.0;
    for (int o = 0; o < 6; ++o) c->bc[o] = ERFX_BC_PERIODIC;
    c->bc[2] = c->bc[5] = ERFX_BC_SLIP_WALL;
    c->ncons = ERFX_NCONS_DRY;
    c->dycore_horiz_adv = c->dycore_vert_adv = ERFX_ADV_UPWIND3;
    c->dryscal_horiz_adv = c->dryscal_vert_adv = ERFX_ADV_UPWIND3;
    c->dycore_horiz_upw_frac = c->dycore_vert_upw_frac = 1.0;
    c->dryscal_horiz_upw_frac = c->dryscal_vert_upw_frac = 1.0;
    c->molec_diff_type = ERFX_MOLEC_NONE;
    c->rho0_trans = 1.0;
    c->les_type = ERFX_LES_NONE;
    c->Pr_t = 1.0 / 3.0;
    c->Sc_t = 1.0;
    c->use_gravity = 1;
    c->c_p = Cp_d;
    c->coriolis_factor = 2.0 * 2.0 * PI / 86400.0;
    c->sinphi = 1.0;
    c->mri_ratio = 6;
    c->force_stage1_single_substep = 1;
    c->implicit_substepping = 1;
    c->rayleigh_dampcoef = 0.2;
    c->rayleigh_zdamp = 500.0;
    c->moisture_type = ERFX_MOIST_NONE;
    c->buoyancy_type = 1;
    c->moistscal_horiz_adv = c->moistscal_vert_adv = ERFX_ADV_WENO3;
    c->moistscal_horiz_upw_frac = c->moistscal_vert_upw_frac = 1.0;
}

int erfx_ghost_cells(const erfx_config* c) {
    if (!c) return -1;
    auto wide = [](int t) {
        return t == ERFX_ADV_UPWIND5 || t == ERFX_ADV_CENTERED6 || t == ERFX_ADV_WENO5 || t == ERFX_ADV_WENOZ5;
    };
    int g = (wide(c->dycore_horiz_adv) || wide(c->dycore_vert_adv) || wide(c->dryscal_horiz_adv) ||
             wide(c->dryscal_vert_adv) || wide(c->moistscal_horiz_adv)) ? 3 : 2;
    // ERF tests the moist vertical type only against WENO5/5Z
    if (c->moistscal_vert_adv == ERFX_ADV_WENO5 || c->moistscal_vert_adv == ERFX_ADV_WENOZ5) g = 3;
    return g + 1;
}

int erfx_layout_for(const erfx_config* c, erfx_layout* out) {
    if (!c || !out) { set_error("null argument"); return ERFX_ERR_ARG; }
    int ng = erfx_ghost_cells(c);
    out->nx = c->nx; out->ny = c->ny; out->nz = c->nz; out->ng = ng;
    long row = c->nx + 1 + 2 * ng;
    out->sy = (row + 15) / 16 * 16;
    out->sz = out->sy * (c->ny + 1 + 2 * ng);
    out->nelem = out->sz * (c->nz + 1 + 2 * ng);
    return ERFX_OK;
}

static int launch_status() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return ERFX_OK;
}

// Device allocation of n doubles, zeroed.  With `shift`, 16 more doubles are
// allocated and the pointer handed out is the allocation plus `shift`: for a
// field of the layout, shift = 16 - ng (mod 16) puts element (0, j, k) of every
// row on a 128-byte line (rows and planes are multiples of 16 doubles long, and
// the allocation itself starts on a line), so a wave's 64 consecutive cells of
// a row touch 4 cache lines instead of 5.
static int dalloc(std::vector<void*>& list, double** p, size_t n, size_t shift = 0) {
    void* q = nullptr;
    const size_t total = n + (shift ? 16 : 0);
    hipError_t e = hipMalloc(&q, total * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    list.push_back(q);
    // zero once so never-written ghost layers hold defined values
    e = hipMemset(q, 0, total * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMemset");
    *p = static_cast<double*>(q) + shift;
    return ERFX_OK;
}
// the shift of a layout field (see dalloc)
static size_t row_shift(const Lay& L) { return (size_t)((16 - L.ng % 16) % 16); }

int erfx_ctx_create(const erfx_config* cfg, void* stream, erfx_ctx** out) {
    if (!cfg || !out) { set_error("null argument"); return ERFX_ERR_ARG; }
    int rc = validate_config(*cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { set_error("no HIP device"); return ERFX_ERR_NODEVICE; }
    auto* c = new erfx_ctx();
    erfx_layout lay;
    erfx_layout_for(cfg, &lay);
    make_params(*cfg, lay, c->P);
    c->stream = static_cast<hipStream_t>(stream);
    const size_t ne = (size_t)lay.nelem;
    const size_t sh = row_shift(c->P.L);
    double** bufs[] = {&c->fw.extrap, &c->fw.drw, &c->fw.soln, &c->sw.pp,
                       &c->sw.tau[0], &c->sw.tau[1], &c->sw.tau[2], &c->sw.tau[3], &c->sw.tau[4], &c->sw.tau[5]};
    for (double** b : bufs)
        if ((rc = dalloc(c->allocs, b, ne, sh))) { erfx_ctx_destroy(c); return rc; }
    if ((rc = dalloc(c->allocs, &c->fw.trhs, 2 * ne, sh))) { erfx_ctx_destroy(c); return rc; }
    c->sw.xsrc = c->sw.ysrc = c->sw.zsrc = nullptr;
    if (c->P.moist) {
        const size_t nz = (size_t)lay.nz;
        double* kb = nullptr;
        if ((rc = dalloc(c->allocs, &c->kfz, ne, sh)) ||
            (rc = dalloc(c->allocs, &c->sw.pavg, (size_t)PLANE_NOUT * nz)) ||
            (rc = dalloc(c->allocs, &c->sw.pavg_part, nz * PLANE_SPLIT_MAX * 2 * PLANE_ROWS)) ||
            (rc = dalloc(c->allocs, &kb, 1))) {
            erfx_ctx_destroy(c);
            return rc;
        }
        c->kbad = reinterpret_cast<int*>(kb);
    }
    if (c->P.terrain) {
        double** tb[] = {&c->z_nd, &c->detJ, &c->ax, &c->ay, &c->az, &c->z_cc, &c->mrx, &c->mry, &c->sw.omega,
                         &c->tw.du, &c->tw.dv, &c->tw.nu, &c->tw.nv, &c->tw.om};
        for (double** b : tb)
            if ((rc = dalloc(c->allocs, b, ne, sh))) { erfx_ctx_destroy(c); return rc; }
        if ((rc = dalloc(c->allocs, &c->tw.wtop, (size_t)lay.sz, sh))) { erfx_ctx_destroy(c); return rc; }
        if (c->P.use_diff) {
            double** db[] = {&c->sw.tau_t[0], &c->sw.tau_t[1], &c->sw.tau_t[2], &c->sw.er, &c->sw.dfx, &c->sw.dfy};
            for (double** b : db)
                if ((rc = dalloc(c->allocs, b, ne, sh))) { erfx_ctx_destroy(c); return rc; }
            if ((rc = dalloc(c->allocs, &c->sw.dstate, (size_t)cfg->ncons * ne, sh))) { erfx_ctx_destroy(c); return rc; }
        }
        c->P.z_nd = c->z_nd; c->P.detJ = c->detJ; c->P.ax = c->ax; c->P.ay = c->ay; c->P.az = c->az;
        c->P.z_cc = c->z_cc;
        c->P.mrx = c->mrx;
        c->P.mry = c->mry;
        // flat node heights until erfx_set_terrain gives the real ones
        std::vector<double> zf(ne);
        for (size_t q = 0; q < ne; ++q) zf[q] = cfg->prob_lo[2] + (double)((long)(q / lay.sz) - lay.ng) * cfg->dz;
        hipError_t e = hipMemcpy(c->z_nd, zf.data(), ne * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) { erfx_ctx_destroy(c); return hip_fail(e, "hipMemcpy"); }
        terrain_metrics(c->P, c->z_nd, c->detJ, c->ax, c->ay, c->az, c->z_cc, c->mrx, c->mry, c->stream);
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { erfx_ctx_destroy(c); return hip_fail(e, "terrain metrics"); }
    }
    *out = c;
    return ERFX_OK;
}

// Per-point Rayleigh factors of a terrain grid: the reference evaluates the
// damping at each cell's z_cc and each z-node's height, with ztop the height
// of node (0, 0, nz) (ERF_make_sources.cpp:203-225, ERF_make_mom_sources.cpp:247-296,
// ERF_init_rayleigh.cpp:17); computed here on the host with the C library's sin.
static int build_ray_points(erfx_ctx* c) {
    if (!c->P.terrain || !c->ray) return ERFX_OK;
    const Lay L = c->P.L;
    const size_t ne = (size_t)L.ne;
    std::vector<double> zc(ne), zn(ne), h((size_t)RPT_NFIELDS * ne, 0.0);
    ERFX_HIP(hipMemcpyAsync(zc.data(), c->z_cc, ne * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ERFX_HIP(hipMemcpyAsync(zn.data(), c->z_nd, ne * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ERFX_HIP(hipStreamSynchronize(c->stream));
    const erfx_config& cf = c->P.cfg;
    const double ztop = zn[(size_t)L.ix(0, 0, L.nz)];
    const double zdamp = cf.rayleigh_zdamp, dampcoef = cf.rayleigh_dampcoef;
    for (size_t q = 0; q < ne; ++q) {
        double zfrac = 1 - (ztop - zc[q]) / zdamp;
        if (zfrac > 0) {
            double sinefac = std::sin(PIoTwo * zfrac);
            h[RPT_FAC_CC * ne + q] = dampcoef * sinefac * sinefac;
            h[RPT_ACT_CC * ne + q] = 1.0;
        }
        double zfrac_n = 1 - (ztop - zn[q]) / zdamp;
        if (zfrac_n > 0) {
            double sinefac = std::sin(PIoTwo * zfrac_n);
            h[RPT_FAC_ND * ne + q] = dampcoef * sinefac * sinefac;
            h[RPT_ACT_ND * ne + q] = 1.0;
        }
    }
    if (!c->ray_pt) {
        int rc = dalloc(c->allocs, &c->ray_pt, h.size());
        if (rc) return rc;
    }
    ERFX_HIP(hipMemcpy(c->ray_pt, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    c->P.ray_pt = c->ray_pt;
    return ERFX_OK;
}

int erfx_set_terrain(erfx_ctx* c, const double* z_nd) {
    if (!c || !z_nd) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (!c->P.terrain) { set_error("the context was made without use_terrain"); return ERFX_ERR_STATE; }
    if (z_nd != c->z_nd)
        ERFX_HIP(hipMemcpyAsync(c->z_nd, z_nd, (size_t)c->P.L.ne * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    terrain_metrics(c->P, c->z_nd, c->detJ, c->ax, c->ay, c->az, c->z_cc, c->mrx, c->mry, c->stream);
    int rc = launch_status();
    if (rc) return rc;
    // dz_min of the time-step estimate (ERF.cpp:898-906): dz times the smallest
    // detJ of the valid cells
    const Lay L = c->P.L;
    std::vector<double> J((size_t)L.ne);
    ERFX_HIP(hipMemcpyAsync(J.data(), c->detJ, (size_t)L.ne * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ERFX_HIP(hipStreamSynchronize(c->stream));
    double jmin = 1.e300;
    for (int k = 0; k < L.nz; ++k)
        for (int j = 0; j < L.ny; ++j)
            for (int i = 0; i < L.nx; ++i) jmin = std::min(jmin, J[(size_t)L.ix(i, j, k)]);
    c->P.dz_min = c->P.cfg.dz * jmin;
    return build_ray_points(c);
}

int erfx_ctx_destroy(erfx_ctx* c) {
    if (!c) return ERFX_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    views_release(c);
    if (c->ecount_host) (void)hipHostFree(c->ecount_host);
    for (void* p : c->allocs) (void)hipFree(p);
    delete c;
    return ERFX_OK;
}

int erfx_ctx_layout(const erfx_ctx* c, erfx_layout* out) {
    if (!c || !out) { set_error("null argument"); return ERFX_ERR_ARG; }
    out->nx = c->P.L.nx; out->ny = c->P.L.ny; out->nz = c->P.L.nz; out->ng = c->P.L.ng;
    out->sy = c->P.L.sy; out->sz = c->P.L.sz; out->nelem = c->P.L.ne;
    return ERFX_OK;
}

int erfx_set_base_state(erfx_ctx* c, const double* r0, const double* p0, const double* pi0) {
    if (!c || !r0 || !p0 || !pi0) { set_error("null argument"); return ERFX_ERR_ARG; }
    c->r0 = r0; c->p0 = p0; c->pi0 = pi0;
    return ERFX_OK;
}

// Per-level Rayleigh table: ubar, vbar, wbar, thetabar, then the factor
// dampcoef * sin^2(pi/2 zfrac) and a 0/1 activity flag at cell centres and at
// z-faces, each nz+1 long.  The factors depend on k only, so they are computed
// here on the host with the C library's sin, exactly as the reference
// evaluates them per cell (ERF_make_sources.cpp:213-222, ERF_make_mom_sources.cpp:255-296).
int erfx_set_rayleigh(erfx_ctx* c, const double* ub, const double* vb, const double* wb, const double* tb) {
    if (c && ub && vb && wb && tb) {
        const size_t n1 = (size_t)c->P.L.nz + 1;
        c->ray_prof.assign(4 * n1, 0.0);
        for (size_t k = 0; k < n1; ++k) {
            c->ray_prof[k] = ub[k]; c->ray_prof[n1 + k] = vb[k];
            c->ray_prof[2 * n1 + k] = wb[k]; c->ray_prof[3 * n1 + k] = tb[k];
        }
    }
    if (!c || !ub || !vb || !wb || !tb) { set_error("null argument"); return ERFX_ERR_ARG; }
    const erfx_config& cf = c->P.cfg;
    const int n1 = cf.nz + 1;
    std::vector<double> h((size_t)RAY_NFIELDS * n1, 0.0);
    for (int k = 0; k < n1; ++k) {
        h[RAY_UBAR * n1 + k] = ub[k];
        h[RAY_VBAR * n1 + k] = vb[k];
        h[RAY_WBAR * n1 + k] = wb[k];
        h[RAY_THETABAR * n1 + k] = tb[k];
    }
    const double zlo = cf.prob_lo[2], dz = cf.dz;
    // ERF's rayleigh_ztop = geom.ProbHi(2) (ERF_init_rayleigh.cpp:17)
    const double ztop = cf.prob_hi[2] > cf.prob_lo[2] ? cf.prob_hi[2] : zlo + cf.nz * dz;
    const double zdamp = cf.rayleigh_zdamp, dampcoef = cf.rayleigh_dampcoef;
    for (int k = 0; k < n1; ++k) {
        double zcc = zlo + (k + 0.5) * dz;
        double zfrac = 1 - (ztop - zcc) / zdamp;
        if (zfrac > 0) {
            double sinefac = std::sin(PIoTwo * zfrac);
            h[RAY_FAC_CC * n1 + k] = dampcoef * sinefac * sinefac;
            h[RAY_ACT_CC * n1 + k] = 1.0;
        }
        double zstag = zlo + k * dz;
        double zfrac_n = 1 - (ztop - zstag) / zdamp;
        if (zfrac_n > 0) {
            double sinefac = std::sin(PIoTwo * zfrac_n);
            h[RAY_FAC_ND * n1 + k] = dampcoef * sinefac * sinefac;
            h[RAY_ACT_ND * n1 + k] = 1.0;
        }
    }
    if (!c->ray) {
        int rc = dalloc(c->allocs, &c->ray, h.size());
        if (rc) return rc;
    }
    ERFX_HIP(hipMemcpy(c->ray, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    c->P.ray = c->ray;
    return build_ray_points(c);
}

static bool rayleigh_on(const erfx_config& c) {
    return c.rayleigh_damp_U || c.rayleigh_damp_V || c.rayleigh_damp_W || c.rayleigh_damp_T;
}

static bool bad_state(const erfx_state* s) { return !s || !s->cons || !s->xmom || !s->ymom || !s->zmom; }

// after a stream sync: report (and clear) a temperature out of the saturation fit's range
static int kessler_status(erfx_ctx* c) {
    if (!c->kbad) return ERFX_OK;
    int h = 0;
    ERFX_HIP(hipMemcpy(&h, c->kbad, sizeof(int), hipMemcpyDeviceToHost));
    if (h) {
        ERFX_HIP(hipMemset(c->kbad, 0, sizeof(int)));
        set_error("Kessler: temperature below -85 C, outside the saturation fit (ERF_Microphysics_Utils.H:62)");
        return ERFX_ERR_STATE;
    }
    return ERFX_OK;
}

int erfx_cons_to_prim(erfx_ctx* c, const double* cons, double* prim, double* pi_stage, int ng) {
    if (!c || !cons || !prim || !pi_stage) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (ng < 1 || ng > c->P.ng_cons) { set_error("ng out of range"); return ERFX_ERR_ARG; }
    cons_to_prim(c->P, cons, prim, pi_stage, ng, c->stream);
    return launch_status();
}

int erfx_apply_bcs(erfx_ctx* c, erfx_state* S, double* xv, double* yv, double* zv, int ng_cons, int ng_vel,
                   int fast_only, int synced) {
    if (!c || bad_state(S) || !xv || !yv || !zv) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (ng_cons < 1 || ng_cons > c->P.ng_cons || ng_vel < 0 || ng_vel > c->P.ng_vel) {
        set_error("ghost widths exceed the layout"); return ERFX_ERR_ARG;
    }
    apply_bcs(c->P, *S, xv, yv, zv, ng_cons, ng_vel, fast_only != 0, synced != 0, c->stream);
    return launch_status();
}

int erfx_make_fast_coeffs(erfx_ctx* c, double* fc, const erfx_state* stage, const double* prim,
                          const double* pi_stage, double dtau, double beta_s) {
    if (!c || !fc || bad_state(stage) || !prim || !pi_stage) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (!c->r0) { set_error("base state not set"); return ERFX_ERR_STATE; }
    make_fast_coeffs(c->P, fc, *stage, prim, pi_stage, c->r0, c->pi0, dtau, beta_s, c->stream);
    return launch_status();
}

int erfx_fast_rhs_N(erfx_ctx* c, int step, int nrk, const erfx_state* slow_rhs, const erfx_state* prev,
                    const erfx_state* stage, const double* prim, const double* pi_stage, const double* fc,
                    erfx_state* data, erfx_state* scratch, double dtau, double beta_s, double facinv) {
    if (!c || bad_state(slow_rhs) || bad_state(prev) || bad_state(stage) || bad_state(data) || bad_state(scratch) ||
        !prim || !pi_stage || !fc) {
        set_error("null argument"); return ERFX_ERR_ARG;
    }
    if (!(nrk > 0 || step == 0)) { set_error("the first RK stage takes a single substep (ERF_fast_rhs_N.cpp:106)"); return ERFX_ERR_ARG; }
    if (c->P.terrain) { set_error("a terrain-following grid takes erf_fast_rhs_T"); return ERFX_ERR_STATE; }
    fast_rhs_N(c->P, c->fw, step, nrk, *slow_rhs, *prev, *stage, prim, pi_stage, fc, *data, *scratch, dtau, beta_s,
               facinv, c->stream);
    return launch_status();
}

int erfx_fast_rhs_T(erfx_ctx* c, int step, int nrk, const erfx_state* slow_rhs, const erfx_state* prev,
                    const erfx_state* stage, const double* prim, const double* pi_stage, const double* fc,
                    erfx_state* data, erfx_state* scratch, double dtau, double beta_s, double facinv) {
    if (!c || bad_state(slow_rhs) || bad_state(prev) || bad_state(stage) || bad_state(data) || bad_state(scratch) ||
        !prim || !pi_stage || !fc) {
        set_error("null argument"); return ERFX_ERR_ARG;
    }
    if (!c->P.terrain) { set_error("erf_fast_rhs_T needs a terrain-following grid (use_terrain)"); return ERFX_ERR_STATE; }
    if (!(nrk > 0 || step == 0)) { set_error("the first RK stage takes a single substep (ERF_fast_rhs_T.cpp:129)"); return ERFX_ERR_ARG; }
    fast_rhs_T(c->P, c->fw, c->tw, step, nrk, *slow_rhs, *prev, *stage, prim, pi_stage, fc, *data, *scratch, dtau,
               beta_s, facinv, c->stream);
    return launch_status();
}

int erfx_turbulent_viscosity(erfx_ctx* c, const double* cons, const double* xv, const double* yv,
                             const double* zv, double* eddy, double* hfx3) {
    (void)hfx3;
    if (!c || !cons || !xv || !yv || !zv || !eddy) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (!c->P.use_turb) { set_error("no LES model configured"); return ERFX_ERR_ARG; }
    turbulent_viscosity(c->P, c->sw, cons, xv, yv, zv, eddy, c->stream);
    return launch_status();
}

int erfx_slow_rhs_pre(erfx_ctx* c, int nrk, double slow_dt, erfx_state* rhs, const erfx_state* data,
                      const double* prim, erfx_state* scratch, const double* xv, const double* yv,
                      const double* zv, const double* eddy, double* hfx3) {
    (void)hfx3;
    if (!c || bad_state(rhs) || bad_state(data) || bad_state(scratch) || !prim || !xv || !yv || !zv) {
        set_error("null argument"); return ERFX_ERR_ARG;
    }
    if (c->P.use_turb && !eddy) { set_error("LES model needs the eddy viscosity"); return ERFX_ERR_ARG; }
    if (!c->r0) { set_error("base state not set"); return ERFX_ERR_STATE; }
    if (rayleigh_on(c->P.cfg) && !c->ray) { set_error("Rayleigh damping profiles not set (erfx_set_rayleigh)"); return ERFX_ERR_STATE; }
    slow_rhs_pre(c->P, c->sw, nrk, slow_dt, *rhs, *data, prim, *scratch, xv, yv, zv, eddy, c->r0, c->p0, c->stream);
    return launch_status();
}

int erfx_kessler_advance(erfx_ctx* c, double* cons, double dt, double* rain_accum) {
    if (!c || !cons) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (!c->P.moist) { set_error("no moisture model configured"); return ERFX_ERR_ARG; }
    if (!(dt > 0)) { set_error("dt must be positive"); return ERFX_ERR_ARG; }
    kessler_advance(c->P, cons, c->kfz, rain_accum, dt, c->stream, c->kbad);
    int rc = launch_status();
    if (rc) return rc;
    // the range check of the saturation fit needs the result (ERF aborts there)
    ERFX_HIP(hipStreamSynchronize(c->stream));
    return kessler_status(c);
}

static int balloc(std::vector<void*>& list, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    list.push_back(*p);
    e = hipMemset(*p, 0, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMemset");
    return ERFX_OK;
}

int erfx_mlp_create(int capacity, erfx_mlp** out) {
    using namespace erfx::mlp;
    if (!out || capacity < 1) { set_error("null argument or capacity < 1"); return ERFX_ERR_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { set_error("no HIP device"); return ERFX_ERR_NODEVICE; }
    auto* m = new erfx_mlp();
    Net& N = m->net;
    N.cap = capacity;
    const size_t cap = (size_t)capacity;
    const size_t nblk = (cap + 63) / 64 + 1;     // per-block gradient partials + one block to stage the sum
    struct { void** p; size_t bytes; } req[] = {
        {(void**)&N.P.w, NPARAM * 4}, {(void**)&N.P.m, NPARAM * 4}, {(void**)&N.P.v, NPARAM * 4},
        {(void**)&N.P.w1b, NI * NH * 2}, {(void**)&N.P.w2b, NH * NH * 2}, {(void**)&N.P.w2tb, NH * NH * 2},
        {(void**)&N.P.w3b, NH * NO * 2},
        {(void**)&N.T.z1, cap * NH * 4}, {(void**)&N.T.h1, cap * NH * 2}, {(void**)&N.T.z2, cap * NH * 4},
        {(void**)&N.T.h2, cap * NH * 2}, {(void**)&N.T.dz2, cap * NH * 2}, {(void**)&N.T.dz1, cap * NH * 4},
        {(void**)&N.T.dy, cap * NO * 4}, {(void**)&N.T.loss, cap * 4},
        {(void**)&N.ybuf, cap * NO * 4}, {(void**)&N.partial, nblk * NPARAM * 4}, {(void**)&N.loss_dev, 4},
    };
    for (auto& r : req) {
        int rc = balloc(m->allocs, r.p, r.bytes);
        if (rc) { erfx_mlp_destroy(m); return rc; }
    }
    N.W = Weights{N.P.w1b, N.P.w + OFF_B1, N.P.w2b, N.P.w2tb, N.P.w + OFF_B2, N.P.w3b, N.P.w + OFF_B3};
    for (int k = 0; k < NI; ++k) { m->sc.x_mean[k] = 0.0; m->sc.x_scale[k] = 1.0; m->sc.x_lo[k] = -1e300; m->sc.x_hi[k] = 1e300; }
    for (int o = 0; o < NO; ++o) m->sc.y_scale[o] = 1.0;
    *out = m;
    return ERFX_OK;
}

int erfx_mlp_destroy(erfx_mlp* m) {
    if (!m) return ERFX_OK;
    (void)hipDeviceSynchronize();
    for (void* p : m->allocs) (void)hipFree(p);
    delete m;
    return ERFX_OK;
}

int erfx_mlp_set_params(erfx_mlp* m, const float* h) {
    if (!m || !h) { set_error("null argument"); return ERFX_ERR_ARG; }
    Net& N = m->net;
    ERFX_HIP(hipMemcpy(N.P.w, h, mlp::NPARAM * sizeof(float), hipMemcpyHostToDevice));
    ERFX_HIP(hipMemset(N.P.m, 0, mlp::NPARAM * sizeof(float)));
    ERFX_HIP(hipMemset(N.P.v, 0, mlp::NPARAM * sizeof(float)));
    N.t = 0;
    mlp_refresh(N, nullptr);
    ERFX_HIP(hipDeviceSynchronize());
    return launch_status();
}

int erfx_mlp_get_params(erfx_mlp* m, float* h) {
    if (!m || !h) { set_error("null argument"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipDeviceSynchronize());
    ERFX_HIP(hipMemcpy(h, m->net.P.w, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return ERFX_OK;
}

int erfx_mlp_set_scaling(erfx_mlp* m, const double* xm, const double* xs, const double* xl, const double* xh,
                         const double* ys) {
    if (!m || !xm || !xs || !xl || !xh || !ys) { set_error("null argument"); return ERFX_ERR_ARG; }
    for (int k = 0; k < mlp::NI; ++k) {
        if (!(xs[k] > 0)) { set_error("x_scale must be positive"); return ERFX_ERR_ARG; }
        m->sc.x_mean[k] = xm[k]; m->sc.x_scale[k] = xs[k]; m->sc.x_lo[k] = xl[k]; m->sc.x_hi[k] = xh[k];
    }
    for (int o = 0; o < mlp::NO; ++o) {
        if (!(ys[o] > 0)) { set_error("y_scale must be positive"); return ERFX_ERR_ARG; }
        m->sc.y_scale[o] = ys[o];
    }
    m->scaled = true;
    return ERFX_OK;
}

int erfx_mlp_forward(erfx_mlp* m, const float* x, float* y, int n) {
    if (!m || (n > 0 && (!x || !y)) || n < 0) { set_error("bad argument"); return ERFX_ERR_ARG; }
    mlp_forward(m->net, x, y, n, nullptr);
    return launch_status();
}

static int mlp_pass(erfx_mlp* m, const float* x, const float* t, int n, bool update, double lr, float* grad_host,
                    double* loss) {
    if (!m || !x || !t || n < 1) { set_error("bad argument"); return ERFX_ERR_ARG; }
    Net& N = m->net;
    if (n > N.cap) { set_error("more samples than the capacity given at creation"); return ERFX_ERR_ARG; }
    if (update) N.t += 1;
    // the summed gradient is staged in the partial buffer's last block, past
    // every block a pass of n <= cap samples uses
    float* gstage = grad_host ? N.partial + (size_t)((N.cap + 63) / 64) * mlp::NPARAM : nullptr;
    mlp_train_pass(N, x, t, n, update, (float)lr, gstage, nullptr);
    int rc = launch_status();
    if (rc) return rc;
    if (grad_host || loss) {
        ERFX_HIP(hipDeviceSynchronize());
        if (grad_host) ERFX_HIP(hipMemcpy(grad_host, gstage, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost));
        if (loss) {
            float ls = 0.0f;
            ERFX_HIP(hipMemcpy(&ls, N.loss_dev, sizeof(float), hipMemcpyDeviceToHost));
            *loss = (double)ls / ((double)n * mlp::NO);
        }
    }
    return ERFX_OK;
}

int erfx_mlp_gradient(erfx_mlp* m, const float* x, const float* t, int n, float* grad_host, double* loss) {
    if (!grad_host) { set_error("null argument"); return ERFX_ERR_ARG; }
    return mlp_pass(m, x, t, n, false, 0.0, grad_host, loss);
}

int erfx_mlp_train_step(erfx_mlp* m, const float* x, const float* t, int n, double lr, double* loss) {
    if (!(lr > 0)) { set_error("lr must be positive"); return ERFX_ERR_ARG; }
    return mlp_pass(m, x, t, n, true, lr, nullptr, loss);
}

static int emulator_space(erfx_ctx* c) {
    if (c->eslot) return ERFX_OK;
    const size_t ne = (size_t)c->P.L.ne;
    const size_t ncell = (size_t)c->P.L.nx * c->P.L.ny * c->P.L.nz;
    int rc;
    if ((rc = balloc(c->allocs, (void**)&c->eslot, ne * sizeof(int))) ||
        (rc = balloc(c->allocs, (void**)&c->ex, ncell * mlp::NI * sizeof(float))) ||
        (rc = balloc(c->allocs, (void**)&c->ey, ncell * mlp::NO * sizeof(float))) ||
        (rc = balloc(c->allocs, (void**)&c->erows, ncell * sizeof(long))) ||
        (rc = balloc(c->allocs, (void**)&c->ecount, 2 * sizeof(unsigned long long))))
        return rc;
    if (hipHostMalloc((void**)&c->ecount_host, 2 * sizeof(unsigned long long)) != hipSuccess) {
        set_error("hipHostMalloc failed"); return ERFX_ERR_HIP;
    }
    return ERFX_OK;
}

int erfx_kessler_samples(erfx_ctx* c, const double* cons, const double* zvel, double dt, float* rows, long max_rows,
                         long* count) {
    if (!c || !cons || !zvel || !rows || !count || max_rows < 0) { set_error("bad argument"); return ERFX_ERR_ARG; }
    if (c->P.cfg.moisture_type != ERFX_MOIST_KESSLER) { set_error("samples need the Kessler model with rain"); return ERFX_ERR_ARG; }
    int rc = emulator_space(c);
    if (rc) return rc;
    ERFX_HIP(hipMemsetAsync(c->ecount, 0, sizeof(unsigned long long), c->stream));
    kessler_samples(c->P, cons, zvel, dt, rows, max_rows, c->ecount, c->kbad, c->stream);
    if ((rc = launch_status())) return rc;
    ERFX_HIP(hipMemcpyAsync(c->ecount_host, c->ecount, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    ERFX_HIP(hipStreamSynchronize(c->stream));
    *count = (long)c->ecount_host[0];
    return kessler_status(c);
}

int erfx_kessler_mlp_advance(erfx_ctx* c, erfx_mlp* m, double* cons, const double* zvel, double dt,
                             double* rain_accum, long* n_mlp, long* n_fallback) {
    if (!c || !m || !cons || !zvel) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (c->P.cfg.moisture_type != ERFX_MOIST_KESSLER) { set_error("the learned step replaces Kessler with rain"); return ERFX_ERR_ARG; }
    if (!m->scaled) { set_error("set the network's scaling first (erfx_mlp_set_scaling)"); return ERFX_ERR_STATE; }
    if (!(dt > 0)) { set_error("dt must be positive"); return ERFX_ERR_ARG; }
    int rc = emulator_space(c);
    if (rc) return rc;
    kessler_mlp_advance(c->P, m->net, m->sc, cons, zvel, c->kfz, rain_accum, dt, c->eslot, c->ex, c->ey, c->erows,
                        c->ecount, c->ecount_host, c->kbad, c->stream);
    if ((rc = launch_status())) return rc;
    ERFX_HIP(hipStreamSynchronize(c->stream));
    if (n_mlp) *n_mlp = (long)c->ecount_host[0];
    if (n_fallback) *n_fallback = (long)c->ecount_host[1];
    return kessler_status(c);
}

int erfx_slow_rhs_post(erfx_ctx* c, int nrk, double slow_dt, erfx_state* rhs, const erfx_state* old,
                       erfx_state* snew, erfx_state* data, const double* prim, const erfx_state* scratch,
                       const double* xv, const double* yv, const double* eddy) {
    (void)xv; (void)yv;
    if (!c || bad_state(rhs) || bad_state(old) || bad_state(snew) || bad_state(data) || bad_state(scratch) || !prim) {
        set_error("null argument"); return ERFX_ERR_ARG;
    }
    if (c->P.use_turb && !eddy) { set_error("LES model needs the eddy viscosity"); return ERFX_ERR_ARG; }
    slow_rhs_post(c->P, c->sw, nrk, slow_dt, *rhs, *old, *snew, *data, prim, *scratch, eddy, c->stream);
    return launch_status();
}

// ---- sim --------------------------------------------------------------------
int erfx_sim_create(const erfx_config* cfg, erfx_sim** out) {
    if (!cfg || !out) { set_error("null argument"); return ERFX_ERR_ARG; }
    auto* s = new erfx_sim();
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete s; return hip_fail(e, "hipStreamCreate"); }
    s->own_stream = true;
    int rc = erfx_ctx_create(cfg, s->stream, &s->ctx);
    if (rc) { erfx_sim_destroy(s); return rc; }
    const Params& P = s->ctx->P;
    const size_t ne = (size_t)P.L.ne;
    const int nc = cfg->ncons;
    const size_t sh = row_shift(P.L);
    auto A = [&](double** p, size_t n) { return dalloc(s->allocs, p, n, sh); };
    for (int b = 0; b < 2 && !rc; ++b) {
        rc = A(&s->cons[b], nc * ne);
        for (int d = 0; d < 3 && !rc; ++d) rc = A(&s->vel[b][d], ne);
    }
    if (!rc) rc = A(&s->r0, ne);
    if (!rc) rc = A(&s->p0, ne);
    if (!rc) rc = A(&s->pi0, ne);
    for (int st = 0; st < 5 && !rc; ++st) {
        if (st != 0 && st != 1) rc = A(&s->S[st].cons, nc * ne);
        if (!rc) rc = A(&s->S[st].xmom, ne);
        if (!rc) rc = A(&s->S[st].ymom, ne);
        if (!rc) rc = A(&s->S[st].zmom, ne);
    }
    if (!rc) rc = A(&s->prim, (nc - 1) * ne);
    if (!rc) rc = A(&s->pi_stage, ne);
    if (!rc) rc = A(&s->fast_coeffs, 5 * ne);
    if (!rc) rc = dalloc(s->allocs, &s->red, 6 * EST_BLOCKS + 6);
    if (!rc && hipHostMalloc((void**)&s->red_host, 6 * sizeof(double)) != hipSuccess) {
        set_error("hipHostMalloc failed"); rc = ERFX_ERR_HIP;
    }
    if (!rc && P.use_turb) rc = A(&s->eddy, 13 * ne);
    if (!rc && P.moist) rc = A(&s->rain_accum, (size_t)P.L.sz);
    if (rc) { erfx_sim_destroy(s); return rc; }
    // S_old and S_new cons alias vars_old / vars_new (ERF_Advance.cpp:127-160)
    s->S[0].cons = s->cons[1 - s->cur];
    s->S[1].cons = s->cons[s->cur];
    erfx_set_base_state(s->ctx, s->r0, s->p0, s->pi0);
    *out = s;
    return ERFX_OK;
}

int erfx_sim_destroy(erfx_sim* s) {
    if (!s) return ERFX_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    graph_free(s);
    tuner_free(s);
    timer_free(s->timer);
    if (s->red_host) (void)hipHostFree(s->red_host);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->ctx) erfx_ctx_destroy(s->ctx);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return ERFX_OK;
}

int erfx_sim_layout(const erfx_sim* s, erfx_layout* out) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    return erfx_ctx_layout(s->ctx, out);
}

erfx_ctx* erfx_sim_ctx(erfx_sim* s) { return s ? s->ctx : nullptr; }
void* erfx_sim_stream(erfx_sim* s) { return s ? (void*)s->stream : nullptr; }

// device pointer of a sim field and its size in doubles
static double* field_ptr(erfx_sim* s, int id, long* count) {
    const int nc = s->ctx->P.cfg.ncons;
    int n = 1;
    double* p = nullptr;
    if (id == ERFX_D_RAIN_ACCUM) {
        if (count) *count = s->rain_accum ? (long)s->ctx->P.L.sz : 0;
        return s->rain_accum;
    }
    switch (id) {
        case ERFX_F_CONS: p = s->cons[s->cur]; n = nc; break;
        case ERFX_F_XVEL: p = s->vel[s->cur][0]; break;
        case ERFX_F_YVEL: p = s->vel[s->cur][1]; break;
        case ERFX_F_ZVEL: p = s->vel[s->cur][2]; break;
        case ERFX_F_R0: p = s->r0; break;
        case ERFX_F_P0: p = s->p0; break;
        case ERFX_F_PI0: p = s->pi0; break;
        case ERFX_D_PRIM: p = s->prim; n = nc - 1; break;
        case ERFX_D_PI_STAGE: p = s->pi_stage; break;
        case ERFX_D_FAST_COEFFS: p = s->fast_coeffs; n = 5; break;
        case ERFX_D_EDDY: p = s->eddy; n = 13; break;
        case ERFX_D_XVEL_OLD: p = s->vel[1 - s->cur][0]; break;
        case ERFX_D_YVEL_OLD: p = s->vel[1 - s->cur][1]; break;
        case ERFX_D_ZVEL_OLD: p = s->vel[1 - s->cur][2]; break;
        case ERFX_F_Z_ND: p = s->ctx->z_nd; break;
        case ERFX_D_DETJ: p = s->ctx->detJ; break;
        case ERFX_D_AX: p = s->ctx->ax; break;
        case ERFX_D_AY: p = s->ctx->ay; break;
        case ERFX_D_AZ: p = s->ctx->az; break;
        case ERFX_D_Z_CC: p = s->ctx->z_cc; break;
        default:
            if (id >= ERFX_D_STATE_BASE && id < ERFX_D_STATE_BASE + 20) {
                int st = (id - ERFX_D_STATE_BASE) / 4, c = (id - ERFX_D_STATE_BASE) % 4;
                erfx_state& S = s->S[st];
                double* f[4] = {S.cons, S.xmom, S.ymom, S.zmom};
                p = f[c];
                n = c == 0 ? nc : 1;
            }
    }
    if (count) *count = (long)n * s->ctx->P.L.ne;
    return p;
}

double* erfx_sim_field_ptr(erfx_sim* s, int id) { return s ? field_ptr(s, id, nullptr) : nullptr; }

long erfx_sim_field_size(const erfx_sim* s, int id) {
    if (!s) return -1;
    long n = 0;
    double* p = field_ptr(const_cast<erfx_sim*>(s), id, &n);
    return p ? n : -1;
}

int erfx_sim_upload(erfx_sim* s, int id, const double* host) {
    if (!s || !host) { set_error("null argument"); return ERFX_ERR_ARG; }
    long n = 0;
    double* p = field_ptr(s, id, &n);
    if (!p) { set_error("unknown field id"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipMemcpyAsync(p, host, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    if (id == ERFX_F_Z_ND) {   // new node heights: remake the metrics
        int rc = erfx_set_terrain(s->ctx, p);
        if (rc) return rc;
    }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    return ERFX_OK;
}

int erfx_sim_download(erfx_sim* s, int id, double* host) {
    if (!s || !host) { set_error("null argument"); return ERFX_ERR_ARG; }
    long n = 0;
    double* p = field_ptr(s, id, &n);
    if (!p) { set_error("unknown field id"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipMemcpyAsync(host, p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    ERFX_HIP(hipStreamSynchronize(s->stream));
    return ERFX_OK;
}

// One step: ERF::Advance (ERF_Advance.cpp:20-233), advance_dycore
// (ERF_advance_dycore.cpp:37-305) and MRISplitIntegrator::advance (ERF_MRI.H:187-334).
// slow_dt of the three RK stages as the MRI loop forms it, time_stage - time
// with time_stage = time + dt/3, time + dt/2, time + dt (ERF_MRI.H:241-299,
// ERF_TI_slow_rhs_fun.H:19):
// the sum rounds to the spacing of time, so the values vary with time
static void stage_dts(double time, double dt, double out[3]) {
    out[0] = (time + dt / 3.0) - time;
    out[1] = (time + dt / 2.0) - time;
    out[2] = (time + dt) - time;
}

static int sim_one_step(erfx_sim* s, double time, double dt, int ratio) {
    erfx_ctx* c = s->ctx;
    const Params& P = c->P;
    hipStream_t st = s->stream;
    const size_t ne = (size_t)P.L.ne;
    const int nc = P.cfg.ncons;
    // swap vars_old / vars_new (ERF_Advance.cpp:25)
    const int old = s->cur, nw = 1 - s->cur;
    double* cons_old = s->cons[old];
    double* uo = s->vel[old][0]; double* vo = s->vel[old][1]; double* wo = s->vel[old][2];
    double* un = s->vel[nw][0]; double* vn = s->vel[nw][1]; double* wn = s->vel[nw][2];
    erfx_state& So = s->S[0];
    erfx_state& Sn = s->S[1];
    erfx_state& Ss = s->S[2];
    erfx_state& Sc = s->S[3];
    erfx_state& Fs = s->S[4];
    So.cons = cons_old;
    Sn.cons = s->cons[nw];
    // FillPatch of the old state and V2M on the full velocity ghost width (:98-108)
    fill_patch(P, cons_old, uo, vo, wo, st);
    velocity_to_momentum(P, uo, vo, wo, P.ng_vel, P.ng_vel, cons_old, So.xmom, So.ymom, So.zmom, st);
    // advance_dycore
    if (P.use_turb) turbulent_viscosity(P, c->sw, cons_old, uo, vo, wo, s->eddy, st);
    {
        CopyList cl;
        cl.add(un, uo, (long)ne); cl.add(vn, vo, (long)ne); cl.add(wn, wo, (long)ne);
        copy_fields(cl, st);
    }
    apply_bcs(P, So, un, vn, wn, P.ng_cons, P.ng_vel, false, true, st);
    cons_to_prim(P, So.cons, s->prim, s->pi_stage, P.ng_cons, st);
    // MRI: S_new <- S_old (ERF_MRI.H:225-230)
    {
        CopyList cl;
        cl.add(Sn.cons, So.cons, (long)(nc * ne)); cl.add(Sn.xmom, So.xmom, (long)ne);
        cl.add(Sn.ymom, So.ymom, (long)ne); cl.add(Sn.zmom, So.zmom, (long)ne);
        copy_fields(cl, st);
    }
    const bool nosub = P.cfg.no_substepping != 0;
    const double sub_timestep = nosub ? 0.0 : dt / ratio;
    // the substep chain folds apply_bcs into the substep's kernels (on a
    // terrain grid the column kernel forms the top face's wall value as the
    // fill would, from the previous substep's momenta)
    const bool chain = fast_bcs_fusable(P) && P.bck[2][5] == EXT_DIR && !nosub;
    const double beta_s = P.cfg.implicit_substepping ? 0.1 : -1.0;
    double sdt[3];
    stage_dts(time, dt, sdt);
    for (int nrk = 0; nrk < 3; ++nrk) {
        int nsub = 1;
        double dtau = 0;
        if (nrk == 0) {
            if (P.cfg.force_stage1_single_substep || nosub) { nsub = 1; dtau = dt / 3.0; }
            else { nsub = ratio / 3; dtau = sub_timestep; }
        } else if (nrk == 1) {
            if (nosub) { nsub = 1; dtau = 0.5 * dt; }
            else { nsub = ratio / 2; dtau = sub_timestep; }
        } else {
            if (nosub) { nsub = 1; dtau = dt; }
            else { nsub = ratio; dtau = sub_timestep; }
        }
        if (nrk > 0) cons_to_prim(P, Sn.cons, s->prim, s->pi_stage, P.ng_cons, st);
        const double slow_dt = sdt[nrk];
        slow_rhs_pre(P, c->sw, nrk, slow_dt, Fs, Sn, s->prim, Sc, un, vn, wn, s->eddy, s->r0, s->p0, st);
        const double inv_fac = 1.0 / static_cast<double>(nsub);
        if (nosub) {
            // no_substep_fun (ERF_TI_no_substep_fun.H:1-146): the fast
            // variables from the slow RHS alone, then apply_bcs on S_sum with
            // its full ghost widths (fast only)
            no_substep(P, Ss, So, Fs, nsub * dtau, st);
            apply_bcs(P, Ss, un, vn, wn, P.ng_cons, P.ng_vel, true, false, st);
        }
        for (int ks = 0; ks < (nosub ? 0 : nsub); ++ks) {
            // On a laterally periodic box the apply_bcs between two substeps is
            // folded into the next substep (round trip of the momenta on read,
            // ghost layer 1 of rho and rho*theta written by the column kernel);
            // the full fill runs after the stage's last substep.
            FastChain ch;
            ch.rt_prev = chain && ks > 0;
            ch.ghost_writes = chain && ks < nsub - 1 && !P.terrain;   // the terrain substep reads images
            ch.w_lo = P.vel_extdir[2][2];
            ch.w_hi = P.vel_extdir[2][5];
            if (ks == 0) make_fast_coeffs(P, s->fast_coeffs, Sn, s->prim, s->pi_stage, s->r0, s->pi0, dtau, beta_s, st);
            if (P.terrain)
                fast_rhs_T(P, c->fw, c->tw, ks, nrk, Fs, ks == 0 ? So : Ss, Sn, s->prim, s->pi_stage, s->fast_coeffs,
                           Ss, Sc, dtau, beta_s, inv_fac, st, ch);
            else
                fast_rhs_N(P, c->fw, ks, nrk, Fs, ks == 0 ? So : Ss, Sn, s->prim, s->pi_stage, s->fast_coeffs, Ss, Sc,
                           dtau, beta_s, inv_fac, st, ch);
            if (!chain || ks == nsub - 1) apply_bcs(P, Ss, un, vn, wn, 1, 1, true, false, st);
        }
        slow_rhs_post(P, c->sw, nrk, slow_dt, Fs, So, Sn, Ss, s->prim, Sc, s->eddy, st);
        apply_bcs(P, Sn, un, vn, wn, P.ng_cons, P.ng_vel, false, false, st);
    }
    // advance_microphysics on the new state (ERF_Advance.cpp:172-174)
    if (P.moist) {
        if (s->samples) {
            kessler_samples(P, Sn.cons, wn, dt, s->samples + 0, s->samples_max, s->samples_count, c->kbad, st);
        }
        if (s->mlp) {
            int rc = emulator_space(c);
            if (rc) return rc;
            kessler_mlp_advance(P, s->mlp->net, s->mlp->sc, Sn.cons, wn, c->kfz, s->rain_accum, dt, c->eslot, c->ex,
                                c->ey, c->erows, c->ecount, c->ecount_host, c->kbad, st);
            s->mlp_rows += (long)c->ecount_host[0];
            s->mlp_fallback += (long)c->ecount_host[1];
        } else {
            kessler_advance(P, Sn.cons, c->kfz, s->rain_accum, dt, st, c->kbad);
        }
    }
    s->cur = nw;
    s->S[0].cons = s->cons[1 - s->cur];
    s->S[1].cons = s->cons[s->cur];
    return launch_status();
}

int erfx_sim_set_mlp(erfx_sim* s, erfx_mlp* m) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (m && s->ctx->P.cfg.moisture_type != ERFX_MOIST_KESSLER) { set_error("the learned step replaces Kessler with rain"); return ERFX_ERR_ARG; }
    if (m && !m->scaled) { set_error("set the network's scaling first (erfx_mlp_set_scaling)"); return ERFX_ERR_STATE; }
    s->mlp = m;
    s->mlp_rows = s->mlp_fallback = 0;
    return ERFX_OK;
}

int erfx_sim_mlp_counts(erfx_sim* s, long* n_mlp, long* n_fallback) {
    if (!s || !n_mlp || !n_fallback) { set_error("null argument"); return ERFX_ERR_ARG; }
    *n_mlp = s->mlp_rows;
    *n_fallback = s->mlp_fallback;
    return ERFX_OK;
}

int erfx_sim_set_sampling(erfx_sim* s, float* rows, long max_rows) {
    if (!s || max_rows < 0 || (max_rows > 0 && !rows)) { set_error("bad argument"); return ERFX_ERR_ARG; }
    if (s->ctx->P.cfg.moisture_type != ERFX_MOIST_KESSLER) { set_error("samples need the Kessler model with rain"); return ERFX_ERR_ARG; }
    if (!s->samples_count) {
        int rc = balloc(s->allocs, (void**)&s->samples_count, sizeof(unsigned long long));
        if (rc) return rc;
    }
    ERFX_HIP(hipMemset(s->samples_count, 0, sizeof(unsigned long long)));
    s->samples = max_rows > 0 ? rows : nullptr;
    s->samples_max = max_rows;
    return ERFX_OK;
}

int erfx_sim_sample_count(erfx_sim* s, long* count) {
    if (!s || !count) { set_error("null argument"); return ERFX_ERR_ARG; }
    unsigned long long h = 0;
    if (s->samples_count) {
        ERFX_HIP(hipStreamSynchronize(s->stream));
        ERFX_HIP(hipMemcpy(&h, s->samples_count, sizeof(h), hipMemcpyDeviceToHost));
    }
    *count = (long)h;
    return ERFX_OK;
}

int erfx_sim_step_async(erfx_sim* s, double t0, double dt, int nsteps) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (rayleigh_on(s->ctx->P.cfg) && !s->ctx->ray) {
        set_error("Rayleigh damping profiles not set (erfx_set_rayleigh on the sim's context)"); return ERFX_ERR_STATE;
    }
    if (!(dt > 0) || nsteps < 0) { set_error("dt must be positive and nsteps >= 0"); return ERFX_ERR_ARG; }
    const int ratio = s->ctx->P.cfg.mri_ratio;
    int n = 0;
    {
        int rc = tuner_poll(s);
        if (rc) return rc;
    }
    const bool graphable = s->graph_on && !s->mlp && !s->samples && !s->ctx->P.timer && !s->ctx->P.tuner;
    while (graphable && n + 2 <= nsteps) {
        double sdt[6];
        stage_dts(t0 + n * dt, dt, sdt);
        stage_dts(t0 + (n + 1) * dt, dt, sdt + 3);
        if (!s->gexec || s->g_dt != dt || s->g_ratio != ratio || s->g_cur != s->cur ||
            std::memcmp(sdt, s->g_sdt, sizeof(sdt)) != 0) {
            graph_free(s);
            const int cur0 = s->cur;
            ERFX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
            int rc = sim_one_step(s, t0 + n * dt, dt, ratio);
            if (!rc) rc = sim_one_step(s, t0 + (n + 1) * dt, dt, ratio);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { s->gexec = nullptr; return hip_fail(e, "hipGraphInstantiate"); }
            s->g_dt = dt; s->g_ratio = ratio; s->g_cur = cur0;
            std::memcpy(s->g_sdt, sdt, sizeof(sdt));
            // capture recorded the two steps without running them
        }
        ERFX_HIP(hipGraphLaunch(s->gexec, s->stream));
        n += 2;
    }
    for (; n < nsteps; ++n) {
        int rc = sim_one_step(s, t0 + n * dt, dt, ratio);
        if (rc) return rc;
    }
    return ERFX_OK;
}

int erfx_sim_autotune(erfx_sim* s, int on) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    tuner_free(s);
    // the terrain column kernel (erf_fast_rhs_T) has one form: nothing to time
    if (!on || s->ctx->P.terrain) return ERFX_OK;
    s->tuner = new ColTuner();
    for (hipEvent_t& e : s->tuner->ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) { tuner_free(s); return hip_fail(r, "hipEventCreate"); }
    }
    s->ctx->P.tuner = s->tuner;
    graph_free(s);
    return ERFX_OK;
}

int erfx_sim_column_variant(erfx_sim* s, int* variant, float* ms) {
    if (!s || !variant) { set_error("null argument"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    int rc = tuner_poll(s);
    if (rc) return rc;
    *variant = (s->tuner && s->tuner->chosen < 0) ? -1 : s->ctx->P.col_variant;
    if (ms)
        for (int v = 0; v < COL_NV; ++v) ms[v] = s->tuner ? s->tuner->ms[v] : 0.f;
    return ERFX_OK;
}

int erfx_sim_set_column_variant(erfx_sim* s, int variant) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (variant < 0 || variant >= COL_NV) { set_error("column variant must be 0.." + std::to_string(COL_NV - 1)); return ERFX_ERR_ARG; }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    tuner_free(s);
    s->ctx->P.col_variant = variant;
    graph_free(s);
    return ERFX_OK;
}

int erfx_sim_set_graph(erfx_sim* s, int on) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    s->graph_on = on != 0;
    if (!s->graph_on) graph_free(s);
    return ERFX_OK;
}

// ERF::estTimeStep (ERF_ComputeTimestep.cpp:49-233): the two maxima come from
// est_dt_maxima on the GPU; the rest is host arithmetic in ERF's order.
int erfx_sim_est_dt(erfx_sim* s, const erfx_dt_control* d, double* dt, int* mri_ratio) {
    if (!s || !d || !dt || !mri_ratio) { set_error("null argument"); return ERFX_ERR_ARG; }
    const Params& P = s->ctx->P;
    est_dt_maxima(P, s->cons[s->cur], s->vel[s->cur][0], s->vel[s->cur][1], s->vel[s->cur][2], s->red,
                  s->red_host, s->stream);
    ERFX_HIP(hipStreamSynchronize(s->stream));
    const double comp_inv = s->red_host[0], lowM_inv = s->red_host[1];
    const double estdt_comp = d->cfl / comp_inv;
    double estdt_lowM = 1.e20;
    if (lowM_inv > 0.0) estdt_lowM = d->cfl / lowM_inv;
    long ratio;
    if (d->fixed_dt > 0. && d->fixed_fast_dt > 0.) ratio = static_cast<long>(d->fixed_dt / d->fixed_fast_dt);
    else if (d->fixed_dt > 0.) ratio = static_cast<long>(std::ceil((d->fixed_dt / estdt_comp)));
    else ratio = (lowM_inv > 0.0) ? static_cast<long>(std::ceil((estdt_lowM / estdt_comp))) : 1;
    if (P.cfg.force_stage1_single_substep) {
        if (ratio % 2 != 0) ratio += 1;
    } else {
        if (ratio % 6 != 0) ratio = static_cast<int>(std::ceil(ratio / 6.0) * 6);
    }
    *mri_ratio = d->fixed_mri_ratio > 0 ? d->fixed_mri_ratio : (int)ratio;
    if (P.cfg.no_substepping) *mri_ratio = 0;   // no ratio without substeps (:200-223)
    *dt = d->fixed_dt > 0.0 ? d->fixed_dt : estdt_comp;
    return launch_status();
}

// ERF::Evolve (ERF.cpp:346-399) with ComputeDt (ERF_ComputeTimestep.cpp:9-44).
int erfx_sim_run(erfx_sim* s, double t0, int max_steps, const erfx_dt_control* d, double* t_end, int* steps_done) {
    if (!s || !d) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (!(d->cfl > 0) && !(d->fixed_dt > 0)) { set_error("need cfl > 0 or fixed_dt > 0"); return ERFX_ERR_ARG; }
    if (rayleigh_on(s->ctx->P.cfg) && !s->ctx->ray) {
        set_error("Rayleigh damping profiles not set (erfx_set_rayleigh on the sim's context)"); return ERFX_ERR_STATE;
    }
    double t = t0;
    double dt_prev = 1.e100;
    int done = 0;
    for (int step = 0; step < max_steps && t < d->stop_time; ++step) {
        double dt_tmp;
        int ratio;
        int rc = erfx_sim_est_dt(s, d, &dt_tmp, &ratio);
        if (rc) return rc;
        if (!s->ctx->P.cfg.no_substepping && (ratio < 2 || ratio % 2 != 0)) {
            set_error("MRI ratio must be even and >= 2"); return ERFX_ERR_ARG;
        }
        double dt_0 = dt_tmp;
        dt_tmp = std::min(dt_tmp, d->change_max * dt_prev);
        dt_0 = std::min(dt_0, 1 * dt_tmp);
        if (step == 0) dt_0 *= d->init_shrink;
        const double eps = 1.e-3 * dt_0;
        if (t + dt_0 > d->stop_time - eps) dt_0 = d->stop_time - t;
        rc = sim_one_step(s, t, dt_0, ratio);
        if (rc) return rc;
        t += dt_0;
        dt_prev = dt_0;
        ++done;
        if (t >= d->stop_time - 1.e-6 * dt_0) break;
    }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    if (t_end) *t_end = t;
    if (steps_done) *steps_done = done;
    return kessler_status(s->ctx);
}

// ERF::sum_integrated_quantities' volume-weighted sums (volWgtSumMF,
// ERF_WriteScalarProfiles.cpp:11-60, 378-410): total mass, rho*theta and scalar.
int erfx_sim_integrals(erfx_sim* s, double* mass, double* rhotheta, double* scalar) {
    if (!s || !mass || !rhotheta || !scalar) { set_error("null argument"); return ERFX_ERR_ARG; }
    volume_sums(s->ctx->P, s->cons[s->cur], s->red, s->red_host, s->stream);
    ERFX_HIP(hipStreamSynchronize(s->stream));
    *mass = s->red_host[0] + s->red_host[1];
    *rhotheta = s->red_host[2] + s->red_host[3];
    *scalar = s->red_host[4] + s->red_host[5];
    return launch_status();
}

// ---- checkpoint / restart -----------------------------------------------------
// The state that a step reads and that no step recomputes: cons, the three
// velocities, the base state, the node heights of a terrain grid, the
// accumulated surface rain and the Rayleigh profiles; every other buffer of
// the sim is made again from these at the start of each step.  The role of
// ERF::WriteCheckpointFile / ReadCheckpointFile (IO/ERF_Checkpoint.cpp:25-590),
// which write vars_new, the base state and z_phys_nd per level with AMReX's
// VisMF; here one text Header and one raw little-endian FP64 file.
static const char* CHK_MAGIC = "erfx-checkpoint-1";

static std::vector<std::pair<int, long>> chk_fields(erfx_sim* s) {
    std::vector<std::pair<int, long>> f;
    for (int id : {ERFX_F_CONS, ERFX_F_XVEL, ERFX_F_YVEL, ERFX_F_ZVEL, ERFX_F_R0, ERFX_F_P0, ERFX_F_PI0}) {
        long n = 0;
        field_ptr(s, id, &n);
        f.push_back({id, n});
    }
    if (s->ctx->P.terrain) f.push_back({ERFX_F_Z_ND, s->ctx->P.L.ne});
    if (s->rain_accum) f.push_back({ERFX_D_RAIN_ACCUM, s->ctx->P.L.sz});
    return f;
}

int erfx_sim_write_checkpoint(erfx_sim* s, const char* dir, double time, int step) {
    if (!s || !dir) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (mkdir(dir, 0755) != 0 && errno != EEXIST) {
        set_error(std::string("cannot create checkpoint directory ") + dir); return ERFX_ERR_ARG;
    }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    const Params& P = s->ctx->P;
    const auto fields = chk_fields(s);
    const std::string d(dir);
    FILE* h = std::fopen((d + "/Header").c_str(), "w");
    if (!h) { set_error("cannot write " + d + "/Header"); return ERFX_ERR_ARG; }
    std::fprintf(h, "%s\n%d %d %d %d %ld %ld %ld\n%d %d %d\n%a\n%d\n%zu\n", CHK_MAGIC, P.L.nx, P.L.ny, P.L.nz,
                 P.L.ng, P.L.sy, P.L.sz, P.L.ne, P.cfg.ncons, P.cfg.moisture_type, P.cfg.use_terrain, time, step,
                 s->ctx->ray_prof.size());
    for (const auto& f : fields) std::fprintf(h, "%d %ld\n", f.first, f.second);
    std::fprintf(h, "# time %.17g, step %d\n", time, step);
    std::fclose(h);
    FILE* b = std::fopen((d + "/state.bin").c_str(), "wb");
    if (!b) { set_error("cannot write " + d + "/state.bin"); return ERFX_ERR_ARG; }
    std::vector<double> host;
    for (const auto& f : fields) {
        host.resize((size_t)f.second);
        double* p = field_ptr(s, f.first, nullptr);
        hipError_t e = hipMemcpy(host.data(), p, host.size() * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { std::fclose(b); return hip_fail(e, "checkpoint download"); }
        if (std::fwrite(host.data(), sizeof(double), host.size(), b) != host.size()) {
            std::fclose(b); set_error("short write of " + d + "/state.bin"); return ERFX_ERR_ARG;
        }
    }
    const auto& rp = s->ctx->ray_prof;
    if (!rp.empty() && std::fwrite(rp.data(), sizeof(double), rp.size(), b) != rp.size()) {
        std::fclose(b); set_error("short write of " + d + "/state.bin"); return ERFX_ERR_ARG;
    }
    if (std::fclose(b) != 0) { set_error("cannot close " + d + "/state.bin"); return ERFX_ERR_ARG; }
    return ERFX_OK;
}

int erfx_sim_read_checkpoint(erfx_sim* s, const char* dir, double* time, int* step) {
    if (!s || !dir || !time || !step) { set_error("null argument"); return ERFX_ERR_ARG; }
    const Params& P = s->ctx->P;
    const std::string d(dir);
    FILE* h = std::fopen((d + "/Header").c_str(), "r");
    if (!h) { set_error("cannot read " + d + "/Header"); return ERFX_ERR_ARG; }
    char magic[64] = {0};
    int nx, ny, nz, ng, nc, mt, ut, st;
    long sy, sz, ne;
    double t;
    size_t nray = 0;
    int got = std::fscanf(h, "%63s %d %d %d %d %ld %ld %ld %d %d %d %la %d %zu", magic, &nx, &ny, &nz, &ng, &sy, &sz,
                          &ne, &nc, &mt, &ut, &t, &st, &nray);
    const auto fields = chk_fields(s);
    bool ok = got == 14 && std::strcmp(magic, CHK_MAGIC) == 0;
    for (const auto& f : fields) {
        int id = -1;
        long n = -1;
        ok = ok && std::fscanf(h, "%d %ld", &id, &n) == 2 && id == f.first && n == f.second;
    }
    std::fclose(h);
    if (!ok || nx != P.L.nx || ny != P.L.ny || nz != P.L.nz || ng != P.L.ng || sy != P.L.sy || sz != P.L.sz ||
        ne != P.L.ne || nc != P.cfg.ncons || mt != P.cfg.moisture_type || ut != P.cfg.use_terrain ||
        (nray != 0 && nray != 4 * (size_t)(nz + 1))) {
        set_error("checkpoint " + d + " does not match this sim's grid, components or fields");
        return ERFX_ERR_ARG;
    }
    FILE* b = std::fopen((d + "/state.bin").c_str(), "rb");
    if (!b) { set_error("cannot read " + d + "/state.bin"); return ERFX_ERR_ARG; }
    std::vector<double> host;
    for (const auto& f : fields) {
        host.resize((size_t)f.second);
        if (std::fread(host.data(), sizeof(double), host.size(), b) != host.size()) {
            std::fclose(b); set_error("short read of " + d + "/state.bin"); return ERFX_ERR_ARG;
        }
        int rc = erfx_sim_upload(s, f.first, host.data());   // node heights remake the metrics
        if (rc) { std::fclose(b); return rc; }
    }
    std::vector<double> rp(nray);
    if (nray && std::fread(rp.data(), sizeof(double), nray, b) != nray) {
        std::fclose(b); set_error("short read of " + d + "/state.bin"); return ERFX_ERR_ARG;
    }
    std::fclose(b);
    if (nray) {
        const size_t n1 = (size_t)nz + 1;
        int rc = erfx_set_rayleigh(s->ctx, rp.data(), rp.data() + n1, rp.data() + 2 * n1, rp.data() + 3 * n1);
        if (rc) return rc;
    }
    *time = t;
    *step = st;
    return ERFX_OK;
}

int erfx_sim_sync(erfx_sim* s) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    return kessler_status(s->ctx);
}

// Times the acoustic column solve, the step's dominant kernel, on up to
// 4096 launches after the reset.
int erfx_sim_timer_reset(erfx_sim* s) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    timer_free(s->timer);
    KTimer& t = s->timer;
    t.ev = new hipEvent_t[8192];
    for (int i = 0; i < 8192; ++i) {
        hipError_t e = hipEventCreate(&t.ev[i]);
        if (e != hipSuccess) { timer_free(t); return hip_fail(e, "hipEventCreate"); }
        t.cap = i + 1;
    }
    t.kernel = KT_FAST_COLUMN;
    s->ctx->P.timer = &t;
    return ERFX_OK;
}

int erfx_sim_timer_stop(erfx_sim* s) {
    if (!s) { set_error("null argument"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    s->ctx->P.timer = nullptr;
    return ERFX_OK;
}

int erfx_sim_timer_read(erfx_sim* s, char* name, int name_len, double* total_ms, long* launches, double* bytes) {
    if (!s || !name || name_len < 1 || !total_ms || !launches || !bytes) { set_error("null argument"); return ERFX_ERR_ARG; }
    ERFX_HIP(hipStreamSynchronize(s->stream));
    const KTimer& t = s->timer;
    double sum = 0;
    for (int i = 0; i + 1 < t.used; i += 2) {
        float ms = 0;
        ERFX_HIP(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]));
        sum += ms;
    }
    std::snprintf(name, (size_t)name_len, "%s", t.name);
    *total_ms = sum;
    *launches = t.used / 2;
    *bytes = t.bytes;
    return ERFX_OK;
}

int erfx_sim_step(erfx_sim* s, double t0, double dt, int nsteps) {
    int rc = erfx_sim_step_async(s, t0, dt, nsteps);
    if (rc) return rc;
    return erfx_sim_sync(s);
}

}  // extern "C"

// ---- batched ensembles ---------------------------------------------------------
// A member's step is recorded by running the sim driver with a Recorder
// active (every launch appends its work item); the n-th items of all members
// form one batched launch.  A recorded step is valid for the dt, stage time
// steps, buffer parities and column variant it was recorded with.
namespace {
struct BatchSlot {
    BatchLaunch fn;
    unsigned width, total;
    int n;
    size_t off;        // the items
    size_t soff;       // the members' first work-group ids
};
struct BatchStep {
    double dt = 0, sdt[3] = {};
    int variant = 0;
    std::vector<int> cur;            // members' buffer parity at the step's start
    std::vector<BatchSlot> slots;
    void* items = nullptr;           // device copy of every slot's items
    hipGraphExec_t gexec = nullptr;
};
void batch_free(BatchStep* b) {
    if (!b) return;
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->items) (void)hipFree(b->items);
    delete b;
}
constexpr size_t ENS_KEEP = 4;       // recorded steps kept (two parities x two stage-dt sets)
}  // namespace

// Members are split into groups of about equal cell count, each batched on a
// stream of its own: the groups' launches overlap, so the tail of one batched
// kernel is filled by another group's work.
struct EnsGroup {
    std::vector<erfx_sim*> m;
    hipStream_t stream = nullptr;
    std::vector<BatchStep*> rec;     // most recently used last
};

struct erfx_ensemble {
    std::vector<erfx_sim*> m;
    std::vector<EnsGroup> g;
    bool graph_on = false;
    int variant = COL_DEFAULT;
    int tune_next = -1;              // next variant to time, -1 when not tuning
    float tune_ms[COL_NV] = {};
    long cells = 0;
    long launches = 0;
};

// the members' parity after one step
static void sim_flip(erfx_sim* s) {
    s->cur = 1 - s->cur;
    s->S[0].cons = s->cons[1 - s->cur];
    s->S[1].cons = s->cons[s->cur];
}

static long sim_cells(const erfx_sim* s) { return (long)s->ctx->P.L.nx * s->ctx->P.L.ny * s->ctx->P.L.nz; }

static int ens_record(const std::vector<erfx_sim*>& mem, double time, double dt, int variant, BatchStep** out) {
    const int n = (int)mem.size();
    std::vector<std::vector<RecItem>> seq((size_t)n);
    auto b = new BatchStep();
    b->dt = dt;
    stage_dts(time, dt, b->sdt);
    b->variant = variant;
    for (int k = 0; k < n; ++k) {
        erfx_sim* s = mem[(size_t)k];
        Params& P = s->ctx->P;
        const int cur = s->cur;
        double* c0 = s->S[0].cons;
        double* c1 = s->S[1].cons;
        KTimer* tm = P.timer;
        ColTuner* tu = P.tuner;
        const int var = P.col_variant;
        P.timer = nullptr;
        P.tuner = nullptr;
        P.col_variant = variant;
        Recorder r;
        g_rec = &r;
        int rc = sim_one_step(s, time, dt, s->ctx->P.cfg.mri_ratio);
        g_rec = nullptr;
        s->cur = cur;
        s->S[0].cons = c0;
        s->S[1].cons = c1;
        P.timer = tm;
        P.tuner = tu;
        P.col_variant = var;
        b->cur.push_back(cur);
        if (!rc && r.refused) {
            set_error(std::string("not batchable: ") + r.refused);
            rc = ERFX_ERR_UNSUPPORTED;
        }
        if (rc) { batch_free(b); return rc; }
        seq[(size_t)k] = std::move(r.items);
    }
    const size_t ns = seq[0].size();
    for (int k = 1; k < n; ++k) {
        bool same = seq[(size_t)k].size() == ns;
        for (size_t i = 0; same && i < ns; ++i) {
            const RecItem& a = seq[0][i];
            const RecItem& c = seq[(size_t)k][i];
            same = a.launch == c.launch && a.width == c.width && a.size == c.size;
        }
        if (!same) {
            batch_free(b);
            set_error("ensemble members record different kernel sequences (scheme options or MRI ratio differ)");
            return ERFX_ERR_UNSUPPORTED;
        }
    }
    // item arrays, one per slot, each 256-byte aligned; t0 made cumulative
    size_t bytes = 0;
    for (size_t i = 0; i < ns; ++i) {
        BatchSlot sl{seq[0][i].launch, seq[0][i].width, 0u, n, (bytes + 255) / 256 * 256, 0};
        sl.soff = (sl.off + (size_t)n * seq[0][i].size + 255) / 256 * 256;
        bytes = sl.soff + (size_t)n * sizeof(unsigned);
        b->slots.push_back(sl);
    }
    std::vector<unsigned char> host(bytes);
    for (size_t i = 0; i < ns; ++i) {
        BatchSlot& sl = b->slots[i];
        unsigned long long t0 = 0;
        for (int k = 0; k < n; ++k) {
            const RecItem& it = seq[(size_t)k][i];
            BHdr h;
            std::memcpy(&h, it.bytes.data(), sizeof(BHdr));
            h.t0 = (unsigned)t0;
            const unsigned start = (unsigned)t0;
            std::memcpy(host.data() + sl.soff + (size_t)k * sizeof(unsigned), &start, sizeof(unsigned));
            t0 += h.tiles;
            unsigned char* dst = host.data() + sl.off + (size_t)k * it.size;
            std::memcpy(dst, it.bytes.data(), it.size);
            std::memcpy(dst, &h, sizeof(BHdr));
        }
        if (t0 * sl.width >= (1ull << 32)) {
            batch_free(b);
            set_error("batched launch exceeds the grid limit");
            return ERFX_ERR_UNSUPPORTED;
        }
        sl.total = (unsigned)t0;
    }
    hipError_t err = hipMalloc(&b->items, bytes > 0 ? bytes : 1);
    if (err != hipSuccess) { b->items = nullptr; batch_free(b); return hip_fail(err, "hipMalloc"); }
    err = hipMemcpy(b->items, host.data(), bytes, hipMemcpyHostToDevice);
    if (err != hipSuccess) { batch_free(b); return hip_fail(err, "hipMemcpy"); }
    *out = b;
    return ERFX_OK;
}

static int ens_launch(const BatchStep* b, hipStream_t st) {
    for (size_t i = 0; i < b->slots.size(); ++i) {
        const BatchSlot& sl = b->slots[i];
        const unsigned char* base = static_cast<const unsigned char*>(b->items);
        if (sl.total)
            sl.fn(base + sl.off, reinterpret_cast<const unsigned*>(base + sl.soff), sl.n, sl.total, sl.width, st);
    }
    return ERFX_OK;
}

// one group's step: the recorded launch list for (dt, stage dts, parities,
// variant), recorded on a miss, replayed directly or as a graph
static int ens_group_step(EnsGroup& G, bool graph_on, double time, double dt, int variant, long* launches) {
    double sdt[3];
    stage_dts(time, dt, sdt);
    BatchStep* b = nullptr;
    for (size_t i = 0; i < G.rec.size() && !b; ++i) {
        BatchStep* r = G.rec[i];
        bool hit = r->dt == dt && r->variant == variant && std::memcmp(r->sdt, sdt, sizeof(sdt)) == 0;
        for (size_t k = 0; hit && k < G.m.size(); ++k) hit = r->cur[k] == G.m[k]->cur;
        if (hit) {
            b = r;
            G.rec.erase(G.rec.begin() + (long)i);
            G.rec.push_back(b);
        }
    }
    if (!b) {
        int rc = ens_record(G.m, time, dt, variant, &b);
        if (rc) return rc;
        if (G.rec.size() >= ENS_KEEP) {
            batch_free(G.rec.front());
            G.rec.erase(G.rec.begin());
        }
        G.rec.push_back(b);
    }
    for (const BatchSlot& sl : b->slots) *launches += sl.total ? 1 : 0;
    if (graph_on) {
        if (!b->gexec) {
            ERFX_HIP(hipStreamBeginCapture(G.stream, hipStreamCaptureModeRelaxed));
            int rc = ens_launch(b, G.stream);
            if (rc) return rc;
            hipGraph_t g = nullptr;
            hipError_t err = hipStreamEndCapture(G.stream, &g);
            if (err != hipSuccess) return hip_fail(err, "hipStreamEndCapture");
            err = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (err != hipSuccess) { b->gexec = nullptr; return hip_fail(err, "hipGraphInstantiate"); }
        }
        ERFX_HIP(hipGraphLaunch(b->gexec, G.stream));
    } else {
        int rc = ens_launch(b, G.stream);
        if (rc) return rc;
    }
    return launch_status();
}

static int ens_sync(erfx_ensemble* e) {
    for (EnsGroup& G : e->g) ERFX_HIP(hipStreamSynchronize(G.stream));
    return ERFX_OK;
}

static int ens_one_step(erfx_ensemble* e, double time, double dt) {
    const int variant = e->tune_next >= 0 ? e->tune_next : e->variant;
    const bool timing = e->tune_next >= 0;
    std::chrono::steady_clock::time_point t0;
    if (timing) {
        int rc = ens_sync(e);
        if (rc) return rc;
        t0 = std::chrono::steady_clock::now();
    }
    long nl = 0;
    for (EnsGroup& G : e->g) {
        int rc = ens_group_step(G, e->graph_on, time, dt, variant, &nl);
        if (rc) return rc;
    }
    e->launches = nl;
    if (timing) {
        int rc = ens_sync(e);
        if (rc) return rc;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        e->tune_ms[e->tune_next] = (float)ms;
        if (++e->tune_next == COL_NV) {
            int best = 0;
            for (int v = 1; v < COL_NV; ++v)
                if (e->tune_ms[v] < e->tune_ms[best]) best = v;
            e->variant = best;
            e->tune_next = -1;
        }
    }
    for (erfx_sim* s : e->m) sim_flip(s);
    return ERFX_OK;
}

static void ens_clear(erfx_ensemble* e) {
    for (EnsGroup& G : e->g) {
        for (BatchStep* b : G.rec) batch_free(b);
        G.rec.clear();
    }
}

// split the members into k groups: largest first, each to the group with the
// fewest cells so far
static int ens_regroup(erfx_ensemble* e, int k) {
    int rc = ens_sync(e);
    if (rc) return rc;
    ens_clear(e);
    for (EnsGroup& G : e->g)
        if (G.stream) (void)hipStreamDestroy(G.stream);
    e->g.clear();
    k = std::max(1, std::min(k, (int)e->m.size()));
    e->g.resize((size_t)k);
    for (EnsGroup& G : e->g) {
        hipError_t err = hipStreamCreate(&G.stream);
        if (err != hipSuccess) { G.stream = nullptr; return hip_fail(err, "hipStreamCreate"); }
    }
    std::vector<erfx_sim*> order = e->m;
    std::stable_sort(order.begin(), order.end(),
                     [](const erfx_sim* a, const erfx_sim* b) { return sim_cells(a) > sim_cells(b); });
    std::vector<long> load((size_t)k, 0);
    for (erfx_sim* s : order) {
        const size_t g = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        e->g[g].m.push_back(s);
        load[g] += sim_cells(s);
    }
    return ERFX_OK;
}

extern "C" {

int erfx_ensemble_create(erfx_sim* const* members, int n, erfx_ensemble** out) {
    if (!members || n < 1 || !out) { set_error("null argument or no members"); return ERFX_ERR_ARG; }
    *out = nullptr;
    for (int k = 0; k < n; ++k) {
        if (!members[k]) { set_error("null member"); return ERFX_ERR_ARG; }
        for (int l = 0; l < k; ++l)
            if (members[l] == members[k]) { set_error("a sim appears twice in the ensemble"); return ERFX_ERR_ARG; }
    }
    auto e = new erfx_ensemble();
    e->m.assign(members, members + n);
    for (erfx_sim* s : e->m) e->cells += sim_cells(s);
    int rc = ens_regroup(e, 2);
    if (rc) { erfx_ensemble_destroy(e); return rc; }
    *out = e;
    return ERFX_OK;
}

int erfx_ensemble_destroy(erfx_ensemble* e) {
    if (!e) return ERFX_OK;
    (void)ens_sync(e);
    ens_clear(e);
    for (EnsGroup& G : e->g)
        if (G.stream) (void)hipStreamDestroy(G.stream);
    delete e;
    return ERFX_OK;
}

int erfx_ensemble_set_streams(erfx_ensemble* e, int nstreams) {
    if (!e) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (nstreams < 1 || nstreams > 8) { set_error("streams must be 1..8"); return ERFX_ERR_ARG; }
    return ens_regroup(e, nstreams);
}

int erfx_ensemble_step(erfx_ensemble* e, double t0, double dt, int nsteps) {
    if (!e) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (!(dt > 0) || nsteps < 0) { set_error("dt must be positive and nsteps >= 0"); return ERFX_ERR_ARG; }
    for (erfx_sim* s : e->m) {
        if (rayleigh_on(s->ctx->P.cfg) && !s->ctx->ray) {
            set_error("Rayleigh damping profiles not set on a member"); return ERFX_ERR_STATE;
        }
        if (s->mlp || s->samples) { set_error("not batchable: learned microphysics or row sampling on a member"); return ERFX_ERR_UNSUPPORTED; }
        ERFX_HIP(hipStreamSynchronize(s->stream));   // members' own uploads and steps done
        // a flag latched by the member's own earlier steps belongs to those steps
        int rc = kessler_status(s->ctx);
        if (rc) return rc;
    }
    for (int n = 0; n < nsteps; ++n) {
        int rc = ens_one_step(e, t0 + n * dt, dt);
        if (rc) { (void)ens_sync(e); return rc; }
    }
    int rc = ens_sync(e);
    if (rc) return rc;
    // a member whose temperature left the saturation fit's range (ERF aborts
    // there): report it as this call's error, naming the member, and clear
    // every member's flag so no later call of a member reports it again
    int bad = -1;
    for (size_t m = 0; m < e->m.size(); ++m) {
        if (kessler_status(e->m[m]->ctx) != ERFX_OK && bad < 0) bad = (int)m;
    }
    if (bad >= 0) {
        set_error("ensemble member " + std::to_string(bad) +
                  ": Kessler temperature below -85 C, outside the saturation fit (ERF_Microphysics_Utils.H:62)");
        return ERFX_ERR_STATE;
    }
    return ERFX_OK;
}

int erfx_ensemble_set_graph(erfx_ensemble* e, int on) {
    if (!e) { set_error("null argument"); return ERFX_ERR_ARG; }
    e->graph_on = on != 0;
    return ERFX_OK;
}

int erfx_ensemble_autotune(erfx_ensemble* e, int on) {
    if (!e) { set_error("null argument"); return ERFX_ERR_ARG; }
    bool terrain = false;
    for (erfx_sim* s : e->m) terrain = terrain || s->ctx->P.terrain;
    e->tune_next = (on && !terrain) ? 0 : -1;
    for (float& v : e->tune_ms) v = 0.f;
    return ERFX_OK;
}

int erfx_ensemble_set_column_variant(erfx_ensemble* e, int variant) {
    if (!e) { set_error("null argument"); return ERFX_ERR_ARG; }
    if (variant < 0 || variant >= COL_NV) { set_error("column variant out of range"); return ERFX_ERR_ARG; }
    e->variant = variant;
    e->tune_next = -1;
    return ERFX_OK;
}

int erfx_ensemble_column_variant(erfx_ensemble* e, int* variant, float* ms) {
    if (!e || !variant) { set_error("null argument"); return ERFX_ERR_ARG; }
    *variant = e->tune_next >= 0 ? -1 : e->variant;
    if (ms)
        for (int v = 0; v < COL_NV; ++v) ms[v] = e->tune_ms[v];
    return ERFX_OK;
}

int erfx_ensemble_info(erfx_ensemble* e, long* launches_per_step, long* cells) {
    if (!e || !launches_per_step || !cells) { set_error("null argument"); return ERFX_ERR_ARG; }
    *launches_per_step = e->launches;
    *cells = e->cells;
    return ERFX_OK;
}

}  // extern "C"
