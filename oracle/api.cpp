// api.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  C entry points for ctypes.
// Field ids follow include/erfx.h (ERFX_F_*) plus the diagnostic ids
// ERFX_D_* for the integrator's internal buffers.
#include "oracle.hpp"
#include <cstdio>

using namespace orc;

namespace {
thread_local std::string g_err;

Field* field_by_id(Level& L, int id) {
    switch (id) {
        case ERFX_F_CONS: return &L.cons;
        case ERFX_F_XVEL: return &L.xvel;
        case ERFX_F_YVEL: return &L.yvel;
        case ERFX_F_ZVEL: return &L.zvel;
        case ERFX_F_R0: return &L.r0;
        case ERFX_F_P0: return &L.p0;
        case ERFX_F_PI0: return &L.pi0;
        case ERFX_D_PRIM: return &L.prim;
        case ERFX_D_PI_STAGE: return &L.pi_stage;
        case ERFX_D_FAST_COEFFS: return &L.fast_coeffs;
        case ERFX_D_EDDY: return &L.eddy;
        case ERFX_D_XVEL_OLD: return &L.xvel_old;
        case ERFX_D_YVEL_OLD: return &L.yvel_old;
        case ERFX_D_ZVEL_OLD: return &L.zvel_old;
        case ERFX_D_XMOM_SRC: return &L.xmom_src;
        case ERFX_D_YMOM_SRC: return &L.ymom_src;
        case ERFX_D_ZMOM_SRC: return &L.zmom_src;
        default: break;
    }
    if (L.terrain) switch (id) {
        case ERFX_F_Z_ND: return &L.z_nd;
        case ERFX_D_DETJ: return &L.detJ;
        case ERFX_D_AX: return &L.ax;
        case ERFX_D_AY: return &L.ay;
        case ERFX_D_AZ: return &L.az;
        case ERFX_D_Z_CC: return &L.z_cc;
        default: break;
    }
    if (id >= ERFX_D_STATE_BASE && id < ERFX_D_STATE_BASE + 20) {
        int s = (id - ERFX_D_STATE_BASE) / 4, c = (id - ERFX_D_STATE_BASE) % 4;
        StateF* st[5] = {&L.S_old, &L.S_new, &L.S_sum, &L.S_scratch, &L.F_slow};
        Field* f[4] = {&st[s]->cons, &st[s]->xmom, &st[s]->ymom, &st[s]->zmom};
        return f[c];
    }
    return nullptr;
}

// The library's whole-step driver keeps vars_new's cons and the integrator's
// S_new cons in one buffer (ERF aliases them, ERF_Advance.cpp:157); the
// oracle keeps two copies, so a hook that writes one of them mirrors it into
// the other.
void mirror_snew_to_cons(Level& L) { L.cons.v = L.S_new.cons.v; }
void mirror_cons_to_snew(Level& L) { L.S_new.cons.v = L.cons.v; }

template <class F> int guard(F&& f) {
    try { f(); return 0; }
    catch (const std::exception& e) { g_err = e.what(); return -1; }
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// out[i] = x[i]**y[i] with the EOS's pow (for the accuracy test against libm).
void orc_pow_n(const double* x, const double* y, double* out, long n) {
    for (long i = 0; i < n; ++i) out[i] = erfx_pow(x[i], y[i]);
}

// ERF's column solve on faces 0..n: factor (A, B, C) as make_fast_coeffs does,
// then the Thomas sweep of erf_fast_rhs_N.  B is overwritten with the factors.
void orc_column_solve(int n, const double* A, double* B, const double* C, const double* rhs, double* soln) {
    column_factor(n, A, B, C, 1);
    column_solve(n, A, B, C, rhs, soln, 1);
}

// Face value of the centred/upwind interpolation family from the six cell
// values s[0..5] at i-3..i+2 (lo face of cell i), advecting velocity sign `upw`.
double orc_face_interp(int type, const double* s, double upw, double upw_frac) {
    return face_interp(type, s[5], s[4], s[3], s[2], s[1], s[0], upw, upw_frac);
}

// Rayleigh reference profiles (nz+1 values each)
int orc_set_rayleigh(void* h, const double* ub, const double* vb, const double* wb, const double* tb) {
    Level& L = *static_cast<Level*>(h);
    const int n1 = L.L.nz + 1;
    L.ray_ubar.assign(ub, ub + n1);
    L.ray_vbar.assign(vb, vb + n1);
    L.ray_wbar.assign(wb, wb + n1);
    L.ray_thetabar.assign(tb, tb + n1);
    return 0;
}

// initial state of the reference's DensityCurrent regression case
int orc_init_density_current(void* h, double x_c, double z_c, double x_r, double z_r, double T_pert) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] { L.init_density_current(x_c, z_c, x_r, z_r, T_pert); });
}

// initial states of the ScalarAdvDiff and TaylorGreenVortex regression cases
int orc_init_scalar_advdiff(void* h, const double* params) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] { L.init_scalar_advdiff(params); });
}
int orc_init_taylor_green(void* h, double rho_0, double V_0, double M_0) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] { L.init_taylor_green(rho_0, V_0, M_0); });
}

int orc_init_couette_poiseuille(void* h, double rho_0, double T_0, double u_0, double v_0, double w_0, int prob_type,
                                int uniform) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] { L.init_couette_poiseuille(rho_0, T_0, u_0, v_0, w_0, prob_type, uniform != 0); });
}

int orc_init_ekman(void* h, double rho_0, double T_0, double u_g, double rot_period, double mu) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] { L.init_ekman(rho_0, T_0, u_g, rot_period, mu); });
}

// ERF::Evolve's time loop (Source/ERF.cpp:346-399) with ComputeDt
// (ERF_ComputeTimestep.cpp:9-44) before every step: dt from estTimeStep,
// limited to change_max times the previous dt, init_shrink on the first step,
// cut to land on stop_time; the MRI ratio from estTimeStep unless
// fixed_mri_ratio > 0 (ERF_advance_dycore.cpp:299).  *t_end gets the final time.
int orc_run(void* h, double t0, int nsteps, double cfl, double fixed_dt, double fixed_fast_dt,
            int fixed_mri_ratio, double change_max, double init_shrink, double stop_time, double* t_end) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] {
        double t = t0;
        double dt_prev = 1.e100;
        const int ratio0 = L.cfg.mri_ratio;
        for (int step = 0; step < nsteps && t < stop_time; ++step) {
            long ratio = 1;
            double dt_tmp = L.est_time_step(cfl, fixed_dt, fixed_fast_dt, ratio);
            double dt_0 = dt_tmp;
            dt_tmp = std::min(dt_tmp, change_max * dt_prev);
            dt_0 = std::min(dt_0, 1 * dt_tmp);
            if (step == 0) dt_0 *= init_shrink;
            const double eps = 1.e-3 * dt_0;
            if (t + dt_0 > stop_time - eps) dt_0 = stop_time - t;
            L.cfg.mri_ratio = fixed_mri_ratio > 0 ? fixed_mri_ratio : (int)ratio;
            L.advance(t, dt_0);
            t += dt_0;
            dt_prev = dt_0;
            if (t >= stop_time - 1.e-6 * dt_0) break;
        }
        L.cfg.mri_ratio = ratio0;
        if (t_end) *t_end = t;
    });
}

void* orc_create(const erfx_config* cfg, const erfx_layout* lay) {
    auto* L = new Level();
    try { L->init(*cfg, *lay); }
    catch (const std::exception& e) { g_err = e.what(); delete L; return nullptr; }
    return L;
}
void orc_destroy(void* h) { delete static_cast<Level*>(h); }

long orc_field_size(void* h, int id) {
    Level& L = *static_cast<Level*>(h);
    if (id == ERFX_D_RAIN_ACCUM) return (long)L.rain_accum.size();
    Field* f = field_by_id(L, id);
    return f ? (long)f->v.size() : -1;
}
int orc_set(void* h, int id, const double* src) {
    Level& L = *static_cast<Level*>(h);
    if (id == ERFX_D_RAIN_ACCUM) { std::memcpy(L.rain_accum.data(), src, L.rain_accum.size() * sizeof(double)); return 0; }
    Field* f = field_by_id(L, id);
    if (!f) { g_err = "bad field id"; return -1; }
    if (id == ERFX_F_Z_ND) return guard([&] { L.set_terrain(src); });
    std::memcpy(f->v.data(), src, f->v.size() * sizeof(double));
    return 0;
}
int orc_get(void* h, int id, double* dst) {
    Level& L = *static_cast<Level*>(h);
    if (id == ERFX_D_RAIN_ACCUM) { std::memcpy(dst, L.rain_accum.data(), L.rain_accum.size() * sizeof(double)); return 0; }
    Field* f = field_by_id(L, id);
    if (!f) { g_err = "bad field id"; return -1; }
    std::memcpy(dst, f->v.data(), f->v.size() * sizeof(double));
    return 0;
}

int orc_step(void* h, double t0, double dt, int nsteps) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] { for (int s = 0; s < nsteps; ++s) L.advance(t0 + s * dt, dt); });
}

// ---- hook-level entry points on the level's own buffers ------------------
int orc_cons_to_prim(void* h, int cons_id, int ng) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] { L.cons_to_prim(L.A(*field_by_id(L, cons_id)), ng); });
}
// state_index: 0 S_old, 1 S_new, 2 S_sum
int orc_apply_bcs(void* h, int state_index, int ng_cons, int ng_vel, int fast_only, int synced) {
    Level& L = *static_cast<Level*>(h);
    StateF* st[3] = {&L.S_old, &L.S_new, &L.S_sum};
    return guard([&] {
        L.apply_bcs(L.A(*st[state_index]), L.A(L.xvel), L.A(L.yvel), L.A(L.zvel), ng_cons, ng_vel,
                    fast_only != 0, synced != 0);
        if (state_index == 1) mirror_snew_to_cons(L);
    });
}
int orc_make_fast_coeffs(void* h, double dtau, double beta_s) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] {
        L.make_fast_coeffs(L.A(L.fast_coeffs), L.A(L.S_new), L.A(L.prim), L.A(L.pi_stage), dtau, beta_s);
    });
}
int orc_fast_rhs_N(void* h, int step, int nrk, double dtau, double beta_s, double facinv) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] {
        StateA prev = step == 0 ? L.A(L.S_old) : L.A(L.S_sum);
        L.fast_rhs_N(step, nrk, L.A(L.F_slow), prev, L.A(L.S_new), L.A(L.prim), L.A(L.pi_stage),
                     L.A(L.fast_coeffs), L.A(L.S_sum), L.A(L.S_scratch), dtau, beta_s, facinv);
    });
}
int orc_fast_rhs_T(void* h, int step, int nrk, double dtau, double beta_s, double facinv) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] {
        if (!L.terrain) throw std::runtime_error("erf_fast_rhs_T needs a terrain grid");
        StateA prev = step == 0 ? L.A(L.S_old) : L.A(L.S_sum);
        L.fast_rhs_T(step, nrk, L.A(L.F_slow), prev, L.A(L.S_new), L.A(L.prim), L.A(L.pi_stage),
                     L.A(L.fast_coeffs), L.A(L.S_sum), L.A(L.S_scratch), dtau, beta_s, facinv);
    });
}
int orc_turbulent_viscosity(void* h) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] {
        L.turbulent_viscosity(L.A(L.S_old.cons), L.A(L.xvel_old), L.A(L.yvel_old), L.A(L.zvel_old));
    });
}
int orc_slow_rhs_pre(void* h, int nrk, double slow_dt) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] {
        L.make_sources(L.A(L.S_new), L.A(L.prim));
        L.make_mom_sources(L.A(L.S_new), L.A(L.prim));
        L.slow_rhs_pre(nrk, slow_dt, L.A(L.F_slow), L.A(L.S_new), L.A(L.prim), L.A(L.S_scratch),
                       L.A(L.xvel), L.A(L.yvel), L.A(L.zvel));
    });
}
int orc_slow_rhs_post(void* h, int nrk, double slow_dt) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] {
        std::fill(L.cc_src.v.begin(), L.cc_src.v.end(), 0.0);
        L.slow_rhs_post(nrk, slow_dt, L.A(L.F_slow), L.A(L.S_old), L.A(L.S_new), L.A(L.S_sum),
                        L.A(L.prim), L.A(L.S_scratch), L.A(L.xvel), L.A(L.yvel));
        mirror_snew_to_cons(L);
    });
}

// Kessler on the level's vars_new cons with step dt (advance_microphysics)
int orc_kessler(void* h, double dt) {
    Level& L = *static_cast<Level*>(h);
    return guard([&] {
        L.kessler_advance(L.A(L.cons), dt);
        mirror_cons_to_snew(L);
    });
}

// saturation functions (K, mbar) for the closed-form tests
int orc_saturation(const double* t, const double* p, long n, double* es, double* des, double* qs, double* dqs) {
    return guard([&] {
        for (long i = 0; i < n; ++i) {
            es[i] = esatw(t[i]);
            des[i] = dtesatw(t[i]);
            qs[i] = qsatw(t[i], p[i]);
            dqs[i] = dtqsatw(t[i], p[i]);
        }
    });
}

// moist EOS: temperature from rho, rho*theta and qv (ERF_EOS.H:42-51)
void orc_temperature(const double* rho, const double* rt, const double* qv, long n, double* T) {
    for (long i = 0; i < n; ++i) T[i] = getTgivenRandRTh(rho[i], rt[i], qv[i]);
}

}  // extern "C"
