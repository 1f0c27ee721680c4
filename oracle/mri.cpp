// mri.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  One ERF time step on a
// single level: ERF::Advance -> advance_dycore -> MRISplitIntegrator::advance.
#include "oracle.hpp"

namespace orc {

static void copy_field(Field& dst, const Field& src) { dst.v = src.v; dst.ncomp = src.ncomp; }

// ERF::advance_dycore (Source/TimeIntegration/ERF_advance_dycore.cpp:37-305)
// with the MRI stage loop of MRISplitIntegrator::advance (ERF_MRI.H:187-334)
// and the hook bodies of ERF_TI_slow_rhs_fun.H / ERF_TI_fast_rhs_fun.H inlined.
void Level::advance_dycore(double old_time, double dt) {
    const bool use_diff = cfg.molec_diff_type != ERFX_MOLEC_NONE || cfg.les_type != ERFX_LES_NONE;
    const bool use_kturb = cfg.les_type != ERFX_LES_NONE;
    Arr u_old = A(xvel_old), v_old = A(yvel_old), w_old = A(zvel_old);
    StateA So = A(S_old), Sn = A(S_new), Ss = A(S_sum), Sc = A(S_scratch), Fs = A(F_slow);
    Arr xv = A(xvel), yv = A(yvel), zv = A(zvel);

    // once-per-step strain + eddy viscosity (:114-223)
    if (use_diff && use_kturb) turbulent_viscosity(So.cons, u_old, v_old, w_old);
    else if (use_diff) {
        // the strain is still computed into Tau (consumed by make_tau_terms' overwrite)
    }
    // V2M with one ghost (:256-264)
    int g1[3] = {1, 1, 1}, g0[3] = {1, 1, 0};
    velocity_to_momentum(u_old, g1, v_old, g1, w_old, g0, So.cons, So.xmom, So.ymom, So.zmom);
    copy_field(xvel, xvel_old); copy_field(yvel, yvel_old); copy_field(zvel, zvel_old);
    apply_bcs(So, xv, yv, zv, ng_cons, ng_vel, false, true);
    cons_to_prim(So.cons, ng_cons);

    // MRISplitIntegrator::advance
    const bool nosub = cfg.no_substepping != 0;
    const int ratio = nosub ? 2 : cfg.mri_ratio;   // not used without substepping
    if (!nosub && !(ratio > 1 && ratio % 2 == 0)) throw std::runtime_error("mri ratio must be even and > 1");
    copy_field(S_new.cons, S_old.cons); copy_field(S_new.xmom, S_old.xmom);
    copy_field(S_new.ymom, S_old.ymom); copy_field(S_new.zmom, S_old.zmom);
    const double timestep = dt;
    const double sub_timestep = timestep / ratio;
    double time_stage = old_time, old_time_stage;
    const double beta_s = cfg.implicit_substepping ? 0.1 : -1.0;
    for (int nrk = 0; nrk < 3; nrk++) {
        old_time_stage = time_stage;
        int nsubsteps = 1; double dtau = 0;
        if (nrk == 0) {
            if (cfg.force_stage1_single_substep || nosub) { nsubsteps = 1; dtau = timestep / 3.0; }
            else { nsubsteps = ratio / 3; dtau = sub_timestep; }
            time_stage = old_time + timestep / 3.0;
        }
        if (nrk == 1) {
            if (nosub) { nsubsteps = 1; dtau = 0.5 * timestep; }
            else { nsubsteps = ratio / 2; dtau = sub_timestep; }
            time_stage = old_time + timestep / 2.0;
        }
        if (nrk == 2) {
            if (nosub) { nsubsteps = 1; dtau = timestep; }
            else { nsubsteps = ratio; dtau = sub_timestep; }
            time_stage = old_time + timestep;
        }
        if (nrk > 0) cons_to_prim(Sn.cons, ng_cons);   // pre_update
        // slow_rhs_fun_pre: make_sources, make_mom_sources, erf_slow_rhs_pre
        double slow_dt = time_stage - old_time;
        make_sources(Sn, A(prim));
        make_mom_sources(Sn, A(prim));
        slow_rhs_pre(nrk, slow_dt, Fs, Sn, A(prim), Sc, xv, yv, zv);
        double inv_fac = 1.0 / static_cast<double>(nsubsteps);
        if (nosub) {
            // no_substep_fun (ERF_TI_no_substep_fun.H:1-146), fixed or no terrain:
            // S_sum = S_old + dt_stage * F_slow for rho, rho*theta and the
            // momenta, then apply_bcs with the full ghost widths (fast only)
            no_substep(Ss, So, Fs, nsubsteps * dtau);
            apply_bcs(Ss, xv, yv, zv, ng_cons, ng_vel, true, false);
        }
        for (int ks = 0; ks < (nosub ? 0 : nsubsteps); ++ks) {
            // ERF_TI_fast_rhs_fun.H:122-175: erf_fast_rhs_T on a static terrain grid
            if (ks == 0) make_fast_coeffs(A(fast_coeffs), Sn, A(prim), A(pi_stage), dtau, beta_s);
            StateA prev = (ks == 0) ? So : Ss;
            if (terrain)
                fast_rhs_T(ks, nrk, Fs, prev, Sn, A(prim), A(pi_stage), A(fast_coeffs), Ss, Sc, dtau, beta_s, inv_fac);
            else
                fast_rhs_N(ks, nrk, Fs, prev, Sn, A(prim), A(pi_stage), A(fast_coeffs), Ss, Sc, dtau, beta_s, inv_fac);
            apply_bcs(Ss, xv, yv, zv, 1, 1, true, false);
        }
        slow_rhs_post(nrk, slow_dt, Fs, So, Sn, Ss, A(prim), Sc, xv, yv);
        apply_bcs(Sn, xv, yv, zv, ng_cons, ng_vel, false, false);   // post_update
        (void)old_time_stage;
    }
    (void)use_diff;
}

// no_substep_fun (Source/TimeIntegration/ERF_TI_no_substep_fun.H:88-126), fixed
// or no terrain: the fast variables advanced with the slow right-hand side
// alone, cons comps 0 and 1 on the valid cells and the momenta on the nodal
// tile boxes.
void Level::no_substep(StateA ssum, StateA sold, StateA fslow, double slow_dt) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        for (int n = 0; n < 2; ++n)
            ssum.cons(i, j, k, n) = sold.cons(i, j, k, n) + slow_dt * (fslow.cons(i, j, k, n));
    });
    loop(Box(0, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) {
        ssum.xmom(i, j, k) = sold.xmom(i, j, k) + slow_dt * fslow.xmom(i, j, k);
    });
    loop(Box(0, 0, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) {
        ssum.ymom(i, j, k) = sold.ymom(i, j, k) + slow_dt * fslow.ymom(i, j, k);
    });
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz), [&](int i, int j, int k) {
        ssum.zmom(i, j, k) = sold.zmom(i, j, k) + slow_dt * fslow.zmom(i, j, k);
    });
}

// ERF::Advance, level 0, dycore and microphysics (Source/TimeIntegration/ERF_Advance.cpp:20-233):
// the old state is the previous new state; FillPatch + V2M of the old
// velocities, then the dycore step writes the new cons and velocities.
void Level::advance(double time, double dt) {
    copy_field(cons_old_vars, cons);
    copy_field(xvel_old, xvel); copy_field(yvel_old, yvel); copy_field(zvel_old, zvel);
    Arr c = A(cons_old_vars), u = A(xvel_old), v = A(yvel_old), w = A(zvel_old);
    fill_patch(c, u, v, w);
    StateA So = A(S_old);
    int gv[3] = {ng_vel, ng_vel, ng_vel};
    velocity_to_momentum(u, gv, v, gv, w, gv, c, So.xmom, So.ymom, So.zmom);
    copy_field(S_old.cons, cons_old_vars);   // cons_mf copy incl. ghosts (:127-128)
    advance_dycore(time, dt);
    kessler_advance(A(S_new.cons), dt);       // advance_microphysics (:172-174) on vars_new
    copy_field(cons, S_new.cons);             // state_new cons aliases vars_new
}

}  // namespace orc

namespace orc {

// ERF::estTimeStep (Source/TimeIntegration/ERF_ComputeTimestep.cpp:49-233),
// one level, flat grid, compressible with substepping: the acoustic CFL
// estimate cfl / max((|u|+c)/dx, (|v|+c)/dy) (implicit substepping leaves w
// out), the advective estimate cfl / max(|u|/dx, |v|/dy, |w|/dz), and the MRI
// ratio from their quotient (or from fixed_dt / fixed_fast_dt), made even.
double Level::est_time_step(double cfl, double fixed_dt, double fixed_fast_dt, long& ratio) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    Arr c = A(cons), u = A(xvel), v = A(yvel), w = A(zvel);
    // l_implicit_substepping (:78): false for explicit substeps and for none
    const bool implicit = cfg.implicit_substepping != 0 && !cfg.no_substepping;
    const double dzinv = 1.0 / dz_min;   // ERF_ComputeTimestep.cpp:66
    double comp_inv = -1.e100, lowM_inv = -1.e100;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                // average_face_to_cellcenter
                double uc = 0.5 * (u(i, j, k) + u(i + 1, j, k));
                double vc = 0.5 * (v(i, j, k) + v(i, j + 1, k));
                double wc = 0.5 * (w(i, j, k) + w(i, j, k + 1));
                const double rho = c(i, j, k, 0);
                const double rhotheta = c(i, j, k, 1);
                double pressure = getPgivenRTh(rhotheta);
                double cs = std::sqrt(Gamma * pressure / rho);
                double a = (std::abs(uc) + cs) * dxInv[0];
                double b = (std::abs(vc) + cs) * dxInv[1];
                double e = (std::abs(wc) + cs) * dzinv;
                if (implicit) {
                    if (nx > 1 && ny > 1) comp_inv = std::max({a, b, comp_inv});
                    else if (nx > 1) comp_inv = std::max(a, comp_inv);
                    else comp_inv = std::max(b, comp_inv);
                } else {
                    if (nx > 1 && ny > 1) comp_inv = std::max({a, b, e, comp_inv});
                    else if (nx > 1) comp_inv = std::max({a, e, comp_inv});
                    else if (ny > 1) comp_inv = std::max({b, e, comp_inv});
                    else comp_inv = std::max(e, comp_inv);
                }
                lowM_inv = std::max({std::abs(uc) * dxInv[0], std::abs(vc) * dxInv[1], std::abs(wc) * dzinv,
                                     lowM_inv});
            }
    double estdt_comp = cfl / comp_inv;
    double estdt_lowM = 1.e20;
    if (lowM_inv > 0.0) estdt_lowM = cfl / lowM_inv;
    if (cfg.no_substepping) return fixed_dt > 0.0 ? fixed_dt : estdt_comp;   // no ratio (:200-223)
    if (fixed_dt > 0. && fixed_fast_dt > 0.) ratio = static_cast<long>(fixed_dt / fixed_fast_dt);
    else if (fixed_dt > 0.) ratio = static_cast<long>(std::ceil((fixed_dt / estdt_comp)));
    else ratio = (lowM_inv > 0.0) ? static_cast<long>(std::ceil((estdt_lowM / estdt_comp))) : 1;
    if (cfg.force_stage1_single_substep) {
        if (ratio % 2 != 0) ratio += 1;
    } else {
        if (ratio % 6 != 0) ratio = static_cast<int>(std::ceil(ratio / 6.0) * 6);
    }
    return fixed_dt > 0.0 ? fixed_dt : estdt_comp;
}

}  // namespace orc
