// fast.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  Acoustic substep on a flat
// grid: make_fast_coeffs and erf_fast_rhs_N, GPU branches of the reference.
#include "oracle.hpp"

namespace orc {

// make_fast_coeffs (Source/TimeIntegration/ERF_make_fast_coeffs.cpp:27-294),
// flat branch :158-200, GPU column recurrence :214-242, inversion :287-290.
// fc components: 0 = A, 1 = B (1/B after inversion on interior faces), 2 = C, 3 = P, 4 = Q.
void Level::make_fast_coeffs(Arr fc, StateA stage, Arr prim_, Arr pi_st, double dtau, double beta_s) {
    const double beta_2 = 0.5 * (1.0 + beta_s);
    const double c_v = cfg.c_p - R_d;
    const double dzi = dxInv[2];
    const double halfg = std::abs(0.5 * (-gravity));
    Arr r0a = A(r0), pi0a = A(pi0);
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    Arr sc = stage.cons;
    const int q1 = ERFX_RHOQ1 - 1, q2 = ERFX_RHOQ2 - 1;
    Arr J = A(detJ);
    loop(Box(0, 0, 1, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double rhobar_lo = r0a(i, j, k - 1), rhobar_hi = r0a(i, j, k);
        double pibar_lo = pi0a(i, j, k - 1), pibar_hi = pi0a(i, j, k);
        double pi_c = 0.5 * (pi_st(i, j, k - 1) + pi_st(i, j, k));
        double detJ_on_kface = 1.0;
        double coeff_P, coeff_Q;
        if (terrain) {   // :113-137
            detJ_on_kface = 0.5 * (J(i, j, k) + J(i, j, k - 1));
            double inv_detJ_on_kface = 1. / detJ_on_kface;
            coeff_P = -Gamma * R_d * dzi * inv_detJ_on_kface
                    + halfg * R_d * rhobar_hi / (c_v * pibar_hi * sc(i, j, k, 1));
            coeff_Q = Gamma * R_d * dzi * inv_detJ_on_kface
                    + halfg * R_d * rhobar_lo / (c_v * pibar_lo * sc(i, j, k - 1, 1));
        } else {
            coeff_P = -Gamma * R_d * dzi
                    + halfg * R_d * rhobar_hi / (c_v * pibar_hi * sc(i, j, k, 1));
            coeff_Q = Gamma * R_d * dzi
                    + halfg * R_d * rhobar_lo / (c_v * pibar_lo * sc(i, j, k - 1, 1));
        }
        coeff_P *= pi_c;
        coeff_Q *= pi_c;
        fc(i, j, k, 3) = coeff_P;
        fc(i, j, k, 4) = coeff_Q;
        if (moist) {   // :183-188
            double q = 0.5 * (prim_(i, j, k, q1) + prim_(i, j, k - 1, q1) + prim_(i, j, k, q2) + prim_(i, j, k - 1, q2));
            coeff_P /= (1.0 + q);
            coeff_Q /= (1.0 + q);
        }
        double theta_t_lo  = 0.5 * (prim_(i, j, k - 2, 0) + prim_(i, j, k - 1, 0));
        double theta_t_mid = 0.5 * (prim_(i, j, k - 1, 0) + prim_(i, j, k, 0));
        double theta_t_hi  = 0.5 * (prim_(i, j, k, 0) + prim_(i, j, k + 1, 0));
        double D = dtau * dtau * beta_2 * beta_2 * dzi;
        fc(i, j, k, 0) = D * (halfg - coeff_Q * theta_t_lo);
        fc(i, j, k, 2) = D * (-halfg + coeff_P * theta_t_hi);
        fc(i, j, k, 1) = (terrain ? detJ_on_kface : 1.0) + D * (coeff_Q - coeff_P) * theta_t_mid;
    });
    const bool top_outflow = (cfg.bc[5] == ERFX_BC_OUTFLOW);
    loop(Box(0, 0, 0, nx - 1, ny - 1, 0), [&](int i, int j, int) {
        fc(i, j, 0, 0) = 0.0; fc(i, j, 0, 1) = 1.0; fc(i, j, 0, 2) = 0.0;
        fc(i, j, nz, 0) = 0.0; fc(i, j, nz, 1) = 1.0; fc(i, j, nz, 2) = 0.0;
        if (top_outflow) fc(i, j, nz, 0) = -1.0;
        column_factor(nz, &fc(i, j, 0, 0), &fc(i, j, 0, 1), &fc(i, j, 0, 2), L.sz);
    });
}

// Column recurrence of make_fast_coeffs, GPU branch (ERF_make_fast_coeffs.cpp:214-242)
// followed by the inversion (:287-290): on faces 0..n with stride `st`, B[k]
// becomes bet_k and then 1/bet_k on the interior faces 1..n-1; B[0] and B[n]
// keep bet_0 and bet_n (ERF does not invert them).
void column_factor(int n, const double* A, double* B, const double* C, long st) {
    double bet = B[0];
    for (int k = 1; k <= n; k++) {
        double gam = C[(k - 1) * st] / bet;
        bet = B[k * st] - A[k * st] * gam;
        B[k * st] = bet;
    }
    for (int k = 1; k <= n - 1; k++) B[k * st] = 1.0 / B[k * st];
}

// Thomas sweep of erf_fast_rhs_N, GPU branch (ERF_fast_rhs_N.cpp:435-453), on the
// factors column_factor leaves: forward soln_k = (rhs_k - A_k soln_{k-1}) iB_k,
// backward soln_k -= (C_k iB_k) soln_{k+1}.
void column_solve(int n, const double* A, const double* iB, const double* C, const double* rhs,
                  double* soln, long st) {
    soln[0] = rhs[0] * iB[0];
    for (int k = 1; k <= n; k++) soln[k * st] = (rhs[k * st] - A[k * st] * soln[(k - 1) * st]) * iB[k * st];
    for (int k = n - 1; k >= 0; k--) soln[k * st] -= (C[k * st] * iB[k * st]) * soln[(k + 1) * st];
}

// erf_fast_rhs_N (Source/TimeIntegration/ERF_fast_rhs_N.cpp:36-614), GPU
// branch of the column solve (:435-453).  For step > 0, prev and data alias.
void Level::fast_rhs_N(int step, int nrk, StateA slow_rhs, StateA prev, StateA stage, Arr prim_,
                       Arr pi_st, Arr fc, StateA data, StateA scratch, double dtau, double beta_s,
                       double facinv) {
    const double beta_1 = 0.5 * (1.0 - beta_s);
    const double beta_2 = 0.5 * (1.0 + beta_s);
    const double beta_d = 0.1;
    const double dxi = dxInv[0], dyi = dxInv[1], dzi = dxInv[2];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const bool implicit = cfg.implicit_substepping != 0;
    const int q1 = ERFX_RHOQ1 - 1, q2 = ERFX_RHOQ2 - 1;

    // temporaries (valid-size plus what the stencils read)
    Field f_drt, f_drw, f_extrap, f_trhs, f_tx, f_ty, f_rhs, f_soln;
    f_drt.alloc(L, 1); f_drw.alloc(L, 1); f_extrap.alloc(L, 1); f_trhs.alloc(L, 2);
    f_tx.alloc(L, 1); f_ty.alloc(L, 1); f_rhs.alloc(L, 1); f_soln.alloc(L, 1);
    Arr drt = f_drt.a(L), drw = f_drw.a(L), extrap = f_extrap.a(L), trhs = f_trhs.a(L);
    Arr tx = f_tx.a(L), ty = f_ty.a(L), RHS = f_rhs.a(L), soln = f_soln.a(L);
    Arr lagged = scratch.cons;

    // :115-152
    loop(Box(-1, -1, -1, nx, ny, nz), [&](int i, int j, int k) {
        drt(i, j, k) = prev.cons(i, j, k, 1) - stage.cons(i, j, k, 1);
        if (step == 0) extrap(i, j, k) = drt(i, j, k);
        else extrap(i, j, k) = drt(i, j, k) + beta_d * (drt(i, j, k) - lagged(i, j, k, 1));
        lagged(i, j, k, 1) = drt(i, j, k);
    });
    loop(Box(-1, -1, 0, nx, ny, nz), [&](int i, int j, int k) {
        drw(i, j, k) = prev.zmom(i, j, k) - stage.zmom(i, j, k);
    });

    // :161-257
    Arr ax = scratch.xmom, ay = scratch.ymom, az = scratch.zmom;
    if (nrk == 0 && step == 0) {
        loop(Box(0, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) {
            double new_drho_u = prev.xmom(i, j, k) - stage.xmom(i, j, k) + dtau * slow_rhs.xmom(i, j, k);
            ax(i, j, k) += facinv * new_drho_u;
            tx(i, j, k) = stage.xmom(i, j, k) + new_drho_u;
        });
        loop(Box(0, 0, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) {
            double new_drho_v = prev.ymom(i, j, k) - stage.ymom(i, j, k) + dtau * slow_rhs.ymom(i, j, k);
            ay(i, j, k) += facinv * new_drho_v;
            ty(i, j, k) = stage.ymom(i, j, k) + new_drho_v;
        });
    } else {
        loop(Box(0, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) {
            double gpx = (extrap(i, j, k) - extrap(i - 1, j, k)) * dxi;
            if (moist) {   // :215-219
                double q = 0.5 * (prim_(i, j, k, q1) + prim_(i - 1, j, k, q1) + prim_(i, j, k, q2) + prim_(i - 1, j, k, q2));
                gpx /= (1.0 + q);
            }
            double pi_c = 0.5 * (pi_st(i - 1, j, k) + pi_st(i, j, k));
            double fast_rhs_rho_u = -Gamma * R_d * pi_c * gpx;
            double new_drho_u = prev.xmom(i, j, k) - stage.xmom(i, j, k)
                              + dtau * fast_rhs_rho_u + dtau * slow_rhs.xmom(i, j, k);
            ax(i, j, k) += facinv * new_drho_u;
            tx(i, j, k) = stage.xmom(i, j, k) + new_drho_u;
        });
        loop(Box(0, 0, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) {
            double gpy = (extrap(i, j, k) - extrap(i, j - 1, k)) * dyi;
            if (moist) {   // :238-242
                double q = 0.5 * (prim_(i, j, k, q1) + prim_(i, j - 1, k, q1) + prim_(i, j, k, q2) + prim_(i, j - 1, k, q2));
                gpy /= (1.0 + q);
            }
            double pi_c = 0.5 * (pi_st(i, j - 1, k) + pi_st(i, j, k));
            double fast_rhs_rho_v = -Gamma * R_d * pi_c * gpy;
            double new_drho_v = prev.ymom(i, j, k) - stage.ymom(i, j, k)
                              + dtau * fast_rhs_rho_v + dtau * slow_rhs.ymom(i, j, k);
            ay(i, j, k) += facinv * new_drho_v;
            ty(i, j, k) = stage.ymom(i, j, k) + new_drho_v;
        });
    }

    // :328-357 horizontal flux divergence of the momentum increments
    Arr pr = prim_;
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double xflux_lo = (tx(i, j, k) - stage.xmom(i, j, k));
        double xflux_hi = (tx(i + 1, j, k) - stage.xmom(i + 1, j, k));
        double yflux_lo = (ty(i, j, k) - stage.ymom(i, j, k));
        double yflux_hi = (ty(i, j + 1, k) - stage.ymom(i, j + 1, k));
        trhs(i, j, k, 0) = (xflux_hi - xflux_lo) * dxi + (yflux_hi - yflux_lo) * dyi;
        trhs(i, j, k, 1) = ((xflux_hi * (pr(i, j, k, 0) + pr(i + 1, j, k, 0)) -
                             xflux_lo * (pr(i, j, k, 0) + pr(i - 1, j, k, 0))) * dxi +
                            (yflux_hi * (pr(i, j, k, 0) + pr(i, j + 1, k, 0)) -
                             yflux_lo * (pr(i, j, k, 0) + pr(i, j - 1, k, 0))) * dyi) * 0.5;
    });

    const double halfg = std::abs(0.5 * (-gravity));
    Arr sr = slow_rhs.cons;
    // :374-415
    loop(Box(0, 0, 1, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double coeff_P = fc(i, j, k, 3);
        double coeff_Q = fc(i, j, k, 4);
        if (moist) {   // :379-384
            double q = 0.5 * (pr(i, j, k, q1) + pr(i, j, k - 1, q1) + pr(i, j, k, q2) + pr(i, j, k - 1, q2));
            coeff_P /= (1.0 + q);
            coeff_Q /= (1.0 + q);
        }
        double theta_t_lo  = 0.5 * (pr(i, j, k - 2, 0) + pr(i, j, k - 1, 0));
        double theta_t_mid = 0.5 * (pr(i, j, k - 1, 0) + pr(i, j, k, 0));
        double theta_t_hi  = 0.5 * (pr(i, j, k, 0) + pr(i, j, k + 1, 0));
        double Omega_kp1 = prev.zmom(i, j, k + 1) - stage.zmom(i, j, k + 1);
        double Omega_k   = prev.zmom(i, j, k) - stage.zmom(i, j, k);
        double Omega_km1 = prev.zmom(i, j, k - 1) - stage.zmom(i, j, k - 1);
        double old_drho_k   = prev.cons(i, j, k, 0) - stage.cons(i, j, k, 0);
        double old_drho_km1 = prev.cons(i, j, k - 1, 0) - stage.cons(i, j, k - 1, 0);
        double R0_tmp = coeff_P * drt(i, j, k) + coeff_Q * drt(i, j, k - 1)
                      - halfg * (old_drho_k + old_drho_km1);
        double R1_tmp = halfg * (-sr(i, j, k, 0) - sr(i, j, k - 1, 0) + trhs(i, j, k, 0) + trhs(i, j, k - 1, 0))
                      + (coeff_P * (sr(i, j, k, 1) - trhs(i, j, k, 1)) +
                         coeff_Q * (sr(i, j, k - 1, 1) - trhs(i, j, k - 1, 1)));
        R1_tmp += beta_1 * dzi * ((Omega_kp1 - Omega_km1) * halfg
                                  - (Omega_kp1 * theta_t_hi - Omega_k * theta_t_mid) * coeff_P
                                  - (Omega_k * theta_t_mid - Omega_km1 * theta_t_lo) * coeff_Q);
        RHS(i, j, k) = Omega_k + dtau * (slow_rhs.zmom(i, j, k) + R0_tmp + dtau * beta_2 * R1_tmp);
    });
    // :423-432
    loop(Box(0, 0, 0, nx - 1, ny - 1, 0), [&](int i, int j, int) {
        RHS(i, j, 0) = prev.zmom(i, j, 0) - stage.zmom(i, j, 0) + dtau * slow_rhs.zmom(i, j, 0);
        RHS(i, j, nz) = prev.zmom(i, j, nz) - stage.zmom(i, j, nz) + dtau * slow_rhs.zmom(i, j, nz);
    });
    // :435-464 (GPU branch)
    Arr cz = data.zmom;
    if (implicit) {
        loop(Box(0, 0, 0, nx - 1, ny - 1, 0), [&](int i, int j, int) {
            column_solve(nz, &fc(i, j, 0, 0), &fc(i, j, 0, 1), &fc(i, j, 0, 2), &RHS(i, j, 0), &soln(i, j, 0), L.sz);
            for (int k = 0; k <= nz; k++) cz(i, j, k) = stage.zmom(i, j, k) + soln(i, j, k);
        });
    } else {
        loop(Box(0, 0, 0, nx - 1, ny - 1, 0), [&](int i, int j, int) {
            for (int k = 0; k <= nz; k++) {
                soln(i, j, k) = RHS(i, j, k);
                cz(i, j, k) = stage.zmom(i, j, k) + soln(i, j, k);
            }
        });
    }
    // :518-536
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double zflux_lo = beta_2 * soln(i, j, k) + beta_1 * drw(i, j, k);
        double zflux_hi = beta_2 * soln(i, j, k + 1) + beta_1 * drw(i, j, k + 1);
        az(i, j, k) += facinv * zflux_lo / (1.0 * 1.0);
        if (k == nz - 1) az(i, j, k + 1) += facinv * zflux_hi / (1.0 * 1.0);
        trhs(i, j, k, 0) += dzi * (zflux_hi - zflux_lo);
        trhs(i, j, k, 1) += 0.5 * dzi * (zflux_hi * (pr(i, j, k) + pr(i, j, k + 1))
                                       - zflux_lo * (pr(i, j, k) + pr(i, j, k - 1)));
    });
    // :566-613
    Arr cc = data.cons;
    if (step == 0) {
        loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
            cc(i, j, k, 0) = prev.cons(i, j, k, 0) + dtau * (sr(i, j, k, 0) - trhs(i, j, k, 0));
            cc(i, j, k, 1) = prev.cons(i, j, k, 1) + dtau * (sr(i, j, k, 1) - trhs(i, j, k, 1));
        });
    } else {
        loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
            cc(i, j, k, 0) += dtau * (sr(i, j, k, 0) - trhs(i, j, k, 0));
            cc(i, j, k, 1) += dtau * (sr(i, j, k, 1) - trhs(i, j, k, 1));
        });
    }
    loop(Box(0, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) { data.xmom(i, j, k) = tx(i, j, k); });
    loop(Box(0, 0, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) { data.ymom(i, j, k) = ty(i, j, k); });
}

}  // namespace orc
