// terrain.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  Terrain-following grid:
// the metrics of the node heights and the acoustic substep erf_fast_rhs_T.
#include "oracle.hpp"

namespace orc {

// make_J, make_areas, make_zcc (Source/Utils/ERF_TerrainMetrics.cpp:583-715):
// detJ on cells, ax / ay on x- / y-faces, az = 1 on z-faces, z_cc on cells,
// on the valid box grown by one, k >= 0; ERF sets every metric to 1 first
// (ERF_make_new_arrays.cpp:87-96), which the k = -1 layer keeps.
// z_nodes: every node of the layout, ghosts included (periodic images
// laterally, linear extrapolation in z), as the caller built it.
void Level::set_terrain(const double* z_nodes) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    z_nd.alloc(L, 1);
    std::memcpy(z_nd.v.data(), z_nodes, z_nd.v.size() * sizeof(double));
    for (Field* f : {&detJ, &ax, &ay, &az, &z_cc}) {
        f->alloc(L, 1);
        std::fill(f->v.begin(), f->v.end(), 1.0);
    }
    std::fill(z_cc.v.begin(), z_cc.v.end(), 0.0);
    Arr z = A(z_nd), J = A(detJ), Ax = A(ax), Ay = A(ay), Zc = A(z_cc);
    const double dzInv = 1.0 / cfg.dz;
    loop(Box(-1, -1, 0, nx, ny, nz), [&](int i, int j, int k) {
        J(i, j, k) = .25 * dzInv * (z(i, j, k + 1) + z(i + 1, j, k + 1) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1)
                                    - z(i, j, k) - z(i + 1, j, k) - z(i, j + 1, k) - z(i + 1, j + 1, k));
        Zc(i, j, k) = .125 * (z(i, j, k) + z(i + 1, j, k) + z(i, j + 1, k) + z(i + 1, j + 1, k)
                              + z(i, j, k + 1) + z(i + 1, j, k + 1) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1));
    });
    loop(Box(-1, -1, 0, nx + 1, ny, nz), [&](int i, int j, int k) {
        Ax(i, j, k) = .5 * dzInv * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k) - z(i, j + 1, k));
    });
    loop(Box(-1, -1, 0, nx, ny + 1, nz), [&](int i, int j, int k) {
        Ay(i, j, k) = .5 * dzInv * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k) - z(i + 1, j, k));
    });
    double jmin = 1.e300;
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) { jmin = std::min(jmin, J(i, j, k)); });
    dz_min = cfg.dz * jmin;
    terrain = true;
}

// erf_fast_rhs_T (Source/TimeIntegration/ERF_fast_rhs_T.cpp:39-704), static
// terrain, unit map factors.  For step > 0, prev and data alias.
void Level::fast_rhs_T(int step, int nrk, StateA slow_rhs, StateA prev, StateA stage, Arr prim_, Arr pi_st,
                       Arr fc, StateA data, StateA scratch, double dtau, double beta_s, double facinv) {
    (void)nrk;
    const double beta_1 = 0.5 * (1.0 - beta_s);
    const double beta_2 = 0.5 * (1.0 + beta_s);
    const double beta_d = 0.1;
    const double dxi = dxInv[0], dyi = dxInv[1], dzi = dxInv[2];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int q1 = ERFX_RHOQ1 - 1, q2 = ERFX_RHOQ2 - 1;
    Arr z = A(z_nd), J = A(detJ);
    Arr cc = data.cons, lagged = scratch.cons;

    Field f_du, f_dv, f_dw, f_drho, f_drt, f_nu, f_nv, f_extrap, f_trhs, f_rhs, f_soln, f_omega;
    f_du.alloc(L, 1); f_dv.alloc(L, 1); f_dw.alloc(L, 1); f_drho.alloc(L, 1); f_drt.alloc(L, 1);
    f_nu.alloc(L, 1); f_nv.alloc(L, 1); f_extrap.alloc(L, 1); f_trhs.alloc(L, 2); f_rhs.alloc(L, 1);
    f_soln.alloc(L, 1); f_omega.alloc(L, 1);
    Arr du = f_du.a(L), dv = f_dv.a(L), dw = f_dw.a(L), drho = f_drho.a(L), drt = f_drt.a(L);
    Arr nu = f_nu.a(L), nv = f_nv.a(L), extrap = f_extrap.a(L), trhs = f_trhs.a(L);
    Arr RHS = f_rhs.a(L), soln = f_soln.a(L), om = f_omega.a(L);

    // :139-144
    if (step == 0) {
        loop(Box(-1, -1, -1, nx, ny, nz), [&](int i, int j, int k) {
            cc(i, j, k, 0) = prev.cons(i, j, k, 0);
            cc(i, j, k, 1) = prev.cons(i, j, k, 1);
        });
    }
    // :146-175: momentum increments on faces grown by one laterally; the
    // top-of-box layer is copied into the ghost layer above
    loop(Box(-1, -1, 0, nx + 1, ny, nz - 1), [&](int i, int j, int k) {
        du(i, j, k) = prev.xmom(i, j, k) - stage.xmom(i, j, k);
        if (k == nz - 1) du(i, j, k + 1) = du(i, j, k);
    });
    loop(Box(-1, -1, 0, nx, ny + 1, nz - 1), [&](int i, int j, int k) {
        dv(i, j, k) = prev.ymom(i, j, k) - stage.ymom(i, j, k);
        if (k == nz - 1) dv(i, j, k + 1) = dv(i, j, k);
    });
    loop(Box(-1, -1, 0, nx, ny, nz), [&](int i, int j, int k) {
        dw(i, j, k) = prev.zmom(i, j, k) - stage.zmom(i, j, k);
    });
    // :177-186
    loop(Box(-1, -1, -1, nx, ny, nz), [&](int i, int j, int k) {
        drho(i, j, k) = cc(i, j, k, 0) - stage.cons(i, j, k, 0);
        drt(i, j, k) = cc(i, j, k, 1) - stage.cons(i, j, k, 1);
        if (step == 0) extrap(i, j, k) = drt(i, j, k);
        else extrap(i, j, k) = drt(i, j, k) + beta_d * (drt(i, j, k) - lagged(i, j, k, 1));
    });
    // :196-204
    loop(Box(-1, -1, -1, nx, ny, nz), [&](int i, int j, int k) { lagged(i, j, k, 1) = drt(i, j, k); });

    // :262-331 horizontal momenta with the terrain-corrected pressure gradient
    Arr ax_m = scratch.xmom, ay_m = scratch.ymom, az_m = scratch.zmom;
    loop(Box(0, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) {
        double met_h_xi = met::h_xi_Iface(z, i, j, k, dxi);
        double met_h_zeta = met::h_zeta_Iface(z, i, j, k, dzi);
        double gp_xi = (extrap(i, j, k) - extrap(i - 1, j, k)) * dxi;
        double gp_zeta_on_iface = (k == 0)
            ? 0.5 * dzi * (extrap(i - 1, j, k + 1) + extrap(i, j, k + 1) - extrap(i - 1, j, k) - extrap(i, j, k))
            : 0.25 * dzi * (extrap(i - 1, j, k + 1) + extrap(i, j, k + 1) - extrap(i - 1, j, k - 1) - extrap(i, j, k - 1));
        double gpx = gp_xi - (met_h_xi / met_h_zeta) * gp_zeta_on_iface;
        gpx *= 1.0;
        if (moist) {
            double q = 0.5 * (prim_(i, j, k, q1) + prim_(i - 1, j, k, q1) + prim_(i, j, k, q2) + prim_(i - 1, j, k, q2));
            gpx /= (1.0 + q);
        }
        double pi_c = 0.5 * (pi_st(i - 1, j, k) + pi_st(i, j, k));
        double fast_rhs_rho_u = -Gamma * R_d * pi_c * gpx;
        nu(i, j, k) = du(i, j, k) + dtau * fast_rhs_rho_u + dtau * slow_rhs.xmom(i, j, k);
        if (k == nz - 1) nu(i, j, k + 1) = nu(i, j, k);
        ax_m(i, j, k) += facinv * nu(i, j, k);
        data.xmom(i, j, k) = stage.xmom(i, j, k) + nu(i, j, k);
    });
    loop(Box(0, 0, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) {
        double met_h_eta = met::h_eta_Jface(z, i, j, k, dyi);
        double met_h_zeta = met::h_zeta_Jface(z, i, j, k, dzi);
        double gp_eta = (extrap(i, j, k) - extrap(i, j - 1, k)) * dyi;
        double gp_zeta_on_jface = (k == 0)
            ? 0.5 * dzi * (extrap(i, j, k + 1) + extrap(i, j - 1, k + 1) - extrap(i, j, k) - extrap(i, j - 1, k))
            : 0.25 * dzi * (extrap(i, j, k + 1) + extrap(i, j - 1, k + 1) - extrap(i, j, k - 1) - extrap(i, j - 1, k - 1));
        double gpy = gp_eta - (met_h_eta / met_h_zeta) * gp_zeta_on_jface;
        gpy *= 1.0;
        if (moist) {
            double q = 0.5 * (prim_(i, j, k, q1) + prim_(i, j - 1, k, q1) + prim_(i, j, k, q2) + prim_(i, j - 1, k, q2));
            gpy /= (1.0 + q);
        }
        double pi_c = 0.5 * (pi_st(i, j - 1, k) + pi_st(i, j, k));
        double fast_rhs_rho_v = -Gamma * R_d * pi_c * gpy;
        nv(i, j, k) = dv(i, j, k) + dtau * fast_rhs_rho_v + dtau * slow_rhs.ymom(i, j, k);
        if (k == nz - 1) nv(i, j, k + 1) = nv(i, j, k);
        ay_m(i, j, k) += facinv * nv(i, j, k);
        data.ymom(i, j, k) = stage.ymom(i, j, k) + nv(i, j, k);
    });

    // :414-457 horizontal divergence of the area-weighted increments
    Arr pr = prim_;
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double h_zeta_cc_xface_hi = 0.5 * dzi * (z(i + 1, j, k + 1) + z(i + 1, j + 1, k + 1) - z(i + 1, j, k) - z(i + 1, j + 1, k));
        double h_zeta_cc_xface_lo = 0.5 * dzi * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k) - z(i, j + 1, k));
        double h_zeta_cc_yface_hi = 0.5 * dzi * (z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1) - z(i, j + 1, k) - z(i + 1, j + 1, k));
        double h_zeta_cc_yface_lo = 0.5 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k) - z(i + 1, j, k));
        double xflux_lo = nu(i, j, k) * h_zeta_cc_xface_lo / 1.0;
        double xflux_hi = nu(i + 1, j, k) * h_zeta_cc_xface_hi / 1.0;
        double yflux_lo = nv(i, j, k) * h_zeta_cc_yface_lo / 1.0;
        double yflux_hi = nv(i, j + 1, k) * h_zeta_cc_yface_hi / 1.0;
        double mfsq = 1.0 * 1.0;
        trhs(i, j, k, 0) = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq;
        trhs(i, j, k, 1) = ((xflux_hi * (pr(i, j, k, 0) + pr(i + 1, j, k, 0)) -
                             xflux_lo * (pr(i, j, k, 0) + pr(i - 1, j, k, 0))) * dxi * mfsq +
                            (yflux_hi * (pr(i, j, k, 0) + pr(i, j + 1, k, 0)) -
                             yflux_lo * (pr(i, j, k, 0) + pr(i, j - 1, k, 0))) * dyi * mfsq) * 0.5;
    });

    // :460-484 Omega of the old increments on the z-faces of the valid box
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz), [&](int i, int j, int k) {
        if (k == 0) om(i, j, k) = 0.;
        else if (k == nz) om(i, j, k) = dw(i, j, k);
        else om(i, j, k) = met::omega_from_w(i, j, k, dw(i, j, k), du, dv, z, dxInv);
    });

    // :499-552 tridiagonal RHS on the interior faces
    const double halfg = std::abs(0.5 * (-gravity));
    Arr sr = slow_rhs.cons;
    loop(Box(0, 0, 1, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double detJ_on_kface = 0.5 * (J(i, j, k) + J(i, j, k - 1));
        double coeff_P = fc(i, j, k, 3);
        double coeff_Q = fc(i, j, k, 4);
        if (moist) {
            double q = 0.5 * (pr(i, j, k, q1) + pr(i, j, k - 1, q1) + pr(i, j, k, q2) + pr(i, j, k - 1, q2));
            coeff_P /= (1.0 + q);
            coeff_Q /= (1.0 + q);
        }
        double theta_t_lo = 0.5 * (pr(i, j, k - 2, 0) + pr(i, j, k - 1, 0));
        double theta_t_mid = 0.5 * (pr(i, j, k - 1, 0) + pr(i, j, k, 0));
        double theta_t_hi = 0.5 * (pr(i, j, k, 0) + pr(i, j, k + 1, 0));
        double R0_tmp = (-halfg * drho(i, j, k) + coeff_P * drt(i, j, k)) * J(i, j, k)
                      + (-halfg * drho(i, j, k - 1) + coeff_Q * drt(i, j, k - 1)) * J(i, j, k - 1);
        double R1_tmp = -halfg * (sr(i, j, k, 0) * J(i, j, k) + sr(i, j, k - 1, 0) * J(i, j, k - 1))
                      + (coeff_P * sr(i, j, k, 1) * J(i, j, k) + coeff_Q * sr(i, j, k - 1, 1) * J(i, j, k - 1));
        double Omega_kp1 = om(i, j, k + 1);
        double Omega_k = om(i, j, k);
        double Omega_km1 = om(i, j, k - 1);
        R1_tmp += (halfg) * (beta_1 * dzi * (Omega_kp1 - Omega_km1) + trhs(i, j, k, 0) + trhs(i, j, k - 1, 0));
        R1_tmp += -(coeff_P * (beta_1 * dzi * (Omega_kp1 * theta_t_hi - Omega_k * theta_t_mid) + trhs(i, j, k, 1)) +
                    coeff_Q * (beta_1 * dzi * (Omega_k * theta_t_mid - Omega_km1 * theta_t_lo) + trhs(i, j, k - 1, 1)));
        RHS(i, j, k) = detJ_on_kface * dw(i, j, k) +
                       dtau * (detJ_on_kface * slow_rhs.zmom(i, j, k) + R0_tmp + dtau * beta_2 * R1_tmp);
        RHS(i, j, k) += detJ_on_kface * met::omega_from_w(i, j, k, 0., nu, nv, z, dxInv);
    });

    // :561-581 (GPU branch) the column solve
    Arr cz = data.zmom;
    loop(Box(0, 0, 0, nx - 1, ny - 1, 0), [&](int i, int j, int) {
        RHS(i, j, 0) = dtau * slow_rhs.zmom(i, j, 0);
        RHS(i, j, nz) = dtau * slow_rhs.zmom(i, j, nz);
        soln(i, j, 0) = RHS(i, j, 0) * fc(i, j, 0, 1);
        for (int k = 1; k <= nz; k++)
            soln(i, j, k) = (RHS(i, j, k) - fc(i, j, k, 0) * soln(i, j, k - 1)) * fc(i, j, k, 1);
        cz(i, j, 0) = stage.zmom(i, j, 0) + soln(i, j, 0);
        cz(i, j, nz) = stage.zmom(i, j, nz) + soln(i, j, nz);
        for (int k = nz - 1; k >= 0; k--) soln(i, j, k) -= (fc(i, j, k, 2) * fc(i, j, k, 1)) * soln(i, j, k + 1);
    });
    // :620-633: ERF then copies the stage value into every face and adds the
    // increment, converted from Omega, on the interior faces only, so faces
    // 0 and nz end at the stage value
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz), [&](int i, int j, int k) { cz(i, j, k) = stage.zmom(i, j, k); });
    loop(Box(0, 0, 1, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double wpp = met::w_from_omega(i, j, k, soln(i, j, k), nu, nv, z, dxInv);
        cz(i, j, k) += wpp;
    });

    // :640-669 vertical fluxes, averaged rho*w, rho and rho*theta
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double zflux_lo = beta_2 * soln(i, j, k) + beta_1 * om(i, j, k);
        double zflux_hi = beta_2 * soln(i, j, k + 1) + beta_1 * om(i, j, k + 1);
        az_m(i, j, k) += facinv * zflux_lo / (1.0 * 1.0);
        if (k == nz - 1) az_m(i, j, k + 1) += facinv * zflux_hi / (1.0 * 1.0);
        double fast_rhs_rho = -(trhs(i, j, k, 0) + (zflux_hi - zflux_lo) * dzi) / J(i, j, k);
        cc(i, j, k, 0) += dtau * (sr(i, j, k, 0) + fast_rhs_rho);
        double fast_rhs_rhotheta = -(trhs(i, j, k, 1) + 0.5 * (zflux_hi * (pr(i, j, k) + pr(i, j, k + 1)) -
                                                              zflux_lo * (pr(i, j, k) + pr(i, j, k - 1))) * dzi) / J(i, j, k);
        cc(i, j, k, 1) += dtau * (sr(i, j, k, 1) + fast_rhs_rhotheta);
    });
}

}  // namespace orc
