// slow.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  Slow right-hand side of one
// RK stage on a flat, dry grid: momentum sources, erf_slow_rhs_pre and
// erf_slow_rhs_post.
#include "oracle.hpp"
#include <cstdio>

namespace orc {

// make_mom_sources (Source/SourceTerms/ERF_make_mom_sources.cpp:40-450):
// buoyancy (make_buoyancy, Source/SourceTerms/ERF_make_buoyancy.cpp:32-380:
// type 1 dry :95-107 and moist :178-209, moist types 2-4 :211-376 on plane
// averages), Coriolis, Rayleigh damping and geostrophic forcing.
void Level::make_mom_sources(StateA data, Arr prim_) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    Arr xs = A(xmom_src), ys = A(ymom_src), zs = A(zmom_src);
    std::fill(xmom_src.v.begin(), xmom_src.v.end(), 0.0);
    std::fill(ymom_src.v.begin(), ymom_src.v.end(), 0.0);
    std::fill(zmom_src.v.begin(), zmom_src.v.end(), 0.0);
    Arr c = data.cons, r0a = A(r0);
    const double grav_z = -gravity;
    const Box faces_in(0, 0, 1, nx - 1, ny - 1, nz - 1);
    if (!moist) {
        loop(faces_in, [&](int i, int j, int k) {
            zs(i, j, k) = grav_z * 0.5 * ((c(i, j, k) - r0a(i, j, k)) + (c(i, j, k - 1) - r0a(i, j, k - 1)));
        });
    } else if (cfg.buoyancy_type == 1) {
        loop(faces_in, [&](int i, int j, int k) {
            double rhop_hi = c(i, j, k, ERFX_RHO) + c(i, j, k, ERFX_RHOQ1) + c(i, j, k, ERFX_RHOQ2) - r0a(i, j, k);
            double rhop_lo = c(i, j, k - 1, ERFX_RHO) + c(i, j, k - 1, ERFX_RHOQ1) + c(i, j, k - 1, ERFX_RHOQ2) -
                             r0a(i, j, k - 1);
            for (int q_offset = 2; q_offset < n_qstate; ++q_offset) {
                rhop_hi += c(i, j, k, ERFX_RHOQ1 + q_offset);
                rhop_lo += c(i, j, k - 1, ERFX_RHOQ1 + q_offset);
            }
            zs(i, j, k) = grav_z * 0.5 * (rhop_hi + rhop_lo);
        });
    } else {
        std::vector<double> avg;
        plane_averages(data, prim_, avg);
        const double* rho_d = &avg[0];
        const double* theta_d = &avg[(size_t)nz];
        const double* qv_d = &avg[2 * (size_t)nz];
        const double* qc_d = &avg[3 * (size_t)nz];
        const double* qp_d = &avg[4 * (size_t)nz];
        const int btype = cfg.buoyancy_type;
        const int nq = n_qstate;
        // the departure of one cell from the level's plane averages
        auto qdev = [&](int i, int j, int k) {
            double temp1d = getTgivenRandRTh(rho_d[k], rho_d[k] * theta_d[k], qv_d[k]);
            double temp3d = getTgivenRandRTh(c(i, j, k, ERFX_RHO), c(i, j, k, ERFX_RHOTHETA),
                                             c(i, j, k, ERFX_RHOQ1) / c(i, j, k, ERFX_RHO));
            double qv = (nq >= 1) ? prim_(i, j, k, ERFX_RHOQ1 - 1) : 0.0;
            double qc = (nq >= 2) ? prim_(i, j, k, ERFX_RHOQ2 - 1) : 0.0;
            double qp = (nq >= 3) ? prim_(i, j, k, ERFX_RHOQ3 - 1) : 0.0;
            if (btype == 2)
                return 0.61 * (qv - qv_d[k]) - (qc - qc_d[k] + qp - qp_d[k]) +
                       (temp3d - temp1d) / temp1d * (1.0 + 0.61 * qv_d[k] - qc_d[k] - qp_d[k]);
            if (btype == 4)
                return 0.61 * (qv - qv_d[k]) - (qc - qc_d[k] + qp - qp_d[k]) +
                       (c(i, j, k, ERFX_RHOTHETA) / c(i, j, k, ERFX_RHO) - theta_d[k]) / theta_d[k];
            return 0.61 * qv - (qc + qp) + (temp3d - temp1d) / temp1d;     // type 3
        };
        loop(faces_in, [&](int i, int j, int k) {
            double qplus = qdev(i, j, k);
            double qminus = qdev(i, j, k - 1);
            double qavg = 0.5 * (qplus + qminus);
            double r0avg = 0.5 * (rho_d[k] + rho_d[k - 1]);
            zs(i, j, k) = -qavg * r0avg * grav_z;
        });
    }
    Box tbx(0, 0, 0, nx, ny - 1, nz - 1), tby(0, 0, 0, nx - 1, ny, nz - 1), tbz(0, 0, 0, nx - 1, ny - 1, nz - 1);
    Arr ru = data.xmom, rv = data.ymom, rw = data.zmom;
    if (cfg.use_coriolis) {
        const double cf = cfg.coriolis_factor, sinphi = cfg.sinphi, cosphi = cfg.cosphi;
        loop(tbx, [&](int i, int j, int k) {
            double rho_v_loc = 0.25 * (rv(i, j + 1, k) + rv(i, j, k) + rv(i - 1, j + 1, k) + rv(i - 1, j, k));
            double rho_w_loc = 0.25 * (rw(i, j, k + 1) + rw(i, j, k) + rw(i, j - 1, k + 1) + rw(i, j - 1, k));
            xs(i, j, k) += cf * (rho_v_loc * sinphi - rho_w_loc * cosphi);
        });
        loop(tby, [&](int i, int j, int k) {
            double rho_u_loc = 0.25 * (ru(i + 1, j, k) + ru(i, j, k) + ru(i + 1, j - 1, k) + ru(i, j - 1, k));
            ys(i, j, k) += -cf * rho_u_loc * sinphi;
        });
        loop(tbz, [&](int i, int j, int k) {
            double rho_u_loc = 0.25 * (ru(i + 1, j, k) + ru(i, j, k) + ru(i + 1, j, k - 1) + ru(i, j, k - 1));
            zs(i, j, k) += cf * rho_u_loc * cosphi;
        });
    }
    // Rayleigh damping (ERF_make_mom_sources.cpp:247-296)
    if ((cfg.rayleigh_damp_U || cfg.rayleigh_damp_V || cfg.rayleigh_damp_W) && (int)ray_wbar.size() < nz + 1)
        throw std::runtime_error("Rayleigh damping profiles not set");
    {
        const double zlo = cfg.prob_lo[2], dz = cfg.dz;
        const double ztop = terrain ? A(z_nd)(0, 0, nz) : flat_ztop(cfg);   // ERF_init_rayleigh.cpp:17 (ProbHi(2))
        const double zdamp = cfg.rayleigh_zdamp, dampcoef = cfg.rayleigh_dampcoef;
        if (cfg.rayleigh_damp_U) {
            loop(tbx, [&](int i, int j, int k) {
                double zcc = terrain ? A(z_cc)(i, j, k) : zlo + (k + 0.5) * dz;
                double zfrac = 1 - (ztop - zcc) / zdamp;
                if (zfrac > 0) {
                    double rho_on_u_face = 0.5 * (c(i, j, k) + c(i - 1, j, k));
                    double uu = ru(i, j, k) / rho_on_u_face;
                    double sinefac = std::sin(PIoTwo * zfrac);
                    xs(i, j, k) -= dampcoef * sinefac * sinefac * (uu - ray_ubar[k]) * rho_on_u_face;
                }
            });
        }
        if (cfg.rayleigh_damp_V) {
            loop(tby, [&](int i, int j, int k) {
                double zcc = terrain ? A(z_cc)(i, j, k) : zlo + (k + 0.5) * dz;
                double zfrac = 1 - (ztop - zcc) / zdamp;
                if (zfrac > 0) {
                    double rho_on_v_face = 0.5 * (c(i, j, k) + c(i, j - 1, k));
                    double vv = rv(i, j, k) / rho_on_v_face;
                    double sinefac = std::sin(PIoTwo * zfrac);
                    ys(i, j, k) -= dampcoef * sinefac * sinefac * (vv - ray_vbar[k]) * rho_on_v_face;
                }
            });
        }
        if (cfg.rayleigh_damp_W) {
            loop(tbz, [&](int i, int j, int k) {
                double zstag = terrain ? A(z_nd)(i, j, k) : zlo + k * dz;
                double zfrac = 1 - (ztop - zstag) / zdamp;
                if (zfrac > 0) {
                    double rho_on_w_face = 0.5 * (c(i, j, k) + c(i, j, k - 1));
                    double ww = rw(i, j, k) / rho_on_w_face;
                    double sinefac = std::sin(PIoTwo * zfrac);
                    zs(i, j, k) -= dampcoef * sinefac * sinefac * (ww - ray_wbar[k]) * rho_on_w_face;
                }
            });
        }
    }
    const double* g = cfg.abl_geo_forcing;
    loop(tbx, [&](int i, int j, int k) {
        double rho_on_u_face = 0.5 * (c(i, j, k) + c(i - 1, j, k));
        xs(i, j, k) += rho_on_u_face * g[0];
    });
    loop(tby, [&](int i, int j, int k) {
        double rho_on_v_face = 0.5 * (c(i, j, k) + c(i, j - 1, k));
        ys(i, j, k) += rho_on_v_face * g[1];
    });
    loop(tbz, [&](int i, int j, int k) {
        double rho_on_w_face = 0.5 * (c(i, j, k) + c(i, j, k - 1));
        zs(i, j, k) += rho_on_w_face * g[2];
    });
}

// make_sources (Source/SourceTerms/ERF_make_sources.cpp:32-418) for the dry
// flat step: cc_src is zero except the Rayleigh damping of rho*theta (:203-225).
void Level::make_sources(StateA data, Arr prim_) {
    std::fill(cc_src.v.begin(), cc_src.v.end(), 0.0);
    if (!cfg.rayleigh_damp_T) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    if ((int)ray_thetabar.size() < nz + 1) throw std::runtime_error("Rayleigh damping profiles not set");
    Arr cs = A(cc_src), c = data.cons;
    const double zlo = cfg.prob_lo[2], dz = cfg.dz;
    const double ztop = terrain ? A(z_nd)(0, 0, nz) : flat_ztop(cfg);
    const double zdamp = cfg.rayleigh_zdamp, dampcoef = cfg.rayleigh_dampcoef;
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double zcc = terrain ? A(z_cc)(i, j, k) : zlo + (k + 0.5) * dz;
        double zfrac = 1 - (ztop - zcc) / zdamp;
        if (zfrac > 0) {
            double theta = prim_(i, j, k, 0);
            double sinefac = std::sin(PIoTwo * zfrac);
            cs(i, j, k, 1) -= dampcoef * sinefac * sinefac * (theta - ray_thetabar[k]) * c(i, j, k, 0);
        }
    });
}

// erf_slow_rhs_pre (Source/TimeIntegration/ERF_slow_rhs_pre.cpp:67-774), flat,
// dry, no terrain: p', Omega, AdvectionSrcForRho (ERF_AdvectionSrcForState.cpp:29-93),
// rho*theta advection and diffusion, momentum advection and diffusion,
// pressure gradient plus sources.
void Level::slow_rhs_pre(int nrk, double dt, StateA rhs, StateA data, Arr prim_, StateA scratch,
                         Arr u, Arr v, Arr w) {
    (void)nrk;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const bool use_diff = cfg.molec_diff_type != ERFX_MOLEC_NONE || cfg.les_type != ERFX_LES_NONE;
    if (use_diff && terrain && cfg.les_type != ERFX_LES_NONE)
        throw std::runtime_error("an LES model on a terrain-following grid is not restated");
    if (use_diff) {
        if (terrain) make_tau_terms_T(data, u, v, w);
        else make_tau_terms(data, u, v, w);
    }
    Arr rwr = rhs.zmom;
    // S_rhs[zmom].setVal(0.0) (:232)
    {
        double* p = rwr.p;
        std::fill(p, p + L.nelem, 0.0);
    }
    Arr c = data.cons;
    Box bx(0, 0, 0, nx - 1, ny - 1, nz - 1);
    // p' on the box grown by one, clipped at k >= 0 (:337-358)
    Field fpp; fpp.alloc(L, 1);
    Arr pp = fpp.a(L), p0a = A(p0);
    int bad_theta = 0;
    loop(Box(-1, -1, 0, nx, ny, nz), [&](int i, int j, int k) {
        if (c(i, j, k, 1) <= 0.) {
#pragma omp critical
            { if (!bad_theta) { bad_theta = 1; std::fprintf(stderr, "BAD THETA AT %d %d %d %e\n", i, j, k, c(i, j, k, 1)); } }
        }
        double qv_for_p = moist ? c(i, j, k, ERFX_RHOQ1) / c(i, j, k, ERFX_RHO) : 0.0;
        pp(i, j, k) = getPgivenRTh(c(i, j, k, 1), qv_for_p) - p0a(i, j, k);
    });
    if (bad_theta) throw std::runtime_error("Bad theta in slow_rhs_pre");
    // Omega on z-faces grown by one (:367-414): rho*w on a flat grid; with
    // terrain 0 at and below the bottom face, rho*w above the top face and
    // OmegaFromW in between (the top face included)
    Arr om = A(omega);
    if (!terrain) {
        loop(Box(-1, -1, -1, nx, ny, nz + 1), [&](int i, int j, int k) { om(i, j, k) = data.zmom(i, j, k); });
    } else {
        Arr zn = A(z_nd);
        loop(Box(-1, -1, -1, nx, ny, nz + 1), [&](int i, int j, int k) {
            if (k <= 0) om(i, j, k) = 0.;
            else if (k == nz + 1) om(i, j, k) = data.zmom(i, j, k);
            else om(i, j, k) = met::omega_from_w(i, j, k, data.zmom(i, j, k), data.xmom, data.ymom, zn, dxInv);
        });
    }
    Arr Ax = A(ax), Ay = A(ay), Az = A(az), Jc = A(detJ);
    // AdvectionSrcForRho
    Arr ax = scratch.xmom, ay = scratch.ymom, az = scratch.zmom;
    const double dxi = dxInv[0], dyi = dxInv[1], dzi = dxInv[2];
    Field ffx, ffy, ffz; ffx.alloc(L, 1); ffy.alloc(L, 1); ffz.alloc(L, 1);
    Arr fx = ffx.a(L), fy = ffy.a(L), fz = ffz.a(L);
    // AdvectionSrcForRho (ERF_AdvectionSrcForState.cpp:29-93): area-weighted
    // fluxes (areas and detJ are 1 on a flat grid)
    loop(Box(0, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) {
        fx(i, j, k) = (terrain ? Ax(i, j, k) : 1.0) * data.xmom(i, j, k) / 1.0;
        ax(i, j, k) = fx(i, j, k);
    });
    loop(Box(0, 0, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) {
        fy(i, j, k) = (terrain ? Ay(i, j, k) : 1.0) * data.ymom(i, j, k) / 1.0;
        ay(i, j, k) = fy(i, j, k);
    });
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz), [&](int i, int j, int k) {
        double mfsq = 1.0 * 1.0;
        fz(i, j, k) = (terrain ? Az(i, j, k) : 1.0) * om(i, j, k) / mfsq;
        az(i, j, k) = fz(i, j, k);
    });
    Arr cr = rhs.cons;
    loop(bx, [&](int i, int j, int k) {
        double mfsq = 1.0 * 1.0;
        const double dJ = terrain ? Jc(i, j, k) : 1.0;
        if (dJ > 0.)
            cr(i, j, k, 0) = -mfsq / dJ * ((fx(i + 1, j, k) - fx(i, j, k)) * dxi +
                                          (fy(i, j + 1, k) - fy(i, j, k)) * dyi +
                                          (fz(i, j, k + 1) - fz(i, j, k)) * dzi);
        else
            cr(i, j, k, 0) = 0.;
    });
    advection_src_scalars(bx, 1, 1, ax, ay, az, prim_, cr, cfg.dycore_horiz_adv, cfg.dycore_vert_adv,
                          cfg.dycore_horiz_upw_frac, cfg.dycore_vert_upw_frac);
    if (use_diff) {
        if (terrain) diffusion_src_state_T(bx, 1, 1, c, prim_, cr);
        else diffusion_src_state(bx, 1, 1, c, prim_, cr);
    }
    Arr cs = A(cc_src);
    loop(bx, [&](int i, int j, int k) {
        cr(i, j, k, 0) += cs(i, j, k, 0);
        cr(i, j, k, 1) += cs(i, j, k, 1);
    });
    // momentum
    Box tbx(0, 0, 0, nx, ny - 1, nz - 1), tby(0, 0, 0, nx - 1, ny, nz - 1), tbz(0, 0, 1, nx - 1, ny - 1, nz - 1);
    Arr rur = rhs.xmom, rvr = rhs.ymom;
    if (terrain) advection_src_mom_T(tbx, tby, tbz, rur, rvr, rwr, u, v, w, data.xmom, data.ymom, om);
    else advection_src_mom(tbx, tby, tbz, rur, rvr, rwr, u, v, w, data.xmom, data.ymom, om);
    if (use_diff && terrain) {   // DiffusionSrcForMom_T (ERF_DiffusionSrcForMom_T.cpp:30-79)
        Arr t11 = A(tau11), t22 = A(tau22), t33 = A(tau33), t12 = A(tau12), t13 = A(tau13), t23 = A(tau23);
        Arr t21 = A(tau21), t31 = A(tau31), t32 = A(tau32), Jc = A(detJ);
        loop(tbx, [&](int i, int j, int k) {
            double mf = 1.0;
            double diffContrib = ((t11(i, j, k) - t11(i - 1, j, k)) * dxi * mf
                                + (t12(i, j + 1, k) - t12(i, j, k)) * dyi * mf
                                + (t13(i, j, k + 1) - t13(i, j, k)) * dzi);
            diffContrib /= 0.5 * (Jc(i, j, k) + Jc(i - 1, j, k));
            rur(i, j, k) -= diffContrib;
        });
        loop(tby, [&](int i, int j, int k) {
            double mf = 1.0;
            double diffContrib = ((t21(i + 1, j, k) - t21(i, j, k)) * dxi * mf
                                + (t22(i, j, k) - t22(i, j - 1, k)) * dyi * mf
                                + (t23(i, j, k + 1) - t23(i, j, k)) * dzi);
            diffContrib /= 0.5 * (Jc(i, j, k) + Jc(i, j - 1, k));
            rvr(i, j, k) -= diffContrib;
        });
        loop(tbz, [&](int i, int j, int k) {
            double mf = 1.0;
            double diffContrib = ((t31(i + 1, j, k) - t31(i, j, k)) * dxi * mf
                                + (t32(i, j + 1, k) - t32(i, j, k)) * dyi * mf
                                + (t33(i, j, k) - t33(i, j, k - 1)) * dzi);
            diffContrib /= 0.5 * (Jc(i, j, k) + Jc(i, j, k - 1));
            rwr(i, j, k) -= diffContrib;
        });
    } else if (use_diff) {
        Arr t11 = A(tau11), t22 = A(tau22), t33 = A(tau33), t12 = A(tau12), t13 = A(tau13), t23 = A(tau23);
        loop(tbx, [&](int i, int j, int k) {
            double mf = 1.0;
            rur(i, j, k) -= ((t11(i, j, k) - t11(i - 1, j, k)) * dxi * mf
                           + (t12(i, j + 1, k) - t12(i, j, k)) * dyi * mf
                           + (t13(i, j, k + 1) - t13(i, j, k)) * dzi);
        });
        loop(tby, [&](int i, int j, int k) {
            double mf = 1.0;
            rvr(i, j, k) -= ((t12(i + 1, j, k) - t12(i, j, k)) * dxi * mf
                           + (t22(i, j, k) - t22(i, j - 1, k)) * dyi * mf
                           + (t23(i, j, k + 1) - t23(i, j, k)) * dzi);
        });
        loop(tbz, [&](int i, int j, int k) {
            double mf = 1.0;
            rwr(i, j, k) -= ((t13(i + 1, j, k) - t13(i, j, k)) * dxi * mf
                           + (t23(i, j + 1, k) - t23(i, j, k)) * dyi * mf
                           + (t33(i, j, k) - t33(i, j, k - 1)) * dzi);
        });
    }
    Arr xs = A(xmom_src), ys = A(ymom_src), zs = A(zmom_src);
    const double apg0 = cfg.abl_pressure_grad[0], apg1 = cfg.abl_pressure_grad[1], apg2 = cfg.abl_pressure_grad[2];
    // q = qv + qc averaged to the face (:622-626, :671-675, :732-736)
    const int q1 = ERFX_RHOQ1 - 1, q2 = ERFX_RHOQ2 - 1;
    Arr cp = prim_;
    Arr zn = A(z_nd);
    loop(tbx, [&](int i, int j, int k) {
        double gp_xi = dxi * (pp(i, j, k) - pp(i - 1, j, k));
        double gpx = gp_xi;
        if (terrain) {   // :600-620
            double met_h_xi = met::h_xi_Iface(zn, i, j, k, dxi);
            double met_h_zeta = met::h_zeta_Iface(zn, i, j, k, dzi);
            double gp_zeta_on_iface;
            if (k == 0)
                gp_zeta_on_iface = 0.5 * dzi * (pp(i - 1, j, k + 1) + pp(i, j, k + 1) - pp(i - 1, j, k) - pp(i, j, k));
            else if (k == nz - 1)
                gp_zeta_on_iface = 0.5 * dzi * (pp(i - 1, j, k) + pp(i, j, k) - pp(i - 1, j, k - 1) - pp(i, j, k - 1));
            else
                gp_zeta_on_iface = 0.25 * dzi * (pp(i - 1, j, k + 1) + pp(i, j, k + 1) - pp(i - 1, j, k - 1) - pp(i, j, k - 1));
            gpx -= (met_h_xi / met_h_zeta) * gp_zeta_on_iface;
        }
        gpx *= 1.0;
        double q = 0.0;
        if (moist) q = 0.5 * (cp(i, j, k, q1) + cp(i - 1, j, k, q1) + cp(i, j, k, q2) + cp(i - 1, j, k, q2));
        rur(i, j, k) += (-gpx - apg0) / (1.0 + q) + xs(i, j, k);
    });
    loop(tby, [&](int i, int j, int k) {
        double gp_eta = dyi * (pp(i, j, k) - pp(i, j - 1, k));
        double gpy = gp_eta;
        if (terrain) {   // :650-669
            double met_h_eta = met::h_eta_Jface(zn, i, j, k, dyi);
            double met_h_zeta = met::h_zeta_Jface(zn, i, j, k, dzi);
            double gp_zeta_on_jface;
            if (k == 0)
                gp_zeta_on_jface = 0.5 * dzi * (pp(i, j, k + 1) + pp(i, j - 1, k + 1) - pp(i, j, k) - pp(i, j - 1, k));
            else if (k == nz - 1)
                gp_zeta_on_jface = 0.5 * dzi * (pp(i, j, k) + pp(i, j - 1, k) - pp(i, j, k - 1) - pp(i, j - 1, k - 1));
            else
                gp_zeta_on_jface = 0.25 * dzi * (pp(i, j, k + 1) + pp(i, j - 1, k + 1) - pp(i, j, k - 1) - pp(i, j - 1, k - 1));
            gpy -= (met_h_eta / met_h_zeta) * gp_zeta_on_jface;
        }
        gpy *= 1.0;
        double q = 0.0;
        if (moist) q = 0.5 * (cp(i, j, k, q1) + cp(i, j - 1, k, q1) + cp(i, j, k, q2) + cp(i, j - 1, k, q2));
        rvr(i, j, k) += (-gpy - apg1) / (1.0 + q) + ys(i, j, k);
    });
    // zero the rhs on Dirichlet lateral faces (:697-726)
    if (bck[0][0] == EXT_DIR) loop(Box(0, 0, 0, 0, ny - 1, nz - 1), [&](int i, int j, int k) { rur(i, j, k) = 0.; });
    if (bck[0][3] == EXT_DIR) loop(Box(nx, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) { rur(i, j, k) = 0.; });
    if (bck[1][1] == EXT_DIR) loop(Box(0, 0, 0, nx - 1, 0, nz - 1), [&](int i, int j, int k) { rvr(i, j, k) = 0.; });
    if (bck[1][4] == EXT_DIR) loop(Box(0, ny, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) { rvr(i, j, k) = 0.; });
    loop(tbz, [&](int i, int j, int k) {
        double met_h_zeta = terrain ? met::h_zeta_Kface(zn, i, j, k, dzi) : 1;
        double gpz = dzi * (pp(i, j, k) - pp(i, j, k - 1)) / met_h_zeta;
        double q = 0.0;
        if (moist) q = 0.5 * (cp(i, j, k, q1) + cp(i, j, k - 1, q1) + cp(i, j, k, q2) + cp(i, j, k - 1, q2));
        rwr(i, j, k) += (zs(i, j, k) - gpz - apg2) / (1.0 + q);
    });
    (void)dt;
}

// erf_slow_rhs_post (Source/TimeIntegration/ERF_slow_rhs_post.cpp:49-559):
// the slow variables are RhoScalar and, with moisture, the n_qstate species
// as one block with the moist advection types (:168-175, :364-434); the
// species are clipped at 0 after the update (:485-497).
void Level::slow_rhs_post(int nrk, double dt, StateA rhs, StateA old, StateA snew, StateA data,
                          Arr prim_, StateA scratch, Arr u, Arr v) {
    (void)nrk; (void)u; (void)v;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const bool use_diff = cfg.molec_diff_type != ERFX_MOLEC_NONE || cfg.les_type != ERFX_LES_NONE;
    Box tbx(0, 0, 0, nx - 1, ny - 1, nz - 1);
    Arr cur = data.cons, ncons_ = snew.cons, oc = old.cons, cr = rhs.cons;
    int nsv = ncons - 2;
    loop(tbx, [&](int i, int j, int k) {
        for (int nn = 0; nn < nsv; ++nn) cur(i, j, k, 2 + nn) = ncons_(i, j, k, 2 + nn);
    });
    const int ivar = ERFX_RHOSCALAR;
    advection_src_scalars(tbx, ivar, 1, scratch.xmom, scratch.ymom, scratch.zmom, prim_, cr,
                          cfg.dryscal_horiz_adv, cfg.dryscal_vert_adv, cfg.dryscal_horiz_upw_frac,
                          cfg.dryscal_vert_upw_frac);
    if (use_diff) {
        if (terrain) diffusion_src_state_T(tbx, ivar, 1, ncons_, prim_, cr);
        else diffusion_src_state(tbx, ivar, 1, ncons_, prim_, cr);
    }
    if (moist) {
        advection_src_scalars(tbx, ERFX_RHOQ1, n_qstate, scratch.xmom, scratch.ymom, scratch.zmom, prim_, cr,
                              cfg.moistscal_horiz_adv, cfg.moistscal_vert_adv, cfg.moistscal_horiz_upw_frac,
                              cfg.moistscal_vert_upw_frac);
        if (use_diff) {
            if (terrain) diffusion_src_state_T(tbx, ERFX_RHOQ1, n_qstate, ncons_, prim_, cr);
            else diffusion_src_state(tbx, ERFX_RHOQ1, n_qstate, ncons_, prim_, cr);
        }
    }
    Arr cs = A(cc_src);
    loop(tbx, [&](int i, int j, int k) {
        cr(i, j, k, ivar) += cs(i, j, k, ivar);
        cur(i, j, k, ivar) = oc(i, j, k, ivar) + dt * cr(i, j, k, ivar);
    });
    if (moist) {
        loop(tbx, [&](int i, int j, int k) {
            for (int n = ERFX_RHOQ1; n < ncons; ++n) {
                cr(i, j, k, n) += cs(i, j, k, n);
                cur(i, j, k, n) = oc(i, j, k, n) + dt * cr(i, j, k, n);
                cur(i, j, k, n) = std::max(cur(i, j, k, n), 0.0);
            }
        });
    }
    loop(tbx, [&](int i, int j, int k) {
        for (int n = 0; n < ncons; ++n) ncons_(i, j, k, n) = cur(i, j, k, n);
    });
    loop(Box(0, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) { snew.xmom(i, j, k) = data.xmom(i, j, k); });
    loop(Box(0, 0, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) { snew.ymom(i, j, k) = data.ymom(i, j, k); });
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz), [&](int i, int j, int k) { snew.zmom(i, j, k) = data.zmom(i, j, k); });
}

}  // namespace orc
