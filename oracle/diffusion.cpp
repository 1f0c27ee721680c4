// diffusion.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  Strain, stress,
// Smagorinsky eddy viscosity and diffusive fluxes on a flat grid.
#include "oracle.hpp"

namespace orc {

// ComputeStrain_N (Source/Diffusion/ERF_ComputeStrain_N.cpp:28-229), unit map
// factors.  One-sided 8/3 stencils where a velocity is Dirichlet on a face the
// edge box touches.
void Level::compute_strain_N(const Box& bxcc, Box tbxxy, Box tbxxz, Box tbxyz, Arr u, Arr v, Arr w,
                             Arr t11, Arr t22, Arr t33, Arr t12, Arr t13, Arr t23) {
    const double* dI = dxInv;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    auto isdir = [&](int var, int ori) { return bck[var][ori] == EXT_DIR; };
    bool xl_v_dir = isdir(1, 0) && tbxxy.lo[0] == 0;
    bool xh_v_dir = isdir(1, 3) && tbxxy.hi[0] == nx;
    bool xl_w_dir = isdir(2, 0) && tbxxz.lo[0] == 0;
    bool xh_w_dir = isdir(2, 3) && tbxxz.hi[0] == nx;
    bool yl_u_dir = isdir(0, 1) && tbxxy.lo[1] == 0;
    bool yh_u_dir = isdir(0, 4) && tbxxy.hi[1] == ny;
    bool yl_w_dir = isdir(2, 1) && tbxyz.lo[1] == 0;
    bool yh_w_dir = isdir(2, 4) && tbxyz.hi[1] == ny;
    bool zl_u_dir = isdir(0, 2) && tbxxz.lo[2] == 0;
    bool zh_u_dir = isdir(0, 5) && tbxxz.hi[2] == nz;
    bool zl_v_dir = isdir(1, 2) && tbxyz.lo[2] == 0;
    bool zh_v_dir = isdir(1, 5) && tbxyz.hi[2] == nz;
    if (xl_v_dir) {
        Box p = tbxxy; p.hi[0] = p.lo[0]; tbxxy.lo[0] += 1;
        loop(p, [&](int i, int j, int k) {
            t12(i, j, k) = 0.5 * ((u(i, j, k) / 1.0 - u(i, j - 1, k) / 1.0) * dI[1] +
                                  (-(8. / 3.) * v(i - 1, j, k) / 1.0 + 3. * v(i, j, k) / 1.0 -
                                   (1. / 3.) * v(i + 1, j, k) / 1.0) * dI[0]) * 1.0 * 1.0;
        });
    }
    if (xh_v_dir) {
        Box p = tbxxy; p.lo[0] = p.hi[0]; tbxxy.hi[0] -= 1;
        loop(p, [&](int i, int j, int k) {
            t12(i, j, k) = 0.5 * ((u(i, j, k) / 1.0 - u(i, j - 1, k) / 1.0) * dI[1] +
                                  -(-(8. / 3.) * v(i, j, k) / 1.0 + 3. * v(i - 1, j, k) / 1.0 -
                                    (1. / 3.) * v(i - 2, j, k) / 1.0) * dI[0]) * 1.0 * 1.0;
        });
    }
    if (xl_w_dir) {
        Box p = tbxxz; p.hi[0] = p.lo[0]; tbxxz.lo[0] += 1;
        loop(p, [&](int i, int j, int k) {
            t13(i, j, k) = 0.5 * ((u(i, j, k) - u(i, j, k - 1)) * dI[2] +
                                  (-(8. / 3.) * w(i - 1, j, k) + 3. * w(i, j, k) - (1. / 3.) * w(i + 1, j, k)) * dI[0] * 1.0);
        });
    }
    if (xh_w_dir) {
        Box p = tbxxz; p.lo[0] = p.hi[0]; tbxxz.hi[0] -= 1;
        loop(p, [&](int i, int j, int k) {
            t13(i, j, k) = 0.5 * ((u(i, j, k) - u(i, j, k - 1)) * dI[2] +
                                  -(-(8. / 3.) * w(i, j, k) + 3. * w(i - 1, j, k) - (1. / 3.) * w(i - 2, j, k)) * dI[0] * 1.0);
        });
    }
    if (yl_u_dir) {
        Box p = tbxxy; p.hi[1] = p.lo[1]; tbxxy.lo[1] += 1;
        loop(p, [&](int i, int j, int k) {
            t12(i, j, k) = 0.5 * ((-(8. / 3.) * u(i, j - 1, k) / 1.0 + 3. * u(i, j, k) / 1.0 - (1. / 3.) * u(i, j + 1, k) / 1.0) * dI[1] +
                                  (v(i, j, k) / 1.0 - v(i - 1, j, k) / 1.0) * dI[0]) * 1.0 * 1.0;
        });
    }
    if (yh_u_dir) {
        Box p = tbxxy; p.lo[1] = p.hi[1]; tbxxy.hi[1] -= 1;
        loop(p, [&](int i, int j, int k) {
            t12(i, j, k) = 0.5 * (-(-(8. / 3.) * u(i, j, k) / 1.0 + 3. * u(i, j - 1, k) / 1.0 - (1. / 3.) * u(i, j - 2, k) / 1.0) * dI[1] +
                                  (v(i, j, k) / 1.0 - v(i - 1, j, k) / 1.0) * dI[0]) * 1.0 * 1.0;
        });
    }
    if (yl_w_dir) {
        Box p = tbxyz; p.hi[1] = p.lo[1]; tbxyz.lo[1] += 1;
        loop(p, [&](int i, int j, int k) {
            t23(i, j, k) = 0.5 * ((v(i, j, k) - v(i, j, k - 1)) * dI[2] +
                                  (-(8. / 3.) * w(i, j - 1, k) + 3. * w(i, j, k) - (1. / 3.) * w(i, j + 1, k)) * dI[1] * 1.0 * 1.0);
        });
    }
    if (yh_w_dir) {
        Box p = tbxyz; p.lo[1] = p.hi[1]; tbxyz.hi[1] -= 1;
        loop(p, [&](int i, int j, int k) {
            t23(i, j, k) = 0.5 * ((v(i, j, k) - v(i, j, k - 1)) * dI[2] +
                                  -(-(8. / 3.) * w(i, j, k) + 3. * w(i, j - 1, k) - (1. / 3.) * w(i, j - 2, k)) * dI[1] * 1.0 * 1.0);
        });
    }
    if (zl_u_dir) {
        Box p = tbxxz; p.hi[2] = p.lo[2]; tbxxz.lo[2] += 1;
        loop(p, [&](int i, int j, int k) {
            t13(i, j, k) = 0.5 * ((-(8. / 3.) * u(i, j, k - 1) + 3. * u(i, j, k) - (1. / 3.) * u(i, j, k + 1)) * dI[2] +
                                  (w(i, j, k) - w(i - 1, j, k)) * dI[0] * 1.0);
        });
    }
    if (zh_u_dir) {
        Box p = tbxxz; p.lo[2] = p.hi[2]; tbxxz.hi[2] -= 1;
        loop(p, [&](int i, int j, int k) {
            t13(i, j, k) = 0.5 * (-(-(8. / 3.) * u(i, j, k) + 3. * u(i, j, k - 1) - (1. / 3.) * u(i, j, k - 2)) * dI[2] +
                                  (w(i, j, k) - w(i - 1, j, k)) * dI[0] * 1.0);
        });
    }
    if (zl_v_dir) {
        Box p = tbxyz; p.hi[2] = p.lo[2]; tbxyz.lo[2] += 1;
        loop(p, [&](int i, int j, int k) {
            t23(i, j, k) = 0.5 * ((-(8. / 3.) * v(i, j, k - 1) + 3. * v(i, j, k) - (1. / 3.) * v(i, j, k + 1)) * dI[2] +
                                  (w(i, j, k) - w(i, j - 1, k)) * dI[1] * 1.0);
        });
    }
    if (zh_v_dir) {
        Box p = tbxyz; p.lo[2] = p.hi[2]; tbxyz.hi[2] -= 1;
        loop(p, [&](int i, int j, int k) {
            t23(i, j, k) = 0.5 * (-(-(8. / 3.) * v(i, j, k) + 3. * v(i, j, k - 1) - (1. / 3.) * v(i, j, k - 2)) * dI[2] +
                                  (w(i, j, k) - w(i, j - 1, k)) * dI[1] * 1.0);
        });
    }
    loop(bxcc, [&](int i, int j, int k) {
        t11(i, j, k) = (u(i + 1, j, k) / 1.0 - u(i, j, k) / 1.0) * dI[0] * 1.0 * 1.0;
        t22(i, j, k) = (v(i, j + 1, k) / 1.0 - v(i, j, k) / 1.0) * dI[1] * 1.0 * 1.0;
        t33(i, j, k) = (w(i, j, k + 1) - w(i, j, k)) * dI[2];
    });
    loop(tbxxy, [&](int i, int j, int k) {
        t12(i, j, k) = 0.5 * ((u(i, j, k) / 1.0 - u(i, j - 1, k) / 1.0) * dI[1] +
                              (v(i, j, k) / 1.0 - v(i - 1, j, k) / 1.0) * dI[0]) * 1.0 * 1.0;
    });
    loop(tbxxz, [&](int i, int j, int k) {
        t13(i, j, k) = 0.5 * ((u(i, j, k) - u(i, j, k - 1)) * dI[2] + (w(i, j, k) - w(i - 1, j, k)) * dI[0] * 1.0);
    });
    loop(tbxyz, [&](int i, int j, int k) {
        t23(i, j, k) = 0.5 * ((v(i, j, k) - v(i, j, k - 1)) * dI[2] + (w(i, j, k) - w(i, j - 1, k)) * dI[1] * 1.0);
    });
}

// ComputeSmnSmn (Source/Diffusion/ERF_EddyViscosity.H:28-83) without MOST.
static inline double smn_smn(int i, int j, int k, Arr t11, Arr t22, Arr t33, Arr t12, Arr t13, Arr t23) {
    double s11bar = t11(i, j, k), s22bar = t22(i, j, k), s33bar = t33(i, j, k);
    double s12bar = 0.25 * (t12(i, j, k) + t12(i, j + 1, k) + t12(i + 1, j, k) + t12(i + 1, j + 1, k));
    double s13bar = 0.25 * (t13(i, j, k) + t13(i, j, k + 1) + t13(i + 1, j, k) + t13(i + 1, j, k + 1));
    double s23bar = 0.25 * (t23(i, j, k) + t23(i, j, k + 1) + t23(i, j + 1, k) + t23(i, j + 1, k + 1));
    return s11bar * s11bar + s22bar * s22bar + s33bar * s33bar + 2.0 * s12bar * s12bar +
           2.0 * s13bar * s13bar + 2.0 * s23bar * s23bar;
}

// Strain of the old velocity (ERF_advance_dycore.cpp:114-184) then
// ComputeTurbulentViscosityLES, Smagorinsky branch and ghost extrapolation
// (Source/Diffusion/ERF_ComputeTurbulentViscosity.cpp:32-128, 230-330).
void Level::turbulent_viscosity(Arr c, Arr u, Arr v, Arr w) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    Arr t11 = A(tau11), t22 = A(tau22), t33 = A(tau33), t12 = A(tau12), t13 = A(tau13), t23 = A(tau23);
    compute_strain_N(Box(-1, -1, 0, nx, ny - 1 + 1, nz - 1), Box(-1, -1, 0, nx + 1, ny + 1, nz - 1),
                     Box(-1, -1, 0, nx + 1, ny, nz), Box(-1, -1, 0, nx, ny + 1, nz), u, v, w,
                     t11, t22, t33, t12, t13, t23);
    Arr mu = A(eddy), hz = A(hfx3);
    const double Cs = cfg.Cs;
    const double inv_Pr_t = 1.0 / cfg.Pr_t;
    const double inv_Sc_t = 1.0 / cfg.Sc_t;
    const double inv_sigma_k = 1.0 / 1.0;
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), [&](int i, int j, int k) {
        double SmnSmn = smn_smn(i, j, k, t11, t22, t33, t12, t13, t23);
        double dxI = dxInv[0], dyI = dxInv[1], dzI = dxInv[2];
        double cellVolMsf = 1.0 / (dxI * 1.0 * dyI * 1.0 * dzI);
        double DeltaMsf = erfx_pow(cellVolMsf, 1.0 / 3.0);
        double CsDeltaSqrMsf = Cs * Cs * DeltaMsf * DeltaMsf;
        mu(i, j, k, EddyDiff::Mom_h) = CsDeltaSqrMsf * c(i, j, k, 0) * std::sqrt(2.0 * SmnSmn);
        mu(i, j, k, EddyDiff::Mom_v) = mu(i, j, k, EddyDiff::Mom_h);
        double dtheta_dz = 0.5 * (c(i, j, k + 1, 1) / c(i, j, k + 1, 0) - c(i, j, k - 1, 1) / c(i, j, k - 1, 0)) * dzI;
        hz(i, j, k) = -inv_Pr_t * mu(i, j, k, EddyDiff::Mom_v) * dtheta_dz;
    });
    // lateral extrapolation of Mom_h/Mom_v into one ghost layer (:244-280)
    for (int side = 0; side < 2; ++side) {
        int i_g = side == 0 ? -1 : nx, i_s = side == 0 ? 0 : nx - 1;
        loop(Box(i_g, -1, 0, i_g, ny, nz - 1), [&](int i, int j, int k) {
            int lj = std::min(std::max(j, 0), ny - 1);
            mu(i, j, k, EddyDiff::Mom_h) = mu(i_s, lj, k, EddyDiff::Mom_h);
            mu(i, j, k, EddyDiff::Mom_v) = mu(i_s, lj, k, EddyDiff::Mom_v);
        });
    }
    for (int side = 0; side < 2; ++side) {
        int j_g = side == 0 ? -1 : ny, j_s = side == 0 ? 0 : ny - 1;
        loop(Box(-1, j_g, 0, nx, j_g, nz - 1), [&](int i, int j, int k) {
            int li = std::min(std::max(i, 0), nx - 1);
            mu(i, j, k, EddyDiff::Mom_h) = mu(li, j_s, k, EddyDiff::Mom_h);
            mu(i, j, k, EddyDiff::Mom_v) = mu(li, j_s, k, EddyDiff::Mom_v);
        });
    }
    // other diffusivities (:282-296): Theta, Scalar and Q from Mom_h
    const int offset = (EddyDiff::NumDiffs - 1) / 2;
    double fac[5] = {inv_Pr_t, inv_sigma_k, inv_sigma_k, inv_Sc_t, inv_Sc_t};
    for (int n = 1; n < offset; ++n) {
        if (n == EddyDiff::QKE_h || n == EddyDiff::KE_h) continue;
        loop(Box(-1, -1, 0, nx, ny, nz - 1), [&](int i, int j, int k) {
            mu(i, j, k, n) = mu(i, j, k, EddyDiff::Mom_h) * fac[n - 1];
            mu(i, j, k, n + offset) = mu(i, j, k, n);
        });
    }
    fill_boundary_periodic(mu, EddyDiff::NumDiffs, 1, -1);
    // vertical extrapolation into one ghost layer (:298-330)
    for (int n = 0; n < offset; ++n) {
        if (n == EddyDiff::QKE_h || n == EddyDiff::KE_h) continue;
        loop(Box(-1, -1, 0, nx, ny, 0), [&](int i, int j, int) {
            mu(i, j, -1, n) = mu(i, j, 0, n);
            mu(i, j, nz, n) = mu(i, j, nz - 1, n);
            mu(i, j, -1, n + offset) = mu(i, j, 0, n + offset);
            mu(i, j, nz, n + offset) = mu(i, j, nz - 1, n + offset);
        });
    }
}

// erf_make_tau_terms, flat branch (Source/TimeIntegration/ERF_make_tau_terms.cpp:12-428):
// expansion rate, strain, then ComputeStress{Cons,Var}Visc_N
// (Source/Diffusion/ERF_ComputeStress_N.cpp:1-189).
void Level::make_tau_terms(StateA data, Arr u, Arr v, Arr w) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const bool consA = cfg.molec_diff_type == ERFX_MOLEC_CONSTANT_ALPHA;
    const bool turb = cfg.les_type == ERFX_LES_SMAGORINSKY;
    const double mu_eff = consA ? 2.0 * cfg.dynamic_viscosity / cfg.rho0_trans : 2.0 * cfg.dynamic_viscosity;
    Box bxcc(-1, -1, 0, nx, ny, nz - 1);
    Box tbxxy(-1, -1, 0, nx + 1, ny + 1, nz - 1);
    Box tbxxz(-1, -1, 0, nx + 1, ny, nz);
    Box tbxyz(-1, -1, 0, nx, ny + 1, nz);
    Field fer, f11, f22, f33, f12, f13, f23;
    fer.alloc(L, 1); f11.alloc(L, 1); f22.alloc(L, 1); f33.alloc(L, 1); f12.alloc(L, 1); f13.alloc(L, 1); f23.alloc(L, 1);
    Arr er = fer.a(L), s11 = f11.a(L), s22 = f22.a(L), s33 = f33.a(L), s12 = f12.a(L), s13 = f13.a(L), s23 = f23.a(L);
    loop(bxcc, [&](int i, int j, int k) {
        double mfsq = 1.0 * 1.0;
        er(i, j, k) = (u(i + 1, j, k) / 1.0 - u(i, j, k) / 1.0) * dxInv[0] * mfsq +
                      (v(i, j + 1, k) / 1.0 - v(i, j, k) / 1.0) * dxInv[1] * mfsq +
                      (w(i, j, k + 1) - w(i, j, k)) * dxInv[2];
    });
    compute_strain_N(bxcc, tbxxy, tbxxz, tbxyz, u, v, w, s11, s22, s33, s12, s13, s23);
    // stress on edge boxes shrunk back laterally (:262-264); cells keep the grown box
    Box sxy(0, 0, 0, nx, ny, nz - 1), sxz(0, 0, 0, nx, ny - 1, nz), syz(0, 0, 0, nx - 1, ny, nz);
    const double OneThird = (1. / 3.);
    Arr cd = data.cons;
    Arr mt = A(eddy);
    if (!turb) {
        if (consA) {
            loop(bxcc, [&](int i, int j, int k) {
                double rhoAlpha = cd(i, j, k, 0) * mu_eff;
                s11(i, j, k) = -rhoAlpha * (s11(i, j, k) - OneThird * er(i, j, k));
                s22(i, j, k) = -rhoAlpha * (s22(i, j, k) - OneThird * er(i, j, k));
                s33(i, j, k) = -rhoAlpha * (s33(i, j, k) - OneThird * er(i, j, k));
            });
            loop(sxy, [&](int i, int j, int k) {
                double rho_bar = 0.25 * (cd(i - 1, j, k, 0) + cd(i, j, k, 0) + cd(i - 1, j - 1, k, 0) + cd(i, j - 1, k, 0));
                s12(i, j, k) *= -rho_bar * mu_eff;
            });
            loop(sxz, [&](int i, int j, int k) {
                double rho_bar = 0.25 * (cd(i - 1, j, k, 0) + cd(i, j, k, 0) + cd(i - 1, j, k - 1, 0) + cd(i, j, k - 1, 0));
                s13(i, j, k) *= -rho_bar * mu_eff;
            });
            loop(syz, [&](int i, int j, int k) {
                double rho_bar = 0.25 * (cd(i, j - 1, k, 0) + cd(i, j, k, 0) + cd(i, j - 1, k - 1, 0) + cd(i, j, k - 1, 0));
                s23(i, j, k) *= -rho_bar * mu_eff;
            });
        } else {
            loop(bxcc, [&](int i, int j, int k) {
                s11(i, j, k) = -mu_eff * (s11(i, j, k) - OneThird * er(i, j, k));
                s22(i, j, k) = -mu_eff * (s22(i, j, k) - OneThird * er(i, j, k));
                s33(i, j, k) = -mu_eff * (s33(i, j, k) - OneThird * er(i, j, k));
            });
            loop(sxy, [&](int i, int j, int k) { s12(i, j, k) *= -mu_eff; });
            loop(sxz, [&](int i, int j, int k) { s13(i, j, k) *= -mu_eff; });
            loop(syz, [&](int i, int j, int k) { s23(i, j, k) *= -mu_eff; });
        }
    } else {
        using namespace EddyDiff;
        if (consA) {
            loop(bxcc, [&](int i, int j, int k) {
                double rhoAlpha = cd(i, j, k, 0) * mu_eff;
                double mu_11 = rhoAlpha + 2.0 * mt(i, j, k, Mom_h);
                double mu_22 = mu_11;
                double mu_33 = rhoAlpha + 2.0 * mt(i, j, k, Mom_v);
                s11(i, j, k) = -mu_11 * (s11(i, j, k) - OneThird * er(i, j, k));
                s22(i, j, k) = -mu_22 * (s22(i, j, k) - OneThird * er(i, j, k));
                s33(i, j, k) = -mu_33 * (s33(i, j, k) - OneThird * er(i, j, k));
            });
            loop(sxy, [&](int i, int j, int k) {
                double rho_bar = 0.25 * (cd(i - 1, j, k, 0) + cd(i, j, k, 0) + cd(i - 1, j - 1, k, 0) + cd(i, j - 1, k, 0));
                double mu_bar = 0.25 * (mt(i - 1, j, k, Mom_h) + mt(i, j, k, Mom_h) + mt(i - 1, j - 1, k, Mom_h) + mt(i, j - 1, k, Mom_h));
                double mu_12 = rho_bar * mu_eff + 2.0 * mu_bar;
                s12(i, j, k) *= -mu_12;
            });
            loop(sxz, [&](int i, int j, int k) {
                double rho_bar = 0.25 * (cd(i - 1, j, k, 0) + cd(i, j, k, 0) + cd(i - 1, j, k - 1, 0) + cd(i, j, k - 1, 0));
                double mu_bar = 0.25 * (mt(i - 1, j, k, Mom_v) + mt(i, j, k, Mom_v) + mt(i - 1, j, k - 1, Mom_v) + mt(i, j, k - 1, Mom_v));
                double mu_13 = rho_bar * mu_eff + 2.0 * mu_bar;
                s13(i, j, k) *= -mu_13;
            });
            loop(syz, [&](int i, int j, int k) {
                double rho_bar = 0.25 * (cd(i, j - 1, k, 0) + cd(i, j, k, 0) + cd(i, j - 1, k - 1, 0) + cd(i, j, k - 1, 0));
                double mu_bar = 0.25 * (mt(i, j - 1, k, Mom_v) + mt(i, j, k, Mom_v) + mt(i, j - 1, k - 1, Mom_v) + mt(i, j, k - 1, Mom_v));
                double mu_23 = rho_bar * mu_eff + 2.0 * mu_bar;
                s23(i, j, k) *= -mu_23;
            });
        } else {
            loop(bxcc, [&](int i, int j, int k) {
                double mu_11 = mu_eff + 2.0 * mt(i, j, k, Mom_h);
                double mu_22 = mu_11;
                double mu_33 = mu_eff + 2.0 * mt(i, j, k, Mom_v);
                s11(i, j, k) = -mu_11 * (s11(i, j, k) - OneThird * er(i, j, k));
                s22(i, j, k) = -mu_22 * (s22(i, j, k) - OneThird * er(i, j, k));
                s33(i, j, k) = -mu_33 * (s33(i, j, k) - OneThird * er(i, j, k));
            });
            loop(sxy, [&](int i, int j, int k) {
                double mu_bar = 0.25 * (mt(i - 1, j, k, Mom_h) + mt(i, j, k, Mom_h) + mt(i - 1, j - 1, k, Mom_h) + mt(i, j - 1, k, Mom_h));
                double mu_12 = mu_eff + 2.0 * mu_bar;
                s12(i, j, k) *= -mu_12;
            });
            loop(sxz, [&](int i, int j, int k) {
                double mu_bar = 0.25 * (mt(i - 1, j, k, Mom_v) + mt(i, j, k, Mom_v) + mt(i - 1, j, k - 1, Mom_v) + mt(i, j, k - 1, Mom_v));
                double mu_13 = mu_eff + 2.0 * mu_bar;
                s13(i, j, k) *= -mu_13;
            });
            loop(syz, [&](int i, int j, int k) {
                double mu_bar = 0.25 * (mt(i, j - 1, k, Mom_v) + mt(i, j, k, Mom_v) + mt(i, j - 1, k - 1, Mom_v) + mt(i, j, k - 1, Mom_v));
                double mu_23 = mu_eff + 2.0 * mu_bar;
                s23(i, j, k) *= -mu_23;
            });
        }
    }
    // copy into the persistent tau fields (:284-303): cells on valid grown by 1
    // laterally, edges on the nodal tile boxes
    Arr t11 = A(tau11), t22 = A(tau22), t33 = A(tau33), t12 = A(tau12), t13 = A(tau13), t23 = A(tau23);
    loop(Box(-1, -1, 0, nx, ny, nz - 1), [&](int i, int j, int k) {
        t11(i, j, k) = s11(i, j, k); t22(i, j, k) = s22(i, j, k); t33(i, j, k) = s33(i, j, k);
    });
    loop(sxy, [&](int i, int j, int k) { t12(i, j, k) = s12(i, j, k); });
    loop(sxz, [&](int i, int j, int k) { t13(i, j, k) = s13(i, j, k); });
    loop(syz, [&](int i, int j, int k) { t23(i, j, k) = s23(i, j, k); });
}

// DiffusionSrcForState_N (Source/Diffusion/ERF_DiffusionSrcForState_N.cpp:40-738)
// for cons components [start_comp, start_comp+num_comp); no Dirichlet cons
// faces.  When configured, the rho*theta z-flux at k = 0 is the prescribed
// constant surface flux, placed where the explicit-MOST branch puts hfx_z
// (:300-301 / :414).  The SGS flux diagnostic hfx_z is not kept.
void Level::diffusion_src_state(const Box& bx, int start_comp, int num_comp, Arr cd, Arr cp, Arr rhs) {
    const int nx = bx.hi[0], ny = bx.hi[1], nz = bx.hi[2];
    const bool consA = cfg.molec_diff_type == ERFX_MOLEC_CONSTANT_ALPHA;
    const bool turb = cfg.les_type == ERFX_LES_SMAGORINSKY;
    Arr mt = A(eddy);
    // prim indices: theta, KE, QKE, scalar, then the moisture species (:103-186)
    double alpha_eff[8] = {0};
    const double aT = consA ? cfg.alpha_T : cfg.rho0_trans * cfg.alpha_T;
    const double aC = consA ? cfg.alpha_C : cfg.rho0_trans * cfg.alpha_C;
    alpha_eff[0] = aT;
    for (int p = 3; p < 8; ++p) alpha_eff[p] = aC;
    using namespace EddyDiff;
    const int idx_h[8] = {Theta_h, KE_h, QKE_h, Scalar_h, Q_h, Q_h, Q_h, Q_h};
    const int idx_v[8] = {Theta_v, KE_v, QKE_v, Scalar_v, Q_v, Q_v, Q_v, Q_v};
    const double dx_inv = dxInv[0], dy_inv = dxInv[1], dz_inv = dxInv[2];
    // Dirichlet (ext_dir) cons faces of an Inflow boundary: one-sided
    // second-order stencils through the face value held in the ghost cell
    // (:208-227 and the same test in the other three flux branches)
    const bool xlo_dir = bck[3][0] == EXT_DIR, xhi_dir = bck[3][3] == EXT_DIR;
    const bool ylo_dir = bck[3][1] == EXT_DIR, yhi_dir = bck[3][4] == EXT_DIR;
    Field fx, fy, fz;
    fx.alloc(L, 1); fy.alloc(L, 1); fz.alloc(L, 1);
    Arr xflux = fx.a(L), yflux = fy.a(L), zflux = fz.a(L);
    for (int n = 0; n < num_comp; ++n) {
        const int qty = start_comp + n, prim_index = qty - 1;
        const double ae = alpha_eff[prim_index];
        const int ih = idx_h[prim_index], iv = idx_v[prim_index];
        loop(Box(0, 0, 0, nx + 1, ny, nz), [&](int i, int j, int k) {
            double rhoAlpha;
            if (consA) {
                double rhoFace = 0.5 * (cd(i, j, k, 0) + cd(i - 1, j, k, 0));
                rhoAlpha = rhoFace * ae;
            } else rhoAlpha = ae;
            if (turb) rhoAlpha += 0.5 * (mt(i, j, k, ih) + mt(i - 1, j, k, ih));
            if (xlo_dir && i == 0)
                xflux(i, j, k) = -rhoAlpha * (-(8. / 3.) * cp(i - 1, j, k, prim_index) + 3. * cp(i, j, k, prim_index)
                                              - (1. / 3.) * cp(i + 1, j, k, prim_index)) * dx_inv;
            else if (xhi_dir && i == nx + 1)
                xflux(i, j, k) = -rhoAlpha * ((8. / 3.) * cp(i, j, k, prim_index) - 3. * cp(i - 1, j, k, prim_index)
                                              + (1. / 3.) * cp(i - 2, j, k, prim_index)) * dx_inv;
            else
                xflux(i, j, k) = -rhoAlpha * (cp(i, j, k, prim_index) - cp(i - 1, j, k, prim_index)) * dx_inv * 1.0;
        });
        loop(Box(0, 0, 0, nx, ny + 1, nz), [&](int i, int j, int k) {
            double rhoAlpha;
            if (consA) {
                double rhoFace = 0.5 * (cd(i, j, k, 0) + cd(i, j - 1, k, 0));
                rhoAlpha = rhoFace * ae;
            } else rhoAlpha = ae;
            if (turb) rhoAlpha += 0.5 * (mt(i, j, k, ih) + mt(i, j - 1, k, ih));
            if (ylo_dir && j == 0)
                yflux(i, j, k) = -rhoAlpha * (-(8. / 3.) * cp(i, j - 1, k, prim_index) + 3. * cp(i, j, k, prim_index)
                                              - (1. / 3.) * cp(i, j + 1, k, prim_index)) * dy_inv;
            else if (yhi_dir && j == ny + 1)
                yflux(i, j, k) = -rhoAlpha * ((8. / 3.) * cp(i, j, k, prim_index) - 3. * cp(i, j - 1, k, prim_index)
                                              + (1. / 3.) * cp(i, j - 2, k, prim_index)) * dy_inv;
            else
                yflux(i, j, k) = -rhoAlpha * (cp(i, j, k, prim_index) - cp(i, j - 1, k, prim_index)) * dy_inv * 1.0;
        });
        const bool most_lo = cfg.use_surface_heat_flux && qty == 1;
        loop(Box(0, 0, 0, nx, ny, nz + 1), [&](int i, int j, int k) {
            double rhoAlpha;
            if (consA) {
                double rhoFace = 0.5 * (cd(i, j, k, 0) + cd(i, j, k - 1, 0));
                rhoAlpha = rhoFace * ae;
            } else rhoAlpha = ae;
            if (turb) rhoAlpha += 0.5 * (mt(i, j, k, iv) + mt(i, j, k - 1, iv));
            bool most_on_zlo = most_lo && k == 0;
            if (most_on_zlo) zflux(i, j, k) = cfg.surface_heat_flux;
            else zflux(i, j, k) = -rhoAlpha * (cp(i, j, k, prim_index) - cp(i, j, k - 1, prim_index)) * dz_inv;
        });
        loop(bx, [&](int i, int j, int k) {
            rhs(i, j, k, qty) -= (xflux(i + 1, j, k) - xflux(i, j, k)) * dx_inv * 1.0 +
                                 (yflux(i, j + 1, k) - yflux(i, j, k)) * dy_inv * 1.0 +
                                 (zflux(i, j, k + 1) - zflux(i, j, k)) * dz_inv;
        });
    }
}

}  // namespace orc
