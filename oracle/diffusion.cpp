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

namespace orc {

// ---- terrain-following grid ---------------------------------------------------
// ComputeStrain_T (Source/Diffusion/ERF_ComputeStrain_T.cpp:33-507), unit map
// factors.  The lateral Dirichlet branches test the grown tile boxes against
// the domain, so on one box they never apply (the grown boxes start at -1);
// the z branches do.
void Level::compute_strain_T(Box bxcc, Box tbxxy, Box tbxxz, Box tbxyz, Arr u, Arr v, Arr w, Arr t11, Arr t22,
                             Arr t33, Arr t12, Arr t13, Arr t21, Arr t23, Arr t31, Arr t32) {
    const double* dI = dxInv;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    Arr z = A(z_nd), J = A(detJ);
    auto isdir = [&](int var, int ori) { return bck[var][ori] == EXT_DIR; };
    const bool xl_v_dir = isdir(1, 0) && tbxxy.lo[0] == 0, xh_v_dir = isdir(1, 3) && tbxxy.hi[0] == nx;
    const bool xl_w_dir = isdir(2, 0) && tbxxz.lo[0] == 0, xh_w_dir = isdir(2, 3) && tbxxz.hi[0] == nx;
    const bool yl_u_dir = isdir(0, 1) && tbxxy.lo[1] == 0, yh_u_dir = isdir(0, 4) && tbxxy.hi[1] == ny;
    const bool yl_w_dir = isdir(2, 1) && tbxyz.lo[1] == 0, yh_w_dir = isdir(2, 4) && tbxyz.hi[1] == ny;
    const bool zl_u_dir = isdir(0, 2) && tbxxz.lo[2] == 0, zh_u_dir = isdir(0, 5) && tbxxz.hi[2] == nz;
    const bool zl_v_dir = isdir(1, 2) && tbxyz.lo[2] == 0, zh_v_dir = isdir(1, 5) && tbxyz.hi[2] == nz;
    // the strain pieces shared by the branches (:103-108, :452-455, :470-473)
    auto gradUz_xy = [&](int i, int j, int k) {
        return 0.25 * dI[2] * (u(i, j, k + 1) / 1.0 + u(i, j - 1, k + 1) / 1.0 - u(i, j, k - 1) / 1.0 - u(i, j - 1, k - 1) / 1.0);
    };
    auto gradVz_xy = [&](int i, int j, int k) {
        return 0.25 * dI[2] * (v(i, j, k + 1) / 1.0 + v(i - 1, j, k + 1) / 1.0 - v(i, j, k - 1) / 1.0 - v(i - 1, j, k - 1) / 1.0);
    };
    // tau12 with the x-derivative of v and the y-derivative of u given
    auto t12_of = [&](int i, int j, int k, double dudy, double dvdx) {
        double met_h_xi = met::h_xi_EdgeK(z, i, j, k, dI[0]);
        double met_h_eta = met::h_eta_EdgeK(z, i, j, k, dI[1]);
        double met_h_zeta = met::h_zeta_EdgeK(z, i, j, k, dI[2]);
        return 0.5 * (dudy + dvdx - (met_h_eta / met_h_zeta) * gradUz_xy(i, j, k)
                      - (met_h_xi / met_h_zeta) * gradVz_xy(i, j, k)) * 1.0 * 1.0;
    };
    if (xl_v_dir) {
        Box p = tbxxy; p.hi[0] = p.lo[0]; tbxxy.lo[0] += 1;
        loop(p, [&](int i, int j, int k) {
            t12(i, j, k) = t12_of(i, j, k, (u(i, j, k) / 1.0 - u(i, j - 1, k) / 1.0) * dI[1],
                                  (-(8. / 3.) * v(i - 1, j, k) / 1.0 + 3. * v(i, j, k) / 1.0 - (1. / 3.) * v(i + 1, j, k) / 1.0) * dI[0]);
            t21(i, j, k) = t12(i, j, k);
        });
    }
    if (xh_v_dir) {
        Box p = tbxxy; p.lo[0] = p.hi[0]; tbxxy.hi[0] -= 1;
        loop(p, [&](int i, int j, int k) {
            t12(i, j, k) = t12_of(i, j, k, (u(i, j, k) / 1.0 - u(i, j - 1, k) / 1.0) * dI[1],
                                  -(-(8. / 3.) * v(i, j, k) / 1.0 + 3. * v(i - 1, j, k) / 1.0 - (1. / 3.) * v(i - 2, j, k) / 1.0) * dI[0]);
            t21(i, j, k) = t12(i, j, k);
        });
    }
    auto gradWz_xz = [&](int i, int j, int k) {
        return 0.25 * dI[2] * (w(i, j, k + 1) + w(i - 1, j, k + 1) - w(i, j, k - 1) - w(i - 1, j, k - 1));
    };
    auto gradWz_yz = [&](int i, int j, int k) {
        return 0.25 * dI[2] * (w(i, j, k + 1) + w(i, j - 1, k + 1) - w(i, j, k - 1) - w(i, j - 1, k - 1));
    };
    if (xl_w_dir) {
        Box p = tbxxz; p.hi[0] = p.lo[0]; p.lo[2] += 1; p.hi[2] -= 1; tbxxz.lo[0] += 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dI[0]), met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
            t13(i, j, k) = 0.5 * ((u(i, j, k) - u(i, j, k - 1)) * dI[2] / met_h_zeta
                                  + ((-(8. / 3.) * w(i - 1, j, k) + 3. * w(i, j, k) - (1. / 3.) * w(i + 1, j, k)) * dI[0]
                                     - (met_h_xi / met_h_zeta) * gradWz_xz(i, j, k)) * 1.0);
            t31(i, j, k) = t13(i, j, k);
        });
    }
    if (xh_w_dir) {
        Box p = tbxxz; p.lo[0] = p.hi[0]; p.lo[2] += 1; p.hi[2] -= 1; tbxxz.hi[0] -= 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dI[0]), met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
            t13(i, j, k) = 0.5 * ((u(i, j, k) - u(i, j, k - 1)) * dI[2] / met_h_zeta
                                  - ((-(8. / 3.) * w(i, j, k) + 3. * w(i - 1, j, k) - (1. / 3.) * w(i - 2, j, k)) * dI[0]
                                     - (met_h_xi / met_h_zeta) * gradWz_xz(i, j, k)) * 1.0);
            t31(i, j, k) = t13(i, j, k);
        });
    }
    if (yl_u_dir) {
        Box p = tbxxy; p.hi[1] = p.lo[1]; tbxxy.lo[1] += 1;
        loop(p, [&](int i, int j, int k) {
            t12(i, j, k) = t12_of(i, j, k,
                                  (-(8. / 3.) * u(i, j - 1, k) / 1.0 + 3. * u(i, j, k) / 1.0 - (1. / 3.) * u(i, j + 1, k) / 1.0) * dI[1],
                                  (v(i, j, k) / 1.0 - v(i - 1, j, k) / 1.0) * dI[0]);
            t21(i, j, k) = t12(i, j, k);
        });
    }
    if (yh_u_dir) {
        Box p = tbxxy; p.lo[1] = p.hi[1]; tbxxy.hi[1] -= 1;
        loop(p, [&](int i, int j, int k) {
            t12(i, j, k) = t12_of(i, j, k,
                                  -(-(8. / 3.) * u(i, j, k) / 1.0 + 3. * u(i, j - 1, k) / 1.0 - (1. / 3.) * u(i, j - 2, k) / 1.0) * dI[1],
                                  (v(i, j, k) / 1.0 - v(i - 1, j, k) / 1.0) * dI[0]);
            t21(i, j, k) = t12(i, j, k);
        });
    }
    if (yl_w_dir) {
        Box p = tbxyz; p.hi[1] = p.lo[1]; p.lo[2] += 1; p.hi[2] -= 1; tbxyz.lo[1] += 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_eta = met::h_eta_EdgeI(z, i, j, k, dI[1]), met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
            t23(i, j, k) = 0.5 * ((v(i, j, k) - v(i, j, k - 1)) * dI[2] / met_h_zeta
                                  + ((-(8. / 3.) * w(i, j - 1, k) + 3. * w(i, j, k) - (1. / 3.) * w(i, j + 1, k)) * dI[1] * 1.0
                                     - (met_h_eta / met_h_zeta) * gradWz_yz(i, j, k)) * 1.0);
            t32(i, j, k) = t23(i, j, k);
        });
    }
    if (yh_w_dir) {
        Box p = tbxyz; p.lo[1] = p.hi[1]; p.lo[2] += 1; p.hi[2] -= 1; tbxyz.hi[1] -= 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_eta = met::h_eta_EdgeI(z, i, j, k, dI[1]), met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
            t23(i, j, k) = 0.5 * ((v(i, j, k) - v(i, j, k - 1)) * dI[2] / met_h_zeta
                                  - ((-(8. / 3.) * w(i, j, k) + 3. * w(i, j - 1, k) - (1. / 3.) * w(i, j - 2, k)) * dI[1]
                                     - (met_h_eta / met_h_zeta) * gradWz_yz(i, j, k)) * 1.0);
            t32(i, j, k) = t23(i, j, k);
        });
    }
    if (zl_u_dir) {
        Box p = tbxxz; p.hi[2] = p.lo[2]; tbxxz.lo[2] += 1;
        loop(p, [&](int i, int j, int k) {
            double GradWz = 0.5 * dI[2] * (w(i, j, k + 1) + w(i - 1, j, k + 1) - w(i, j, k) - w(i - 1, j, k));
            double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dI[0]), met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
            t13(i, j, k) = 0.5 * ((-(8. / 3.) * u(i, j, k - 1) + 3. * u(i, j, k) - (1. / 3.) * u(i, j, k + 1)) * dI[2] / met_h_zeta
                                  + ((w(i, j, k) - w(i - 1, j, k)) * dI[0] - (met_h_xi / met_h_zeta) * GradWz) * 1.0);
            t31(i, j, k) = t13(i, j, k);
        });
    }
    if (zh_u_dir) {
        Box p = tbxxz; p.lo[2] = p.hi[2]; tbxxz.hi[2] -= 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
            t13(i, j, k) = 0.5 * (-(-(8. / 3.) * u(i, j, k) + 3. * u(i, j, k - 1) - (1. / 3.) * u(i, j, k - 2)) * dI[2] / met_h_zeta
                                  + (w(i, j, k) - w(i - 1, j, k)) * dI[0] * 1.0);
            t31(i, j, k) = t13(i, j, k);
        });
    }
    if (zl_v_dir) {
        Box p = tbxyz; p.hi[2] = p.lo[2]; tbxyz.lo[2] += 1;
        loop(p, [&](int i, int j, int k) {
            double GradWz = 0.5 * dI[2] * (w(i, j, k + 1) + w(i, j - 1, k + 1) - w(i, j, k) - w(i, j - 1, k));
            double met_h_eta = met::h_eta_EdgeI(z, i, j, k, dI[1]), met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
            t23(i, j, k) = 0.5 * ((-(8. / 3.) * v(i, j, k - 1) + 3. * v(i, j, k) - (1. / 3.) * v(i, j, k + 1)) * dI[2] / met_h_zeta
                                  + ((w(i, j, k) - w(i, j - 1, k)) * dI[1] - (met_h_eta / met_h_zeta) * GradWz) * 1.0);
            t32(i, j, k) = t23(i, j, k);
        });
    }
    if (zh_v_dir) {
        Box p = tbxyz; p.lo[2] = p.hi[2]; tbxyz.hi[2] -= 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
            t23(i, j, k) = 0.5 * (-(-(8. / 3.) * v(i, j, k) + 3. * v(i, j, k - 1) - (1. / 3.) * v(i, j, k - 2)) * dI[2] / met_h_zeta
                                  + (w(i, j, k) - w(i, j - 1, k)) * dI[1] * 1.0);
            t32(i, j, k) = t23(i, j, k);
        });
    }
    // one-sided vertical derivatives at a no-slip bottom (:312-352)
    if (zl_u_dir && zl_v_dir) {
        Box pc = bxcc; pc.hi[2] = pc.lo[2]; bxcc.lo[2] += 1;
        loop(pc, [&](int i, int j, int k) {
            double GradUz = 0.5 * dI[2] * ((-(8. / 3.) * u(i, j, k - 1) + 3. * u(i, j, k) - (1. / 3.) * u(i, j, k + 1))
                                         + (-(8. / 3.) * u(i - 1, j, k - 1) + 3. * u(i - 1, j, k) - (1. / 3.) * u(i - 1, j, k + 1)));
            double GradVz = 0.5 * dI[2] * ((-(8. / 3.) * v(i, j, k - 1) + 3. * v(i, j, k) - (1. / 3.) * v(i, j, k + 1))
                                         + (-(8. / 3.) * v(i, j - 1, k - 1) + 3. * v(i, j - 1, k) - (1. / 3.) * v(i, j - 1, k + 1)));
            double met_h_xi = met::h_xi_Cell(z, i, j, k, dI[0]), met_h_eta = met::h_eta_Cell(z, i, j, k, dI[1]);
            double met_h_zeta = J(i, j, k);
            t11(i, j, k) = ((u(i + 1, j, k) / 1.0 - u(i, j, k) / 1.0) * dI[0] - (met_h_xi / met_h_zeta) * GradUz) * 1.0 * 1.0;
            t22(i, j, k) = ((v(i, j + 1, k) / 1.0 - v(i, j, k) / 1.0) * dI[1] - (met_h_eta / met_h_zeta) * GradVz) * 1.0 * 1.0;
            t33(i, j, k) = (w(i, j, k + 1) - w(i, j, k)) * dI[2] / met_h_zeta;
        });
        Box pxy = tbxxy; pxy.hi[2] = pxy.lo[2]; tbxxy.lo[2] += 1;
        loop(pxy, [&](int i, int j, int k) {
            double GradUz = 0.5 * dI[2] * ((-(8. / 3.) * u(i, j, k - 1) + 3. * u(i, j, k) - (1. / 3.) * u(i, j, k + 1))
                                         + (-(8. / 3.) * u(i, j - 1, k - 1) + 3. * u(i, j - 1, k) - (1. / 3.) * u(i, j - 1, k + 1)));
            double GradVz = 0.5 * dI[2] * ((-(8. / 3.) * v(i, j, k - 1) + 3. * v(i, j, k) - (1. / 3.) * v(i, j, k + 1))
                                         + (-(8. / 3.) * v(i - 1, j, k - 1) + 3. * v(i - 1, j, k) - (1. / 3.) * v(i - 1, j, k + 1)));
            double met_h_xi = met::h_xi_EdgeK(z, i, j, k, dI[0]), met_h_eta = met::h_eta_EdgeK(z, i, j, k, dI[1]);
            double met_h_zeta = met::h_zeta_EdgeK(z, i, j, k, dI[2]);
            t12(i, j, k) = 0.5 * ((u(i, j, k) / 1.0 - u(i, j - 1, k) / 1.0) * dI[1]
                                  + (v(i, j, k) / 1.0 - v(i - 1, j, k) / 1.0) * dI[0]
                                  - (met_h_eta / met_h_zeta) * GradUz - (met_h_xi / met_h_zeta) * GradVz) * 1.0 * 1.0;
            t21(i, j, k) = t12(i, j, k);
        });
    }
    // bottom face without a Dirichlet velocity: GradWz one-sided (:370-408)
    if (!zl_u_dir && tbxxz.lo[2] == 0) {
        Box p = tbxxz; p.hi[2] = p.lo[2]; tbxxz.lo[2] += 1;
        loop(p, [&](int i, int j, int k) {
            double GradWz = 0.5 * dI[2] * (w(i, j, k + 1) + w(i - 1, j, k + 1) - w(i, j, k) - w(i - 1, j, k));
            double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dI[0]), met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
            t13(i, j, k) = 0.5 * ((u(i, j, k) - u(i, j, k - 1)) * dI[2] / met_h_zeta
                                  + ((w(i, j, k) - w(i - 1, j, k)) * dI[0] - (met_h_xi / met_h_zeta) * GradWz) * 1.0);
            t31(i, j, k) = t13(i, j, k);
        });
    }
    if (!zl_v_dir && tbxyz.lo[2] == 0) {
        Box p = tbxyz; p.hi[2] = p.lo[2]; tbxyz.lo[2] += 1;
        loop(p, [&](int i, int j, int k) {
            double GradWz = 0.5 * dI[2] * (w(i, j, k + 1) + w(i, j - 1, k + 1) - w(i, j, k) - w(i, j - 1, k));
            double met_h_eta = met::h_eta_EdgeI(z, i, j, k, dI[1]), met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
            t23(i, j, k) = 0.5 * ((v(i, j, k) - v(i, j, k - 1)) * dI[2] / met_h_zeta
                                  + ((w(i, j, k) - w(i, j - 1, k)) * dI[1] - (met_h_eta / met_h_zeta) * GradWz) * 1.0);
            t32(i, j, k) = t23(i, j, k);
        });
    }
    // top face without a Dirichlet velocity, h_xi = h_eta = 0 there (:410-438)
    if (!zh_u_dir && tbxxz.hi[2] == nz) {
        Box p = tbxxz; p.lo[2] = p.hi[2]; tbxxz.hi[2] -= 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
            t13(i, j, k) = 0.5 * ((u(i, j, k) - u(i, j, k - 1)) * dI[2] / met_h_zeta + (w(i, j, k) - w(i - 1, j, k)) * dI[0] * 1.0);
            t31(i, j, k) = t13(i, j, k);
        });
    }
    if (!zh_v_dir && tbxyz.hi[2] == nz) {
        Box p = tbxyz; p.lo[2] = p.hi[2]; tbxyz.hi[2] -= 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
            t23(i, j, k) = 0.5 * ((v(i, j, k) - v(i, j, k - 1)) * dI[2] / met_h_zeta + (w(i, j, k) - w(i, j - 1, k)) * dI[1] * 1.0);
            t32(i, j, k) = t23(i, j, k);
        });
    }
    // interior (:440-506)
    loop(bxcc, [&](int i, int j, int k) {
        double GradUz = 0.25 * dI[2] * (u(i, j, k + 1) / 1.0 + u(i - 1, j, k + 1) / 1.0 - u(i, j, k - 1) / 1.0 - u(i - 1, j, k - 1) / 1.0);
        double GradVz = 0.25 * dI[2] * (v(i, j, k + 1) / 1.0 + v(i, j - 1, k + 1) / 1.0 - v(i, j, k - 1) / 1.0 - v(i, j - 1, k - 1) / 1.0);
        double met_h_xi = met::h_xi_Cell(z, i, j, k, dI[0]), met_h_eta = met::h_eta_Cell(z, i, j, k, dI[1]);
        double met_h_zeta = J(i, j, k);
        t11(i, j, k) = ((u(i + 1, j, k) / 1.0 - u(i, j, k) / 1.0) * dI[0] - (met_h_xi / met_h_zeta) * GradUz) * 1.0 * 1.0;
        t22(i, j, k) = ((v(i, j + 1, k) / 1.0 - v(i, j, k) / 1.0) * dI[1] - (met_h_eta / met_h_zeta) * GradVz) * 1.0 * 1.0;
        t33(i, j, k) = (w(i, j, k + 1) - w(i, j, k)) * dI[2] / met_h_zeta;
    });
    loop(tbxxy, [&](int i, int j, int k) {
        t12(i, j, k) = t12_of(i, j, k, (u(i, j, k) / 1.0 - u(i, j - 1, k) / 1.0) * dI[1],
                              (v(i, j, k) / 1.0 - v(i - 1, j, k) / 1.0) * dI[0]);
        t21(i, j, k) = t12(i, j, k);
    });
    loop(tbxxz, [&](int i, int j, int k) {
        double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dI[0]), met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
        t13(i, j, k) = 0.5 * ((u(i, j, k) - u(i, j, k - 1)) * dI[2] / met_h_zeta
                              + ((w(i, j, k) - w(i - 1, j, k)) * dI[0] - (met_h_xi / met_h_zeta) * gradWz_xz(i, j, k)) * 1.0);
        t31(i, j, k) = t13(i, j, k);
    });
    loop(tbxyz, [&](int i, int j, int k) {
        double met_h_eta = met::h_eta_EdgeI(z, i, j, k, dI[1]), met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
        t23(i, j, k) = 0.5 * ((v(i, j, k) - v(i, j, k - 1)) * dI[2] / met_h_zeta
                              + ((w(i, j, k) - w(i, j - 1, k)) * dI[1] - (met_h_eta / met_h_zeta) * gradWz_yz(i, j, k)) * 1.0);
        t32(i, j, k) = t23(i, j, k);
    });
}

// ComputeStressConsVisc_T (Source/Diffusion/ERF_ComputeStress_T.cpp:29-252):
// S - D, then 2 mu J^T (S - D) with the metric terms; tau13 / tau23 on the
// bottom and top faces from linearly extrapolated tau11 / tau12 (tau21 /
// tau22) averages.  cd null: constant mu (molec "Constant"); else rho * alpha.
void Level::compute_stress_cons_T(Box bxcc, Box tbxxy, Box tbxxz, Box tbxyz, double mu_eff, bool use_rho, Arr cd,
                                  Arr t11, Arr t22, Arr t33, Arr t12, Arr t13, Arr t21, Arr t23, Arr t31, Arr t32,
                                  Arr er) {
    const double* dI = dxInv;
    Arr z = A(z_nd), J = A(detJ);
    Field fra; fra.alloc(L, 1);
    Arr rhoAlpha = fra.a(L);
    Box gbx = bxcc; gbx.lo[2] -= 1; gbx.hi[2] += 1;
    loop(gbx, [&](int i, int j, int k) { rhoAlpha(i, j, k) = use_rho ? cd(i, j, k, 0) * mu_eff : mu_eff; });
    Box bxcc2 = bxcc; bxcc2.lo[0] += 1; bxcc2.lo[1] += 1; bxcc2.hi[0] -= 1; bxcc2.hi[1] -= 1;
    const double OneThird = (1. / 3.);
    loop(bxcc, [&](int i, int j, int k) {
        t11(i, j, k) -= OneThird * er(i, j, k);
        t22(i, j, k) -= OneThird * er(i, j, k);
        t33(i, j, k) -= OneThird * er(i, j, k);
    });
    loop(bxcc2, [&](int i, int j, int k) {
        double met_h_xi = met::h_xi_Cell(z, i, j, k, dI[0]), met_h_eta = met::h_eta_Cell(z, i, j, k, dI[1]);
        double tau31bar = 0.25 * (t31(i, j, k) + t31(i + 1, j, k) + t31(i, j, k + 1) + t31(i + 1, j, k + 1));
        double tau32bar = 0.25 * (t32(i, j, k) + t32(i, j + 1, k) + t32(i, j, k + 1) + t32(i, j + 1, k + 1));
        double mu_tot = rhoAlpha(i, j, k);
        t33(i, j, k) -= met_h_xi * tau31bar + met_h_eta * tau32bar;
        t33(i, j, k) *= -mu_tot;
    });
    // tau13, tau23 at the bottom (extrapolated averages)
    {
        Box p = tbxxz; p.hi[2] = p.lo[2]; tbxxz.lo[2] += 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dI[0]), met_h_eta = met::h_eta_EdgeJ(z, i, j, k, dI[1]);
            double met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
            double tau11lo = 0.5 * (t11(i, j, k) + t11(i - 1, j, k));
            double tau11hi = 0.5 * (t11(i, j, k + 1) + t11(i - 1, j, k + 1));
            double tau11bar = 1.5 * tau11lo - 0.5 * tau11hi;
            double tau12lo = 0.5 * (t12(i, j, k) + t12(i, j + 1, k));
            double tau12hi = 0.5 * (t12(i, j, k + 1) + t12(i, j + 1, k + 1));
            double tau12bar = 1.5 * tau12lo - 0.5 * tau12hi;
            double mu_tot = 0.25 * (rhoAlpha(i - 1, j, k) + rhoAlpha(i, j, k) + rhoAlpha(i - 1, j, k - 1) + rhoAlpha(i, j, k - 1));
            t13(i, j, k) -= met_h_xi * tau11bar + met_h_eta * tau12bar;
            t13(i, j, k) *= -mu_tot;
            t31(i, j, k) *= -mu_tot * met_h_zeta;
        });
        Box q = tbxyz; q.hi[2] = q.lo[2]; tbxyz.lo[2] += 1;
        loop(q, [&](int i, int j, int k) {
            double met_h_xi = met::h_xi_EdgeI(z, i, j, k, dI[0]), met_h_eta = met::h_eta_EdgeI(z, i, j, k, dI[1]);
            double met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
            double tau21lo = 0.5 * (t21(i, j, k) + t21(i + 1, j, k));
            double tau21hi = 0.5 * (t21(i, j, k + 1) + t21(i + 1, j, k + 1));
            double tau21bar = 1.5 * tau21lo - 0.5 * tau21hi;
            double tau22lo = 0.5 * (t22(i, j, k) + t22(i, j - 1, k));
            double tau22hi = 0.5 * (t22(i, j, k + 1) + t22(i, j - 1, k + 1));
            double tau22bar = 1.5 * tau22lo - 0.5 * tau22hi;
            double mu_tot = 0.25 * (rhoAlpha(i, j - 1, k) + rhoAlpha(i, j, k) + rhoAlpha(i, j - 1, k - 1) + rhoAlpha(i, j, k - 1));
            t23(i, j, k) -= met_h_xi * tau21bar + met_h_eta * tau22bar;
            t23(i, j, k) *= -mu_tot;
            t32(i, j, k) *= -mu_tot * met_h_zeta;
        });
    }
    // ... and at the top
    {
        Box p = tbxxz; p.lo[2] = p.hi[2]; tbxxz.hi[2] -= 1;
        loop(p, [&](int i, int j, int k) {
            double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dI[0]), met_h_eta = met::h_eta_EdgeJ(z, i, j, k, dI[1]);
            double met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
            double tau11lo = 0.5 * (t11(i, j, k - 2) + t11(i - 1, j, k - 2));
            double tau11hi = 0.5 * (t11(i, j, k - 1) + t11(i - 1, j, k - 1));
            double tau11bar = 1.5 * tau11hi - 0.5 * tau11lo;
            double tau12lo = 0.5 * (t12(i, j, k - 2) + t12(i, j + 1, k - 2));
            double tau12hi = 0.5 * (t12(i, j, k - 1) + t12(i, j + 1, k - 1));
            double tau12bar = 1.5 * tau12hi - 0.5 * tau12lo;
            double mu_tot = 0.25 * (rhoAlpha(i - 1, j, k) + rhoAlpha(i, j, k) + rhoAlpha(i - 1, j, k - 1) + rhoAlpha(i, j, k - 1));
            t13(i, j, k) -= met_h_xi * tau11bar + met_h_eta * tau12bar;
            t13(i, j, k) *= -mu_tot;
            t31(i, j, k) *= -mu_tot * met_h_zeta;
        });
        Box q = tbxyz; q.lo[2] = q.hi[2]; tbxyz.hi[2] -= 1;
        loop(q, [&](int i, int j, int k) {
            double met_h_xi = met::h_xi_EdgeI(z, i, j, k, dI[0]), met_h_eta = met::h_eta_EdgeI(z, i, j, k, dI[1]);
            double met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
            double tau21lo = 0.5 * (t21(i, j, k - 2) + t21(i + 1, j, k - 2));
            double tau21hi = 0.5 * (t21(i, j, k - 1) + t21(i + 1, j, k - 1));
            double tau21bar = 1.5 * tau21hi - 0.5 * tau21lo;
            double tau22lo = 0.5 * (t22(i, j, k - 2) + t22(i, j - 1, k - 2));
            double tau22hi = 0.5 * (t22(i, j, k - 1) + t22(i, j - 1, k - 1));
            double tau22bar = 1.5 * tau22hi - 0.5 * tau22lo;
            double mu_tot = 0.25 * (rhoAlpha(i, j - 1, k) + rhoAlpha(i, j, k) + rhoAlpha(i, j - 1, k - 1) + rhoAlpha(i, j, k - 1));
            t23(i, j, k) -= met_h_xi * tau21bar + met_h_eta * tau22bar;
            t23(i, j, k) *= -mu_tot;
            t32(i, j, k) *= -mu_tot * met_h_zeta;
        });
    }
    // tau13, tau23 between
    loop(tbxxz, [&](int i, int j, int k) {
        double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dI[0]), met_h_eta = met::h_eta_EdgeJ(z, i, j, k, dI[1]);
        double met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dI[2]);
        double tau11bar = 0.25 * (t11(i, j, k) + t11(i - 1, j, k) + t11(i, j, k - 1) + t11(i - 1, j, k - 1));
        double tau12bar = 0.25 * (t12(i, j, k) + t12(i, j + 1, k) + t12(i, j, k - 1) + t12(i, j + 1, k - 1));
        double mu_tot = 0.25 * (rhoAlpha(i - 1, j, k) + rhoAlpha(i, j, k) + rhoAlpha(i - 1, j, k - 1) + rhoAlpha(i, j, k - 1));
        t13(i, j, k) -= met_h_xi * tau11bar + met_h_eta * tau12bar;
        t13(i, j, k) *= -mu_tot;
        t31(i, j, k) *= -mu_tot * met_h_zeta;
    });
    loop(tbxyz, [&](int i, int j, int k) {
        double met_h_xi = met::h_xi_EdgeI(z, i, j, k, dI[0]), met_h_eta = met::h_eta_EdgeI(z, i, j, k, dI[1]);
        double met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dI[2]);
        double tau21bar = 0.25 * (t21(i, j, k) + t21(i + 1, j, k) + t21(i, j, k - 1) + t21(i + 1, j, k - 1));
        double tau22bar = 0.25 * (t22(i, j, k) + t22(i, j - 1, k) + t22(i, j, k - 1) + t22(i, j - 1, k - 1));
        double mu_tot = 0.25 * (rhoAlpha(i, j - 1, k) + rhoAlpha(i, j, k) + rhoAlpha(i, j - 1, k - 1) + rhoAlpha(i, j, k - 1));
        t23(i, j, k) -= met_h_xi * tau21bar + met_h_eta * tau22bar;
        t23(i, j, k) *= -mu_tot;
        t32(i, j, k) *= -mu_tot * met_h_zeta;
    });
    // tau11, tau22, tau12 / tau21
    loop(bxcc, [&](int i, int j, int k) {
        double met_h_zeta = J(i, j, k);
        double mu_tot = rhoAlpha(i, j, k);
        t11(i, j, k) *= -mu_tot * met_h_zeta;
        t22(i, j, k) *= -mu_tot * met_h_zeta;
    });
    loop(tbxxy, [&](int i, int j, int k) {
        double met_h_zeta = met::h_zeta_EdgeK(z, i, j, k, dI[2]);
        double mu_tot = 0.25 * (rhoAlpha(i - 1, j, k) + rhoAlpha(i, j, k) + rhoAlpha(i - 1, j - 1, k) + rhoAlpha(i, j - 1, k));
        t12(i, j, k) *= -mu_tot * met_h_zeta;
        t21(i, j, k) *= -mu_tot * met_h_zeta;
    });
}

// erf_make_tau_terms, terrain branch (Source/TimeIntegration/ERF_make_tau_terms.cpp:177-323),
// static terrain, unit map factors, molecular diffusion only: Omega of the
// velocities, the expansion rate divided by detJ, ComputeStrain_T,
// ComputeStressConsVisc_T, and the copy into the persistent tau fields.
void Level::make_tau_terms_T(StateA data, Arr u, Arr v, Arr w) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const bool consA = cfg.molec_diff_type == ERFX_MOLEC_CONSTANT_ALPHA;
    const double mu_eff = consA ? 2.0 * cfg.dynamic_viscosity / cfg.rho0_trans : 2.0 * cfg.dynamic_viscosity;
    Arr z = A(z_nd), J = A(detJ);
    Box bxcc(-1, -1, 0, nx, ny, nz - 1);
    Box tbxxy(-1, -1, 0, nx + 1, ny + 1, nz - 1);
    Box tbxxz(-1, -1, 0, nx + 1, ny, nz);
    Box tbxyz(-1, -1, 0, nx, ny + 1, nz);
    Field fom, fer, f11, f22, f33, f12, f13, f23, f21, f31, f32;
    for (Field* f : {&fom, &fer, &f11, &f22, &f33, &f12, &f13, &f23, &f21, &f31, &f32}) f->alloc(L, 1);
    Arr om = fom.a(L), er = fer.a(L);
    Arr s11 = f11.a(L), s22 = f22.a(L), s33 = f33.a(L), s12 = f12.a(L), s13 = f13.a(L), s23 = f23.a(L);
    Arr s21 = f21.a(L), s31 = f31.a(L), s32 = f32.a(L);
    loop(Box(-1, -1, 0, nx, ny, nz), [&](int i, int j, int k) {
        om(i, j, k) = (k == 0) ? 0. : met::omega_from_w(i, j, k, w(i, j, k), u, v, z, dxInv);
    });
    loop(bxcc, [&](int i, int j, int k) {
        double met_u_h_zeta_hi = met::h_zeta_Iface(z, i + 1, j, k, dxInv[2]);
        double met_u_h_zeta_lo = met::h_zeta_Iface(z, i, j, k, dxInv[2]);
        double met_v_h_zeta_hi = met::h_zeta_Jface(z, i, j + 1, k, dxInv[2]);
        double met_v_h_zeta_lo = met::h_zeta_Jface(z, i, j, k, dxInv[2]);
        double Omega_hi = om(i, j, k + 1), Omega_lo = om(i, j, k);
        double mfsq = 1.0 * 1.0;
        double expansionRate = (u(i + 1, j, k) / 1.0 * met_u_h_zeta_hi - u(i, j, k) / 1.0 * met_u_h_zeta_lo) * dxInv[0] * mfsq +
                               (v(i, j + 1, k) / 1.0 * met_v_h_zeta_hi - v(i, j, k) / 1.0 * met_v_h_zeta_lo) * dxInv[1] * mfsq +
                               (Omega_hi - Omega_lo) * dxInv[2];
        er(i, j, k) = expansionRate / J(i, j, k);
    });
    compute_strain_T(bxcc, tbxxy, tbxxz, tbxyz, u, v, w, s11, s22, s33, s12, s13, s21, s23, s31, s32);
    // the edge boxes lose their lateral halo for the stress (:262-264)
    Box sxy(0, 0, 0, nx, ny, nz - 1), sxz(0, 0, 0, nx, ny - 1, nz), syz(0, 0, 0, nx - 1, ny, nz);
    compute_stress_cons_T(bxcc, sxy, sxz, syz, mu_eff, consA, data.cons, s11, s22, s33, s12, s13, s21, s23, s31,
                          s32, er);
    Arr t11 = A(tau11), t22 = A(tau22), t33 = A(tau33), t12 = A(tau12), t13 = A(tau13), t23 = A(tau23);
    Arr t21 = A(tau21), t31 = A(tau31), t32 = A(tau32);
    loop(bxcc, [&](int i, int j, int k) {
        t11(i, j, k) = s11(i, j, k); t22(i, j, k) = s22(i, j, k); t33(i, j, k) = s33(i, j, k);
    });
    loop(sxy, [&](int i, int j, int k) { t12(i, j, k) = s12(i, j, k); t21(i, j, k) = s21(i, j, k); });
    loop(sxz, [&](int i, int j, int k) { t13(i, j, k) = s13(i, j, k); t31(i, j, k) = s31(i, j, k); });
    loop(syz, [&](int i, int j, int k) { t23(i, j, k) = s23(i, j, k); t32(i, j, k) = s32(i, j, k); });
}

// DiffusionSrcForState_T (Source/Diffusion/ERF_DiffusionSrcForState_T.cpp:40-700),
// molecular diffusion only (ConstantAlpha / Constant), no MOST: face fluxes
// with the slope corrections, the z-flux corrected by the averaged x / y
// fluxes (extrapolated at the bottom and top faces), the x / y fluxes scaled by
// h_zeta, the divergence divided by detJ.  On the ConstantAlpha path the z-flux
// is divided by az (the k-face area, 1 on a static terrain grid, :523), on the
// Constant path by h_zeta at the k-face (:639).
void Level::diffusion_src_state_T(const Box& bx, int start_comp, int num_comp, Arr cd, Arr cp, Arr rhs) {
    const int nx = bx.hi[0], ny = bx.hi[1], nz = bx.hi[2];
    const bool consA = cfg.molec_diff_type == ERFX_MOLEC_CONSTANT_ALPHA;
    Arr z = A(z_nd), J = A(detJ), Ax = A(ax), Ay = A(ay), Az = A(az);
    double alpha_eff[8] = {0};
    const double aT = consA ? cfg.alpha_T : cfg.rho0_trans * cfg.alpha_T;
    const double aC = consA ? cfg.alpha_C : cfg.rho0_trans * cfg.alpha_C;
    alpha_eff[0] = aT;
    for (int p = 3; p < 8; ++p) alpha_eff[p] = aC;
    const double dx_inv = dxInv[0], dy_inv = dxInv[1], dz_inv = dxInv[2];
    const bool zlo_dir = bck[3][2] == EXT_DIR, zhi_dir = bck[3][5] == EXT_DIR;
    Field fx, fy, fz;
    fx.alloc(L, 1); fy.alloc(L, 1); fz.alloc(L, 1);
    Arr xflux = fx.a(L), yflux = fy.a(L), zflux = fz.a(L);
    for (int n = 0; n < num_comp; ++n) {
        const int qty = start_comp + n, prim_index = qty - 1;
        const double ae = alpha_eff[prim_index];
        loop(Box(0, 0, 0, nx + 1, ny, nz), [&](int i, int j, int k) {
            double rhoAlpha, met_h_zeta;
            if (consA) {
                double rhoFace = 0.5 * (cd(i, j, k, 0) + cd(i - 1, j, k, 0));
                rhoAlpha = rhoFace * ae;
                met_h_zeta = Ax(i, j, k);
            } else {
                rhoAlpha = ae;
                met_h_zeta = met::h_zeta_Iface(z, i, j, k, dz_inv);
            }
            double met_h_xi = met::h_xi_Iface(z, i, j, k, dx_inv);
            double GradCz = 0.25 * dz_inv * (cp(i, j, k + 1, prim_index) + cp(i - 1, j, k + 1, prim_index)
                                           - cp(i, j, k - 1, prim_index) - cp(i - 1, j, k - 1, prim_index));
            double GradCx = dx_inv * (cp(i, j, k, prim_index) - cp(i - 1, j, k, prim_index));
            xflux(i, j, k) = -rhoAlpha * 1.0 * (GradCx - (met_h_xi / met_h_zeta) * GradCz);
        });
        loop(Box(0, 0, 0, nx, ny + 1, nz), [&](int i, int j, int k) {
            double rhoAlpha, met_h_zeta;
            if (consA) {
                double rhoFace = 0.5 * (cd(i, j, k, 0) + cd(i, j - 1, k, 0));
                rhoAlpha = rhoFace * ae;
                met_h_zeta = Ay(i, j, k);
            } else {
                rhoAlpha = ae;
                met_h_zeta = met::h_zeta_Jface(z, i, j, k, dz_inv);
            }
            double met_h_eta = met::h_eta_Jface(z, i, j, k, dy_inv);
            double GradCz = 0.25 * dz_inv * (cp(i, j, k + 1, prim_index) + cp(i, j - 1, k + 1, prim_index)
                                           - cp(i, j, k - 1, prim_index) - cp(i, j - 1, k - 1, prim_index));
            double GradCy = dy_inv * (cp(i, j, k, prim_index) - cp(i, j - 1, k, prim_index));
            yflux(i, j, k) = -rhoAlpha * 1.0 * (GradCy - (met_h_eta / met_h_zeta) * GradCz);
        });
        loop(Box(0, 0, 0, nx, ny, nz + 1), [&](int i, int j, int k) {
            double rhoAlpha, met_h_zeta;
            if (consA) {
                double rhoFace = 0.5 * (cd(i, j, k, 0) + cd(i, j, k - 1, 0));
                rhoAlpha = rhoFace * ae;
                met_h_zeta = Az(i, j, k);
            } else {
                rhoAlpha = ae;
                met_h_zeta = met::h_zeta_Kface(z, i, j, k, dz_inv);
            }
            double GradCz;
            if (zlo_dir && k == 0)
                GradCz = dz_inv * (-(8. / 3.) * cp(i, j, k - 1, prim_index) + 3. * cp(i, j, k, prim_index)
                                   - (1. / 3.) * cp(i, j, k + 1, prim_index));
            else if (zhi_dir && k == nz + 1)
                GradCz = dz_inv * ((8. / 3.) * cp(i, j, k, prim_index) - 3. * cp(i, j, k - 1, prim_index)
                                   + (1. / 3.) * cp(i, j, k - 2, prim_index));
            else
                GradCz = dz_inv * (cp(i, j, k, prim_index) - cp(i, j, k - 1, prim_index));
            zflux(i, j, k) = -rhoAlpha * GradCz / met_h_zeta;
        });
        // the z-flux's slope terms: extrapolated at the bottom and top faces,
        // averaged between (:702-760)
        const int k_lo = 0, k_hi = nz + 1;
        loop(Box(0, 0, 0, nx, ny, 0), [&](int i, int j, int) {
            {
                double met_h_xi = met::h_xi_Kface(z, i, j, k_lo, dx_inv), met_h_eta = met::h_eta_Kface(z, i, j, k_lo, dy_inv);
                double xfluxlo = 0.5 * (xflux(i, j, k_lo) + xflux(i + 1, j, k_lo));
                double xfluxhi = 0.5 * (xflux(i, j, k_lo + 1) + xflux(i + 1, j, k_lo + 1));
                double xfluxbar = 1.5 * xfluxlo - 0.5 * xfluxhi;
                double yfluxlo = 0.5 * (yflux(i, j, k_lo) + yflux(i, j + 1, k_lo));
                double yfluxhi = 0.5 * (yflux(i, j, k_lo + 1) + yflux(i, j + 1, k_lo + 1));
                double yfluxbar = 1.5 * yfluxlo - 0.5 * yfluxhi;
                zflux(i, j, k_lo) -= met_h_xi * xfluxbar + met_h_eta * yfluxbar;
            }
            {
                double met_h_xi = met::h_xi_Kface(z, i, j, k_hi, dx_inv), met_h_eta = met::h_eta_Kface(z, i, j, k_hi, dy_inv);
                double xfluxlo = 0.5 * (xflux(i, j, k_hi - 2) + xflux(i + 1, j, k_hi - 2));
                double xfluxhi = 0.5 * (xflux(i, j, k_hi - 1) + xflux(i + 1, j, k_hi - 1));
                double xfluxbar = 1.5 * xfluxhi - 0.5 * xfluxlo;
                double yfluxlo = 0.5 * (yflux(i, j, k_hi - 2) + yflux(i, j + 1, k_hi - 2));
                double yfluxhi = 0.5 * (yflux(i, j, k_hi - 1) + yflux(i, j + 1, k_hi - 1));
                double yfluxbar = 1.5 * yfluxhi - 0.5 * yfluxlo;
                zflux(i, j, k_hi) -= met_h_xi * xfluxbar + met_h_eta * yfluxbar;
            }
        });
        loop(Box(0, 0, 1, nx, ny, nz), [&](int i, int j, int k) {
            double met_h_xi = met::h_xi_Kface(z, i, j, k, dx_inv), met_h_eta = met::h_eta_Kface(z, i, j, k, dy_inv);
            double xfluxbar = 0.25 * (xflux(i, j, k) + xflux(i + 1, j, k) + xflux(i, j, k - 1) + xflux(i + 1, j, k - 1));
            double yfluxbar = 0.25 * (yflux(i, j, k) + yflux(i, j + 1, k) + yflux(i, j, k - 1) + yflux(i, j + 1, k - 1));
            zflux(i, j, k) -= met_h_xi * xfluxbar + met_h_eta * yfluxbar;
        });
        loop(Box(0, 0, 0, nx + 1, ny, nz), [&](int i, int j, int k) { xflux(i, j, k) *= met::h_zeta_Iface(z, i, j, k, dz_inv); });
        loop(Box(0, 0, 0, nx, ny + 1, nz), [&](int i, int j, int k) { yflux(i, j, k) *= met::h_zeta_Jface(z, i, j, k, dz_inv); });
        loop(bx, [&](int i, int j, int k) {
            double stateContrib = (xflux(i + 1, j, k) - xflux(i, j, k)) * dx_inv * 1.0 +
                                  (yflux(i, j + 1, k) - yflux(i, j, k)) * dy_inv * 1.0 +
                                  (zflux(i, j, k + 1) - zflux(i, j, k)) * dz_inv;
            stateContrib /= J(i, j, k);
            rhs(i, j, k, qty) -= stateContrib;
        });
    }
}

}  // namespace orc
