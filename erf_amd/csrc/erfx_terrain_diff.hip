// erfx_terrain_diff.hip -- molecular diffusion on a terrain-following grid
// (static terrain, unit map factors): the strain and stress of
// erf_make_tau_terms' terrain branch and the scalar diffusion term of
// DiffusionSrcForState_T.  Every value is the reference's expression in its
// operation order (FP contraction is off), so each point carries the
// reference's rounding; the launch structure is this library's own: one
// kernel for all strain components, three for the stress, and two per
// diffused scalar (the lateral fluxes, then the corrected z-fluxes and the
// divergence of all three).
#include "erfx_internal.h"
#include "erfx_terrain.h"

namespace erfx {

// erf_make_tau_terms, terrain branch (Source/TimeIntegration/ERF_make_tau_terms.cpp:177-323)
// with ComputeStrain_T (Source/Diffusion/ERF_ComputeStrain_T.cpp:33-507) and
// ComputeStressConsVisc_T (Source/Diffusion/ERF_ComputeStress_T.cpp:29-252).
// The strain is formed on the grown tile boxes of the reference (cells
// (-1..nx, -1..ny, 0..nz-1), edges (-1..nx+1, -1..ny+1, 0..nz-1),
// (-1..nx+1, -1..ny, 0..nz), (-1..nx, -1..ny+1, 0..nz)); the grown boxes start
// outside the domain, so the lateral Dirichlet branches never apply there and
// only the z walls select one-sided forms.  The stress then works on the
// cells and on the valid edge boxes, in the reference's order: S - D on the
// cells, tau33 with the averaged tau31 / tau32, tau13 / tau31 and tau23 /
// tau32 (extrapolated averages at the bottom and top faces), and last tau11,
// tau22, tau12 and tau21.  The stresses are written straight into the
// persistent fields; the reference copies them there from local boxes, and
// the diffusion terms read only the copied regions.
void make_tau_terms_T(const Params& P, const SlowWork& W, const double* cons, const double* u, const double* v,
                      const double* w, hipStream_t s) {
    if (rec_refuse("diffusion on a terrain grid")) return;
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long sy = L.sy, sz = L.sz;
    const double dxi = P.dxInv[0], dyi = P.dxInv[1], dzi = P.dxInv[2];
    const double dI[3] = {dxi, dyi, dzi};
    const bool consA = P.consA;
    const double mu_eff = consA ? 2.0 * P.cfg.dynamic_viscosity / P.cfg.rho0_trans : 2.0 * P.cfg.dynamic_viscosity;
    const bool zl_u = P.bck[0][2] == EXT_DIR, zh_u = P.bck[0][5] == EXT_DIR;
    const bool zl_v = P.bck[1][2] == EXT_DIR, zh_v = P.bck[1][5] == EXT_DIR;
    const Zn z{P.z_nd, L};
    const double* J = P.detJ;
    const Fld fu{u, L}, fv{v, L};
    double *t11 = W.tau[0], *t22 = W.tau[1], *t33 = W.tau[2], *t12 = W.tau[3], *t13 = W.tau[4], *t23 = W.tau[5];
    double *t21 = W.tau_t[0], *t31 = W.tau_t[1], *t32 = W.tau_t[2];
    double* er = W.er;
    const double dx0 = dI[0], dx1 = dI[1], dx2 = dI[2];

    // expansion rate over detJ on the cells, with Omega of the velocities on
    // their two z-faces (0 at the bottom face, :201-231)
    launch_box(Box(-1, -1, 0, nx, ny, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        double met_u_h_zeta_hi = met::h_zeta_Iface(z, i + 1, j, k, dx2);
        double met_u_h_zeta_lo = met::h_zeta_Iface(z, i, j, k, dx2);
        double met_v_h_zeta_hi = met::h_zeta_Jface(z, i, j + 1, k, dx2);
        double met_v_h_zeta_lo = met::h_zeta_Jface(z, i, j, k, dx2);
        double Omega_hi = met::omega_from_w(i, j, k + 1, w[q + sz], fu, fv, z, dx0, dx1);
        double Omega_lo = (k == 0) ? 0. : met::omega_from_w(i, j, k, w[q], fu, fv, z, dx0, dx1);
        double mfsq = 1.0 * 1.0;
        double expansionRate = (u[q + 1] / 1.0 * met_u_h_zeta_hi - u[q] / 1.0 * met_u_h_zeta_lo) * dx0 * mfsq +
                               (v[q + sy] / 1.0 * met_v_h_zeta_hi - v[q] / 1.0 * met_v_h_zeta_lo) * dx1 * mfsq +
                               (Omega_hi - Omega_lo) * dx2;
        er[q] = expansionRate / J[q];
    });

    // strain on the grown boxes, all components in one launch
    launch_box(Box(-1, -1, 0, nx + 1, ny + 1, nz), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        const bool ho = zl_u && zl_v && k == 0;   // one-sided d/dz at a no-slip bottom (:312-352)
        if (i <= nx && j <= ny && k <= nz - 1) {
            double GradUz, GradVz;
            if (ho) {
                GradUz = 0.5 * dx2 * ((-(8. / 3.) * u[q - sz] + 3. * u[q] - (1. / 3.) * u[q + sz])
                                     + (-(8. / 3.) * u[q - 1 - sz] + 3. * u[q - 1] - (1. / 3.) * u[q - 1 + sz]));
                GradVz = 0.5 * dx2 * ((-(8. / 3.) * v[q - sz] + 3. * v[q] - (1. / 3.) * v[q + sz])
                                     + (-(8. / 3.) * v[q - sy - sz] + 3. * v[q - sy] - (1. / 3.) * v[q - sy + sz]));
            } else {
                GradUz = 0.25 * dx2 * (u[q + sz] / 1.0 + u[q - 1 + sz] / 1.0 - u[q - sz] / 1.0 - u[q - 1 - sz] / 1.0);
                GradVz = 0.25 * dx2 * (v[q + sz] / 1.0 + v[q - sy + sz] / 1.0 - v[q - sz] / 1.0 - v[q - sy - sz] / 1.0);
            }
            double met_h_xi = met::h_xi_Cell(z, i, j, k, dx0), met_h_eta = met::h_eta_Cell(z, i, j, k, dx1);
            double met_h_zeta = J[q];
            t11[q] = ((u[q + 1] / 1.0 - u[q] / 1.0) * dx0 - (met_h_xi / met_h_zeta) * GradUz) * 1.0 * 1.0;
            t22[q] = ((v[q + sy] / 1.0 - v[q] / 1.0) * dx1 - (met_h_eta / met_h_zeta) * GradVz) * 1.0 * 1.0;
            t33[q] = (w[q + sz] - w[q]) * dx2 / met_h_zeta;
        }
        if (k <= nz - 1) {                                   // xy edges
            double GradUz, GradVz;
            if (ho) {
                GradUz = 0.5 * dx2 * ((-(8. / 3.) * u[q - sz] + 3. * u[q] - (1. / 3.) * u[q + sz])
                                     + (-(8. / 3.) * u[q - sy - sz] + 3. * u[q - sy] - (1. / 3.) * u[q - sy + sz]));
                GradVz = 0.5 * dx2 * ((-(8. / 3.) * v[q - sz] + 3. * v[q] - (1. / 3.) * v[q + sz])
                                     + (-(8. / 3.) * v[q - 1 - sz] + 3. * v[q - 1] - (1. / 3.) * v[q - 1 + sz]));
            } else {
                GradUz = 0.25 * dx2 * (u[q + sz] / 1.0 + u[q - sy + sz] / 1.0 - u[q - sz] / 1.0 - u[q - sy - sz] / 1.0);
                GradVz = 0.25 * dx2 * (v[q + sz] / 1.0 + v[q - 1 + sz] / 1.0 - v[q - sz] / 1.0 - v[q - 1 - sz] / 1.0);
            }
            double met_h_xi = met::h_xi_EdgeK(z, i, j, k, dx0), met_h_eta = met::h_eta_EdgeK(z, i, j, k, dx1);
            double met_h_zeta = met::h_zeta_EdgeK(z, i, j, k, dx2);
            double s12 = 0.5 * ((u[q] / 1.0 - u[q - sy] / 1.0) * dx1 + (v[q] / 1.0 - v[q - 1] / 1.0) * dx0
                                - (met_h_eta / met_h_zeta) * GradUz - (met_h_xi / met_h_zeta) * GradVz) * 1.0 * 1.0;
            t12[q] = s12;
            t21[q] = s12;
        }
        if (j <= ny) {                                       // xz edges
            double met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dx2);
            double s13;
            if (k == nz) {                                   // top face (h_xi = h_eta = 0 there, :410-438)
                if (zh_u)
                    s13 = 0.5 * (-(-(8. / 3.) * u[q] + 3. * u[q - sz] - (1. / 3.) * u[q - 2 * sz]) * dx2 / met_h_zeta
                                 + (w[q] - w[q - 1]) * dx0 * 1.0);
                else
                    s13 = 0.5 * ((u[q] - u[q - sz]) * dx2 / met_h_zeta + (w[q] - w[q - 1]) * dx0 * 1.0);
            } else {
                double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dx0);
                double GradWz = (k == 0) ? 0.5 * dx2 * (w[q + sz] + w[q - 1 + sz] - w[q] - w[q - 1])
                                         : 0.25 * dx2 * (w[q + sz] + w[q - 1 + sz] - w[q - sz] - w[q - 1 - sz]);
                double dudz = (k == 0 && zl_u)
                                  ? (-(8. / 3.) * u[q - sz] + 3. * u[q] - (1. / 3.) * u[q + sz]) * dx2 / met_h_zeta
                                  : (u[q] - u[q - sz]) * dx2 / met_h_zeta;
                s13 = 0.5 * (dudz + ((w[q] - w[q - 1]) * dx0 - (met_h_xi / met_h_zeta) * GradWz) * 1.0);
            }
            t13[q] = s13;
            t31[q] = s13;
        }
        if (i <= nx) {                                       // yz edges
            double met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dx2);
            double s23;
            if (k == nz) {
                if (zh_v)
                    s23 = 0.5 * (-(-(8. / 3.) * v[q] + 3. * v[q - sz] - (1. / 3.) * v[q - 2 * sz]) * dx2 / met_h_zeta
                                 + (w[q] - w[q - sy]) * dx1 * 1.0);
                else
                    s23 = 0.5 * ((v[q] - v[q - sz]) * dx2 / met_h_zeta + (w[q] - w[q - sy]) * dx1 * 1.0);
            } else {
                double met_h_eta = met::h_eta_EdgeI(z, i, j, k, dx1);
                double GradWz = (k == 0) ? 0.5 * dx2 * (w[q + sz] + w[q - sy + sz] - w[q] - w[q - sy])
                                         : 0.25 * dx2 * (w[q + sz] + w[q - sy + sz] - w[q - sz] - w[q - sy - sz]);
                double dvdz = (k == 0 && zl_v)
                                  ? (-(8. / 3.) * v[q - sz] + 3. * v[q] - (1. / 3.) * v[q + sz]) * dx2 / met_h_zeta
                                  : (v[q] - v[q - sz]) * dx2 / met_h_zeta;
                s23 = 0.5 * (dvdz + ((w[q] - w[q - sy]) * dx1 - (met_h_eta / met_h_zeta) * GradWz) * 1.0);
            }
            t23[q] = s23;
            t32[q] = s23;
        }
    });

    // stress: S - D on the cells, and tau33 on the valid cells
    const double OneThird = (1. / 3.);
    launch_box(Box(-1, -1, 0, nx, ny, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        t11[q] -= OneThird * er[q];
        t22[q] -= OneThird * er[q];
        t33[q] -= OneThird * er[q];
        if (i < 0 || j < 0 || i > nx - 1 || j > ny - 1) return;
        double met_h_xi = met::h_xi_Cell(z, i, j, k, dx0), met_h_eta = met::h_eta_Cell(z, i, j, k, dx1);
        double tau31bar = 0.25 * (t31[q] + t31[q + 1] + t31[q + sz] + t31[q + 1 + sz]);
        double tau32bar = 0.25 * (t32[q] + t32[q + sy] + t32[q + sz] + t32[q + sy + sz]);
        double mu_tot = consA ? cons[q] * mu_eff : mu_eff;
        double t = t33[q];
        t -= met_h_xi * tau31bar + met_h_eta * tau32bar;
        t *= -mu_tot;
        t33[q] = t;
    });
    // tau13 / tau31 on the valid xz edges and tau23 / tau32 on the valid yz edges
    auto ra = [=] __device__(long q) { return consA ? cons[q] * mu_eff : mu_eff; };
    launch_box(Box(0, 0, 0, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (j <= ny - 1) {
            double met_h_xi = met::h_xi_EdgeJ(z, i, j, k, dx0), met_h_eta = met::h_eta_EdgeJ(z, i, j, k, dx1);
            double met_h_zeta = met::h_zeta_EdgeJ(z, i, j, k, dx2);
            double tau11bar, tau12bar;
            if (k == 0) {
                double tau11lo = 0.5 * (t11[q] + t11[q - 1]);
                double tau11hi = 0.5 * (t11[q + sz] + t11[q - 1 + sz]);
                tau11bar = 1.5 * tau11lo - 0.5 * tau11hi;
                double tau12lo = 0.5 * (t12[q] + t12[q + sy]);
                double tau12hi = 0.5 * (t12[q + sz] + t12[q + sy + sz]);
                tau12bar = 1.5 * tau12lo - 0.5 * tau12hi;
            } else if (k == nz) {
                double tau11lo = 0.5 * (t11[q - 2 * sz] + t11[q - 1 - 2 * sz]);
                double tau11hi = 0.5 * (t11[q - sz] + t11[q - 1 - sz]);
                tau11bar = 1.5 * tau11hi - 0.5 * tau11lo;
                double tau12lo = 0.5 * (t12[q - 2 * sz] + t12[q + sy - 2 * sz]);
                double tau12hi = 0.5 * (t12[q - sz] + t12[q + sy - sz]);
                tau12bar = 1.5 * tau12hi - 0.5 * tau12lo;
            } else {
                tau11bar = 0.25 * (t11[q] + t11[q - 1] + t11[q - sz] + t11[q - 1 - sz]);
                tau12bar = 0.25 * (t12[q] + t12[q + sy] + t12[q - sz] + t12[q + sy - sz]);
            }
            double mu_tot = 0.25 * (ra(q - 1) + ra(q) + ra(q - 1 - sz) + ra(q - sz));
            double t = t13[q];
            t -= met_h_xi * tau11bar + met_h_eta * tau12bar;
            t *= -mu_tot;
            t13[q] = t;
            t31[q] *= -mu_tot * met_h_zeta;
        }
        if (i <= nx - 1) {
            double met_h_xi = met::h_xi_EdgeI(z, i, j, k, dx0), met_h_eta = met::h_eta_EdgeI(z, i, j, k, dx1);
            double met_h_zeta = met::h_zeta_EdgeI(z, i, j, k, dx2);
            double tau21bar, tau22bar;
            if (k == 0) {
                double tau21lo = 0.5 * (t21[q] + t21[q + 1]);
                double tau21hi = 0.5 * (t21[q + sz] + t21[q + 1 + sz]);
                tau21bar = 1.5 * tau21lo - 0.5 * tau21hi;
                double tau22lo = 0.5 * (t22[q] + t22[q - sy]);
                double tau22hi = 0.5 * (t22[q + sz] + t22[q - sy + sz]);
                tau22bar = 1.5 * tau22lo - 0.5 * tau22hi;
            } else if (k == nz) {
                double tau21lo = 0.5 * (t21[q - 2 * sz] + t21[q + 1 - 2 * sz]);
                double tau21hi = 0.5 * (t21[q - sz] + t21[q + 1 - sz]);
                tau21bar = 1.5 * tau21hi - 0.5 * tau21lo;
                double tau22lo = 0.5 * (t22[q - 2 * sz] + t22[q - sy - 2 * sz]);
                double tau22hi = 0.5 * (t22[q - sz] + t22[q - sy - sz]);
                tau22bar = 1.5 * tau22hi - 0.5 * tau22lo;
            } else {
                tau21bar = 0.25 * (t21[q] + t21[q + 1] + t21[q - sz] + t21[q + 1 - sz]);
                tau22bar = 0.25 * (t22[q] + t22[q - sy] + t22[q - sz] + t22[q - sy - sz]);
            }
            double mu_tot = 0.25 * (ra(q - sy) + ra(q) + ra(q - sy - sz) + ra(q - sz));
            double t = t23[q];
            t -= met_h_xi * tau21bar + met_h_eta * tau22bar;
            t *= -mu_tot;
            t23[q] = t;
            t32[q] *= -mu_tot * met_h_zeta;
        }
    });
    // tau11, tau22 on the cells (-1..nx, -1..ny); tau12, tau21 on the valid xy
    // edges (0..nx, 0..ny), the same box without its lo halo
    launch_box(Box(-1, -1, 0, nx, ny, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        double met_h_zeta = J[q];
        double mu_tot = consA ? cons[q] * mu_eff : mu_eff;
        t11[q] *= -mu_tot * met_h_zeta;
        t22[q] *= -mu_tot * met_h_zeta;
        if (i < 0 || j < 0) return;
        double hz = met::h_zeta_EdgeK(z, i, j, k, dx2);
        double mu12 = 0.25 * (ra(q - 1) + ra(q) + ra(q - 1 - sy) + ra(q - sy));
        t12[q] *= -mu12 * hz;
        t21[q] *= -mu12 * hz;
    });
}

// DiffusionSrcForState_T (Source/Diffusion/ERF_DiffusionSrcForState_T.cpp:40-700),
// molecular diffusion (ConstantAlpha / Constant), no MOST: the diffusion term
// ((flux differences) / detJ, the amount cell_rhs loses) of cons components
// n0..n0+nc-1 into W.dstate.  Per component: the x- and y-fluxes with their
// slope terms, then per cell the z-fluxes of its two faces with the averaged
// (extrapolated at the bottom and top faces) lateral fluxes subtracted, the
// lateral fluxes scaled by h_zeta, and the divergence.  The z-flux divides by
// az (1 on a static terrain grid) on the ConstantAlpha path (:523) and by
// h_zeta at the k-face on the Constant path (:639).
void diffusion_state_T(const Params& P, const SlowWork& W, int n0, int nc, const double* cons, const double* prim,
                       hipStream_t s) {
    if (rec_refuse("diffusion on a terrain grid")) return;
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const double dx_inv = P.dxInv[0], dy_inv = P.dxInv[1], dz_inv = P.dxInv[2];
    const bool consA = P.consA;
    const Zn z{P.z_nd, L};
    const double* J = P.detJ;
    const double* Ax = P.ax;
    const double* Ay = P.ay;
    const double* Az = P.az;
    const bool zlo_dir = P.bck[3][2] == EXT_DIR, zhi_dir = P.bck[3][5] == EXT_DIR;
    double* xflux = W.dfx;
    double* yflux = W.dfy;
    for (int n = n0; n < n0 + nc; ++n) {
        const int prim_index = n - 1;
        const double ae = (prim_index == 0) ? (consA ? P.cfg.alpha_T : P.cfg.rho0_trans * P.cfg.alpha_T)
                                            : (consA ? P.cfg.alpha_C : P.cfg.rho0_trans * P.cfg.alpha_C);
        const double* cp = prim + (long)prim_index * ne;
        double* out = W.dstate + (long)n * ne;
        launch_box(Box(0, 0, 0, nx, ny, nz - 1), s, [=] __device__(int i, int j, int k) {
            const long q = L.ix(i, j, k);
            if (j <= ny - 1) {
                double rhoAlpha, met_h_zeta;
                if (consA) {
                    double rhoFace = 0.5 * (cons[q] + cons[q - 1]);
                    rhoAlpha = rhoFace * ae;
                    met_h_zeta = Ax[q];
                } else {
                    rhoAlpha = ae;
                    met_h_zeta = met::h_zeta_Iface(z, i, j, k, dz_inv);
                }
                double met_h_xi = met::h_xi_Iface(z, i, j, k, dx_inv);
                double GradCz = 0.25 * dz_inv * (cp[q + sz] + cp[q - 1 + sz] - cp[q - sz] - cp[q - 1 - sz]);
                double GradCx = dx_inv * (cp[q] - cp[q - 1]);
                xflux[q] = -rhoAlpha * 1.0 * (GradCx - (met_h_xi / met_h_zeta) * GradCz);
            }
            if (i <= nx - 1) {
                double rhoAlpha, met_h_zeta;
                if (consA) {
                    double rhoFace = 0.5 * (cons[q] + cons[q - sy]);
                    rhoAlpha = rhoFace * ae;
                    met_h_zeta = Ay[q];
                } else {
                    rhoAlpha = ae;
                    met_h_zeta = met::h_zeta_Jface(z, i, j, k, dz_inv);
                }
                double met_h_eta = met::h_eta_Jface(z, i, j, k, dy_inv);
                double GradCz = 0.25 * dz_inv * (cp[q + sz] + cp[q - sy + sz] - cp[q - sz] - cp[q - sy - sz]);
                double GradCy = dy_inv * (cp[q] - cp[q - sy]);
                yflux[q] = -rhoAlpha * 1.0 * (GradCy - (met_h_eta / met_h_zeta) * GradCz);
            }
        });
        // the corrected z-flux of face kf of column (i, j)
        auto zflux = [=] __device__(int i, int j, int kf) {
            const long q = L.ix(i, j, kf);
            double rhoAlpha, met_h_zeta;
            if (consA) {
                double rhoFace = 0.5 * (cons[q] + cons[q - sz]);
                rhoAlpha = rhoFace * ae;
                met_h_zeta = Az[q];
            } else {
                rhoAlpha = ae;
                met_h_zeta = met::h_zeta_Kface(z, i, j, kf, dz_inv);
            }
            double GradCz;
            if (zlo_dir && kf == 0)
                GradCz = dz_inv * (-(8. / 3.) * cp[q - sz] + 3. * cp[q] - (1. / 3.) * cp[q + sz]);
            else if (zhi_dir && kf == nz)
                GradCz = dz_inv * ((8. / 3.) * cp[q] - 3. * cp[q - sz] + (1. / 3.) * cp[q - 2 * sz]);
            else
                GradCz = dz_inv * (cp[q] - cp[q - sz]);
            double f = -rhoAlpha * GradCz / met_h_zeta;
            double met_h_xi = met::h_xi_Kface(z, i, j, kf, dx_inv), met_h_eta = met::h_eta_Kface(z, i, j, kf, dy_inv);
            double xfluxbar, yfluxbar;
            if (kf == 0) {
                double xfluxlo = 0.5 * (xflux[q] + xflux[q + 1]);
                double xfluxhi = 0.5 * (xflux[q + sz] + xflux[q + 1 + sz]);
                xfluxbar = 1.5 * xfluxlo - 0.5 * xfluxhi;
                double yfluxlo = 0.5 * (yflux[q] + yflux[q + sy]);
                double yfluxhi = 0.5 * (yflux[q + sz] + yflux[q + sy + sz]);
                yfluxbar = 1.5 * yfluxlo - 0.5 * yfluxhi;
            } else if (kf == nz) {
                double xfluxlo = 0.5 * (xflux[q - 2 * sz] + xflux[q + 1 - 2 * sz]);
                double xfluxhi = 0.5 * (xflux[q - sz] + xflux[q + 1 - sz]);
                xfluxbar = 1.5 * xfluxhi - 0.5 * xfluxlo;
                double yfluxlo = 0.5 * (yflux[q - 2 * sz] + yflux[q + sy - 2 * sz]);
                double yfluxhi = 0.5 * (yflux[q - sz] + yflux[q + sy - sz]);
                yfluxbar = 1.5 * yfluxhi - 0.5 * yfluxlo;
            } else {
                xfluxbar = 0.25 * (xflux[q] + xflux[q + 1] + xflux[q - sz] + xflux[q + 1 - sz]);
                yfluxbar = 0.25 * (yflux[q] + yflux[q + sy] + yflux[q - sz] + yflux[q + sy - sz]);
            }
            f -= met_h_xi * xfluxbar + met_h_eta * yfluxbar;
            return f;
        };
        launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
            const long q = L.ix(i, j, k);
            double xf0 = xflux[q] * met::h_zeta_Iface(z, i, j, k, dz_inv);
            double xf1 = xflux[q + 1] * met::h_zeta_Iface(z, i + 1, j, k, dz_inv);
            double yf0 = yflux[q] * met::h_zeta_Jface(z, i, j, k, dz_inv);
            double yf1 = yflux[q + sy] * met::h_zeta_Jface(z, i, j + 1, k, dz_inv);
            double zf0 = zflux(i, j, k), zf1 = zflux(i, j, k + 1);
            double stateContrib = (xf1 - xf0) * dx_inv * 1.0 + (yf1 - yf0) * dy_inv * 1.0 + (zf1 - zf0) * dz_inv;
            stateContrib /= J[q];
            out[q] = stateContrib;
        });
    }
}

}  // namespace erfx
