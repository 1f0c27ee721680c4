// init.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  Initial states of the
// reference's regression cases, restated so that the oracle can be checked
// against the reference's gold plotfiles.
#include "oracle.hpp"

namespace orc {

// HSEutils::init_isentropic_hse (Source/Utils/ERF_HSE_utils.H:88-181), flat
// grid with klo = 0: Newton iteration per level for a density that satisfies
// the EOS at constant theta and the discrete hydrostatic balance.
static void isentropic_hse(double r_sfc, double theta, std::vector<double>& r, std::vector<double>& p,
                           double dz, int khi) {
    const int MAX_ITER = 10;           // ERF_HSE_utils.H:13-14
    const double TOL = 1.e-8;
    const double half_dz = 0.5 * dz;
    r[0] = r_sfc;
    p[0] = p_0 - half_dz * r[0] * CONST_GRAV;
    for (int iter = 0; iter < MAX_ITER; iter++) {
        double p_hse = p_0 - half_dz * r[0] * CONST_GRAV;
        double p_eos = getPgivenRTh(r[0] * theta);
        double A = p_hse - p_eos;
        double dpdr = getdPdRgivenConstantTheta(r[0], theta);
        double drho = A / (dpdr + half_dz * CONST_GRAV);
        r[0] = r[0] + drho;
        p[0] = getPgivenRTh(r[0] * theta);
        if (std::abs(drho) < TOL) break;
    }
    for (int k = 1; k <= khi; k++) {
        r[k] = r[k - 1];
        for (int iter = 0; iter < MAX_ITER; iter++) {
            double r_avg = 0.5 * (r[k - 1] + r[k]);
            double p_hse = p[k - 1] - dz * r_avg * CONST_GRAV;
            double p_eos = getPgivenRTh(r[k] * theta);
            double A = p_hse - p_eos;
            double dpdr = getdPdRgivenConstantTheta(r[k], theta);
            double drho = A / (dpdr + dz * CONST_GRAV);
            r[k] = r[k] + drho;
            p[k] = getPgivenRTh(r[k] * theta);
            if (std::abs(drho) < TOL * r[k - 1]) break;
        }
    }
    r[khi + 1] = r[khi];
}

// Exec/RegTests/DensityCurrent: ERF::init_only (Source/ERF.cpp:1258-1346) with
// the default init type: initHSE (erf_init_dens_hse of
// Source/Prob/ERF_init_density_hse_dry.H:11-120, T_sfc = 300 K, and
// erf_enforce_hse, Source/Initialization/ERF_init1d.cpp:153-313), init_from_hse
// (ERF_init_from_hse.cpp:30-50) and the problem's perturbation
// (Exec/RegTests/DensityCurrent/ERF_prob.cpp:34-160): a cold cos^2 bubble
// entering rho only; velocities zero (U_0 = 0).
// HSEutils::init_isentropic_hse_terrain (Source/Utils/ERF_HSE_utils.H:199-298),
// klo = 0: the same Newton iteration on one column of cell-centre heights
// z_cc (the bottom cell's half height is z_cc(0), the spacing between levels
// z_cc(k) - z_cc(k-1)).
static void isentropic_hse_terrain(double r_sfc, double theta, std::vector<double>& r, std::vector<double>& p,
                                   const std::vector<double>& zc, int khi) {
    const int MAX_ITER = 10;
    const double TOL = 1.e-8;
    const double half_dz = zc[0];
    r[0] = r_sfc;
    p[0] = p_0 - half_dz * r[0] * CONST_GRAV;
    for (int iter = 0; iter < MAX_ITER; iter++) {
        double p_hse = p_0 - half_dz * r[0] * CONST_GRAV;
        double p_eos = getPgivenRTh(r[0] * theta);
        double A = p_hse - p_eos;
        double dpdr = getdPdRgivenConstantTheta(r[0], theta);
        double drho = A / (dpdr + half_dz * CONST_GRAV);
        r[0] = r[0] + drho;
        p[0] = getPgivenRTh(r[0] * theta);
        if (std::abs(drho) < TOL) break;
    }
    for (int k = 1; k <= khi; k++) {
        double dz_loc = (zc[k] - zc[k - 1]);
        r[k] = r[k - 1];
        for (int iter = 0; iter < MAX_ITER; iter++) {
            double p_hse = p[k - 1] - dz_loc * 0.5 * (r[k - 1] + r[k]) * CONST_GRAV;
            double p_eos = getPgivenRTh(r[k] * theta);
            double A = p_hse - p_eos;
            double dpdr = getdPdRgivenConstantTheta(r[k], theta);
            double drho = A / (dpdr + dz_loc * CONST_GRAV);
            r[k] = r[k] + drho;
            p[k] = getPgivenRTh(r[k] * theta);
            if (std::abs(drho) < TOL * r[k - 1]) break;
        }
    }
}

// The same case on a terrain-following grid (DensityCurrent_detJ2): the
// terrain branches of erf_init_dens_hse (ERF_init_density_hse_dry.H:29-74) and
// erf_enforce_hse (ERF_init1d.cpp:190-303) per column of the grid's z_cc, the
// lateral ghost columns of p, pi copied from the edge (:248-300), and the
// bubble placed at z_cc (Exec/RegTests/DensityCurrent/ERF_prob.cpp:60-95).
void Level::init_density_current_T(double x_c, double z_c, double x_r, double z_r, double T_pert) {
    const int nx = L.nx, ny = L.ny, nz = L.nz, ng = L.ng;
    const double dx = cfg.dx;
    const double T_sfc = 300.;
    const double rho_sfc = p_0 / (R_d * T_sfc);
    const double Thetabar = T_sfc;
    const double g = gravity;
    Arr r0a = A(r0), p0a = A(p0), pi0a = A(pi0), c = A(cons), zcc = A(z_cc);
    std::vector<double> r(nz + 1), p(nz + 1), zc(nz + 1);
    for (int j = -1; j <= ny; ++j)
        for (int i = -1; i <= nx; ++i) {
            for (int k = 0; k <= nz; ++k) zc[k] = zcc(i, j, k);
            isentropic_hse_terrain(rho_sfc, Thetabar, r, p, zc, nz - 1);
            for (int k = 0; k <= nz - 1; ++k) r0a(i, j, k) = r[k];
            r0a(i, j, -1) = r0a(i, j, 0);
            r0a(i, j, nz) = r0a(i, j, nz - 1);
            // erf_enforce_hse, klo = 0, khi = nz
            double hz = zcc(i, j, 0);
            p0a(i, j, 0) = p_0 - hz * r0a(i, j, 0) * g;
            pi0a(i, j, 0) = getExnergivenP(p0a(i, j, 0), rdOcp);
            p0a(i, j, -1) = p_0 + hz * r0a(i, j, 0) * g;
            pi0a(i, j, -1) = getExnergivenP(p0a(i, j, -1), rdOcp);
            for (int k = 1; k <= nz; ++k) {
                double dz_loc = (zcc(i, j, k) - zcc(i, j, k - 1));
                double dens_interp = 0.5 * (r0a(i, j, k) + r0a(i, j, k - 1));
                p0a(i, j, k) = p0a(i, j, k - 1) - dz_loc * dens_interp * g;
                pi0a(i, j, k) = getExnergivenP(p0a(i, j, k), rdOcp);
            }
        }
    for (int k = 0; k <= nz; ++k) {
        for (int j = -1; j <= ny; ++j) {
            p0a(-1, j, k) = p0a(0, j, k); pi0a(-1, j, k) = pi0a(0, j, k);
            p0a(nx, j, k) = p0a(nx - 1, j, k); pi0a(nx, j, k) = pi0a(nx - 1, j, k);
        }
        for (int i = -1; i <= nx; ++i) {
            p0a(i, -1, k) = p0a(i, 0, k); pi0a(i, -1, k) = pi0a(i, 0, k);
            p0a(i, ny, k) = p0a(i, ny - 1, k); pi0a(i, ny, k) = pi0a(i, ny - 1, k);
        }
    }
    // the remaining ghost columns as their nearest filled column (periodic
    // images in ERF; the case is laterally uniform)
    for (int k = -1; k <= nz; ++k)
        for (int j = -ng; j <= ny + ng; ++j)
            for (int i = -ng; i <= nx + ng; ++i) {
                const int ii = std::min(std::max(i, -1), nx), jj = std::min(std::max(j, -1), ny);
                if (ii == i && jj == j) continue;
                r0a(i, j, k) = r0a(ii, jj, k); p0a(i, j, k) = p0a(ii, jj, k); pi0a(i, j, k) = pi0a(ii, jj, k);
            }
    std::fill(cons.v.begin(), cons.v.end(), 0.0);
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j)
            for (int i = -1; i <= nx; ++i) {
                c(i, j, k, 0) = r0a(i, j, k);
                c(i, j, k, 1) = getRhoThetagivenP(p0a(i, j, k));
            }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double x = cfg.prob_lo[0] + (i + 0.5) * dx;
                const double z = zcc(i, j, k);
                double Lr = std::sqrt(std::pow((x - x_c) / x_r, 2) + std::pow((z - z_c) / z_r, 2));
                double pert = 0.0;
                if (Lr <= 1.0) {
                    double dT = T_pert * (std::cos(PI * Lr) + 1.0) / 2.0;
                    double Tbar_hse = p0a(i, j, k) / (R_d * r0a(i, j, k));
                    double theta_perturbed = (Tbar_hse + dT) * std::pow(p_0 / p0a(i, j, k), rdOcp);
                    pert = getRhoThetagivenP(p0a(i, j, k)) / theta_perturbed - r0a(i, j, k);
                }
                c(i, j, k, 0) += pert;
            }
    std::fill(xvel.v.begin(), xvel.v.end(), 0.0);
    std::fill(yvel.v.begin(), yvel.v.end(), 0.0);
    std::fill(zvel.v.begin(), zvel.v.end(), 0.0);
}

void Level::init_density_current(double x_c, double z_c, double x_r, double z_r, double T_pert) {
    if (terrain) { init_density_current_T(x_c, z_c, x_r, z_r, T_pert); return; }
    const int nx = L.nx, ny = L.ny, nz = L.nz, ng = L.ng;
    const double dz = cfg.dz, dx = cfg.dx;
    // base density
    const double T_sfc = 300.;
    const double rho_sfc = p_0 / (R_d * T_sfc);
    const double Thetabar = T_sfc;
    std::vector<double> r(nz + 2), p(nz + 2);
    isentropic_hse(rho_sfc, Thetabar, r, p, dz, nz - 1);
    // rho_hse on cells k = -1..nz: r[max(k, 0)]
    std::vector<double> rh(nz + 2), ph(nz + 2), pih(nz + 2);
    for (int k = -1; k <= nz; ++k) rh[k + 1] = r[std::max(k, 0)];
    // erf_enforce_hse
    const double g = gravity;
    const double hz = 0.5 * dz;
    ph[1] = p_0 - hz * rh[1] * g;
    pih[1] = getExnergivenP(ph[1], rdOcp);
    ph[0] = p_0 + hz * rh[1] * g;
    pih[0] = getExnergivenP(ph[0], rdOcp);
    for (int k = 1; k <= nz; k++) {
        double dens_interp = 0.5 * (rh[k + 1] + rh[k]);
        ph[k + 1] = ph[k] - dz * dens_interp * g;
        pih[k + 1] = getExnergivenP(ph[k + 1], rdOcp);
    }
    Arr r0a = A(r0), p0a = A(p0), pi0a = A(pi0), c = A(cons);
    for (int k = -1; k <= nz; ++k)
        for (int j = -ng; j <= ny + ng; ++j)
            for (int i = -ng; i <= nx + ng; ++i) {
                r0a(i, j, k) = rh[k + 1];
                p0a(i, j, k) = ph[k + 1];
                pi0a(i, j, k) = pih[k + 1];
            }
    // init_from_hse on the box grown by one; zero elsewhere
    std::fill(cons.v.begin(), cons.v.end(), 0.0);
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j)
            for (int i = -1; i <= nx; ++i) {
                c(i, j, k, 0) = r0a(i, j, k);
                c(i, j, k, 1) = getRhoThetagivenP(p0a(i, j, k));
            }
    // perturbation of rho on the valid cells
    const double rdOcp_l = rdOcp;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double x = cfg.prob_lo[0] + (i + 0.5) * dx;
                const double z = cfg.prob_lo[2] + (k + 0.5) * dz;
                double Lr = std::sqrt(std::pow((x - x_c) / x_r, 2) + std::pow((z - z_c) / z_r, 2));
                double pert = 0.0;
                if (Lr <= 1.0) {
                    double dT = T_pert * (std::cos(PI * Lr) + 1.0) / 2.0;
                    double Tbar_hse = p0a(i, j, k) / (R_d * r0a(i, j, k));
                    double theta_perturbed = (Tbar_hse + dT) * std::pow(p_0 / p0a(i, j, k), rdOcp_l);
                    pert = getRhoThetagivenP(p0a(i, j, k)) / theta_perturbed - r0a(i, j, k);
                }
                c(i, j, k, 0) += pert;
            }
    std::fill(xvel.v.begin(), xvel.v.end(), 0.0);
    std::fill(yvel.v.begin(), yvel.v.end(), 0.0);
    std::fill(zvel.v.begin(), zvel.v.end(), 0.0);
}

}  // namespace orc

namespace orc {

// Constant-density base state without gravity (Source/Prob/ERF_init_constant_density_hse.H
// with erf_enforce_hse at g = 0): r0 = rho_0, p0 = p_0, pi0 = Exner(p_0).
static void constant_base_state(Level& Lv, double rho_0) {
    const erfx_layout& L = Lv.L;
    Arr r0a = Lv.A(Lv.r0), p0a = Lv.A(Lv.p0), pi0a = Lv.A(Lv.pi0);
    const double g = Lv.gravity;
    const double hz = 0.5 * Lv.cfg.dz;
    for (int k = -1; k <= L.nz; ++k)
        for (int j = -L.ng; j <= L.ny + L.ng; ++j)
            for (int i = -L.ng; i <= L.nx + L.ng; ++i) {
                // the hydrostatic recurrence of erf_enforce_hse with rho = rho_0
                double p;
                if (k == -1) p = p_0 + hz * rho_0 * g;
                else {
                    p = p_0 - hz * rho_0 * g;
                    for (int kk = 1; kk <= k; ++kk) p = p - Lv.cfg.dz * (0.5 * (rho_0 + rho_0)) * g;
                }
                r0a(i, j, k) = rho_0;
                p0a(i, j, k) = p;
                pi0a(i, j, k) = getExnergivenP(p, Lv.rdOcp);
            }
}

// Exec/RegTests/ScalarAdvDiff with init_type = uniform: init_uniform
// (ERF_prob_common.H:450-463: rho = rho_0, rho*theta = rho_0 T_0 on the box grown
// by one), constant-density base state, and the perturbation of ERF_prob.cpp:54-150:
// the scalar of prob_type 10 (A_0 exp(-10 r^2) + B_0 sin x), 11 (cos^2 disc in
// x-y), 12 (cos^2 disc in x-z), times rho_0; x-velocity u_0 plus the log profile
// uRef ln((z+z0)/z0)/ln((zRef+z0)/z0), y-velocity v_0, z-velocity w_0.
void Level::init_scalar_advdiff(const double* pr) {
    const double rho_0 = pr[0], T_0 = pr[1], A_0 = pr[2], B_0 = pr[3], u_0 = pr[4], v_0 = pr[5], w_0 = pr[6];
    const double rad_0 = pr[7], z0 = pr[8], zRef = pr[9], uRef = pr[10];
    const double xc_frac = pr[11], yc_frac = pr[12], zc_frac = pr[13];
    const int prob_type = (int)pr[14];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    constant_base_state(*this, rho_0);
    std::fill(cons.v.begin(), cons.v.end(), 0.0);
    Arr c = A(cons);
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j)
            for (int i = -1; i <= nx; ++i) {
                c(i, j, k, 0) = rho_0;
                c(i, j, k, 1) = rho_0 * T_0;
            }
    const double* lo = cfg.prob_lo;
    const double hi[3] = {lo[0] + nx * cfg.dx, lo[1] + ny * cfg.dy, lo[2] + nz * cfg.dz};
    const double dxv[3] = {cfg.dx, cfg.dy, cfg.dz};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double x = lo[0] + (i + 0.5) * dxv[0];
                const double y = lo[1] + (j + 0.5) * dxv[1];
                const double z = lo[2] + (k + 0.5) * dxv[2];
                const double xc = xc_frac * (lo[0] + hi[0]);
                const double yc = yc_frac * (lo[1] + hi[1]);
                const double zc = zc_frac * (lo[2] + hi[2]);
                const double r0 = rad_0 * (hi[0] - lo[0]);
                const double r3d = std::sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc) + (z - zc) * (z - zc));
                const double r2d_xy = std::sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc));
                const double r2d_xz = std::sqrt((x - xc) * (x - xc) + (z - zc) * (z - zc));
                double sc;
                if (prob_type == 10) sc = A_0 * std::exp(-10. * r3d * r3d) + B_0 * std::sin(x);
                else if (prob_type == 11) sc = A_0 * 0.25 * (1.0 + std::cos(PI * std::min(r2d_xy, r0) / r0));
                else if (prob_type == 12) sc = A_0 * 0.25 * (1.0 + std::cos(PI * std::min(r2d_xz, r0) / r0));
                else if (prob_type == 13) {
                    const double r0_z = rad_0 * (hi[2] - lo[2]);
                    const double r2d_xz_ell = std::sqrt((x - xc) * (x - xc) / (r0 * r0) + (z - zc) * (z - zc) / (r0_z * r0_z));
                    sc = A_0 * 0.25 * (1.0 + std::cos(PI * std::min(r2d_xz_ell, r0_z) / r0_z));
                } else if (prob_type == 14) sc = std::cos(PI * x);
                else sc = (r3d < r0) ? A_0 : 0.0;   // a ball of radius r0 (any other prob_type)
                sc *= rho_0;
                c(i, j, k, ERFX_RHOSCALAR) += sc;
            }
    Arr u = A(xvel), v = A(yvel), w = A(zvel);
    std::fill(xvel.v.begin(), xvel.v.end(), 0.0);
    std::fill(yvel.v.begin(), yvel.v.end(), 0.0);
    std::fill(zvel.v.begin(), zvel.v.end(), 0.0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const double z = lo[2] + (k + 0.5) * dxv[2];
                if (uRef != 0.0)
                    u(i, j, k) = u_0 + uRef * std::log((z + z0) / z0) / std::log((zRef + z0) / z0);
                else
                    u(i, j, k) = u_0;
            }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) v(i, j, k) = v_0;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) w(i, j, k) = w_0;
}

// The background state of ERF::init_only (Source/ERF.cpp:1316-1338) on a
// constant-density base state without gravity: init_type uniform sets rho_0
// and rho_0 T_0 (init_uniform, ERF_prob_common.H:450-463), the default type
// the base state itself (init_from_hse); cells grown by one; every other
// field zero.
static void constant_background(Level& Lv, double rho_0, double T_0, bool uniform) {
    const erfx_layout& L = Lv.L;
    constant_base_state(Lv, rho_0);
    std::fill(Lv.cons.v.begin(), Lv.cons.v.end(), 0.0);
    std::fill(Lv.xvel.v.begin(), Lv.xvel.v.end(), 0.0);
    std::fill(Lv.yvel.v.begin(), Lv.yvel.v.end(), 0.0);
    std::fill(Lv.zvel.v.begin(), Lv.zvel.v.end(), 0.0);
    Arr c = Lv.A(Lv.cons), p0a = Lv.A(Lv.p0), r0a = Lv.A(Lv.r0);
    for (int k = -1; k <= L.nz; ++k)
        for (int j = -1; j <= L.ny; ++j)
            for (int i = -1; i <= L.nx; ++i) {
                c(i, j, k, 0) = uniform ? rho_0 : r0a(i, j, k);
                c(i, j, k, 1) = uniform ? rho_0 * T_0 : getRhoThetagivenP(p0a(i, j, k));
            }
}

// Exec/RegTests/Couette_Poiseuille (ERF_prob.cpp:29-140): prob_type 1 (Couette)
// u = u_0 z / prob_hi_z, v = v_0 z / prob_hi_z with z = (k + 1/2) dz, w = w_0;
// prob_type 10 / 11 (Poiseuille) u (resp. v) = 1 - z_h^2 with z_h the cell
// height from prob_lo; on the valid face boxes (init_custom, ERF_init_custom.cpp:56-117).
void Level::init_couette_poiseuille(double rho_0, double T_0, double u_0, double v_0, double w_0, int prob_type,
                                    bool uniform) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    constant_background(*this, rho_0, T_0, uniform);
    Arr u = A(xvel), v = A(yvel), w = A(zvel);
    const double dz = cfg.dz, zlo = cfg.prob_lo[2], zhi = flat_ztop(cfg);
    if (prob_type == 1) {
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i <= nx; ++i) u(i, j, k) = u_0 * ((k + 0.5) * dz) / zhi;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j <= ny; ++j)
                for (int i = 0; i < nx; ++i) v(i, j, k) = v_0 * ((k + 0.5) * dz) / zhi;
        for (int k = 0; k <= nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) w(i, j, k) = w_0;
    } else if (prob_type == 10 || prob_type == 11) {
        for (int k = 0; k < nz; ++k) {
            const double z_h = zlo + (k + 0.5) * dz;
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i <= nx; ++i) u(i, j, k) = (prob_type == 10) ? 1.0 - z_h * z_h : 0.0;
            for (int j = 0; j <= ny; ++j)
                for (int i = 0; i < nx; ++i) v(i, j, k) = (prob_type == 11) ? 1.0 - z_h * z_h : 0.0;
        }
    } else {
        throw std::runtime_error("Couette/Poiseuille prob_type must be 1, 10 or 11");
    }
}

// Exec/RegTests/EkmanSpiral (ERF_prob.cpp:28-100), init_type uniform: the
// Ekman-layer solution u = u_g (1 - e^{-az} cos(-az)), v = -u_g e^{-az}
// sin(-az), a = 1 / sqrt(2 nu / f), nu = mu / rho_0, f = 4 pi / period,
// z = (k + 1/2) dz.
void Level::init_ekman(double rho_0, double T_0, double u_g, double rot_period, double mu) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    constant_background(*this, rho_0, T_0, true);
    Arr u = A(xvel), v = A(yvel);
    const double coriolis_factor = 4.0 * PI / rot_period;
    const double Az = mu / rho_0;
    const double DE = std::sqrt(2.0 * Az / coriolis_factor);
    const double a = 1.0 / DE;
    for (int k = 0; k < nz; ++k) {
        const double z = (k + 0.5) * cfg.dz;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) u(i, j, k) = u_g * (1.0 - std::exp(-a * z) * std::cos(-a * z));
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) v(i, j, k) = -u_g * std::exp(-a * z) * std::sin(-a * z);
    }
}

// Exec/RegTests/TaylorGreenVortex (ERF_prob.cpp:22-100) with the default init
// type: constant-density base state, init_from_hse, then rho*theta from the
// vortex pressure, RhoScalar = rho_0, u = V_0 sin x cos y cos z, v = -V_0 cos x
// sin y cos z, w = 0.
void Level::init_taylor_green(double rho_0, double V_0, double M_0) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    constant_base_state(*this, rho_0);
    std::fill(cons.v.begin(), cons.v.end(), 0.0);
    Arr c = A(cons), p0a = A(p0), r0a = A(r0);
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j)
            for (int i = -1; i <= nx; ++i) {
                c(i, j, k, 0) = r0a(i, j, k);
                c(i, j, k, 1) = getRhoThetagivenP(p0a(i, j, k));
            }
    const double* lo = cfg.prob_lo;
    const double dxv[3] = {cfg.dx, cfg.dy, cfg.dz};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double x = lo[0] + (i + 0.5) * dxv[0];
                const double y = lo[1] + (j + 0.5) * dxv[1];
                const double z = lo[2] + (k + 0.5) * dxv[2];
                const double p = rho_0 * V_0 * V_0 *
                                 (1.0 / (Gamma * M_0 * M_0) +
                                  (1.0 / 16.0) * (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2));
                c(i, j, k, 1) += getRhoThetagivenP(p) - getRhoThetagivenP(p0a(i, j, k));
                c(i, j, k, ERFX_RHOSCALAR) += 1.0 * rho_0;
            }
    Arr u = A(xvel), v = A(yvel), w = A(zvel);
    std::fill(xvel.v.begin(), xvel.v.end(), 0.0);
    std::fill(yvel.v.begin(), yvel.v.end(), 0.0);
    std::fill(zvel.v.begin(), zvel.v.end(), 0.0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const double x = lo[0] + (i + 0.0) * dxv[0];
                const double y = lo[1] + (j + 0.5) * dxv[1];
                const double z = lo[2] + (k + 0.5) * dxv[2];
                u(i, j, k) = V_0 * std::sin(x) * std::cos(y) * std::cos(z);
            }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double x = lo[0] + (i + 0.5) * dxv[0];
                const double y = lo[1] + (j + 0.0) * dxv[1];
                const double z = lo[2] + (k + 0.5) * dxv[2];
                v(i, j, k) = -V_0 * std::cos(x) * std::sin(y) * std::cos(z);
            }
    (void)w;
}

}  // namespace orc
