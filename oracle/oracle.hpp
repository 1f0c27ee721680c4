// oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of ERF's compressible dycore step (flat grid, dry or with a
// Kessler moisture model), used as
// the correctness oracle and as the CPU baseline.  Only tests/, the package's
// smoke check and bench.py's cpu_baseline leg may load it; the product path
// never calls it.
//
// The restatement follows the reference kernel by kernel and keeps its
// floating-point operation order (build with -ffp-contract=off), so that it
// reproduces ERF's results to rounding.  Each function cites the reference
// file:line it restates.  Arrays use the layout of include/erfx.h.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <stdexcept>

#include "../include/erfx.h"
// The library's reproducible pow (correctly rounded to within ~0.5 ulp, checked
// against the C library's pow by tests/test_oracle_cpu.py) stands in for
// std::pow, so that the oracle and the GPU evaluate the EOS bit for bit alike.
#include "../erf_amd/csrc/erfx_pow.h"

namespace orc {

// ---- constants (Source/ERF_Constants.H:7-20) -------------------------------
constexpr double PI       = 3.14159265358979323846264338327950288;
constexpr double PIoTwo   = PI / 2.0;   // ERF_Constants.H
constexpr double R_d      = 287.0;
constexpr double R_v      = 461.505;
constexpr double Cp_d     = 1004.5;
constexpr double p_0      = 1.0e5;
constexpr double Gamma    = 1.4;
constexpr double CONST_GRAV = 9.81;
constexpr double ip_0     = 1. / p_0;
constexpr double iR_d     = 1. / R_d;
constexpr double iGamma   = 1. / Gamma;
constexpr double L_v      = 2.5e6;      // ERF_Constants.H:16
constexpr double lcond    = 2.501e6;    // ERF_Constants.H:66
constexpr double qcw0     = 1.e-3;      // ERF_Constants.H:46
constexpr double alphaelq = 1.e-3;      // ERF_Constants.H:48
constexpr double Rd_on_Rv = R_d / R_v;  // ERF_Constants.H:87

// ---- EOS (Source/Utils/ERF_EOS.H) -------------------------------------------
// height of the domain top on a flat grid: geom.ProbHi(2), the Rayleigh
// layer's rayleigh_ztop (ERF_init_rayleigh.cpp:17); prob_lo + nz dz when the
// configuration gives no upper corner
inline double flat_ztop(const erfx_config& c) {
    return c.prob_hi[2] > c.prob_lo[2] ? c.prob_hi[2] : c.prob_lo[2] + c.nz * c.dz;
}

inline double getPgivenRTh(double rhotheta, double qv = 0.) {            // ERF_EOS.H:89-93
    return p_0 * erfx_pow(R_d * rhotheta * (1.0 + (R_v / R_d) * qv) * ip_0, Gamma);
}
inline double getExnergivenRTh(double rhotheta, double rdOcp, double qv = 0.0) { // ERF_EOS.H:130-135
    return erfx_pow(R_d * rhotheta * (1.0 + R_v / R_d * qv) * ip_0, Gamma * rdOcp);
}
inline double getTgivenRandRTh(double rho, double rhotheta, double qv = 0.0) {  // ERF_EOS.H:42-51
    double p_loc = p_0 * erfx_pow(R_d * rhotheta * (1.0 + R_v / R_d * qv) * ip_0, Gamma);
    return p_loc / (R_d * rho * (1.0 + R_v / R_d * qv));
}
inline double getExnergivenP(double P, double rdOcp) {                    // ERF_EOS.H:117-121
    return erfx_pow(P * ip_0, rdOcp);
}
inline double getRhoThetagivenP(double p, double qv = 0.0) {              // ERF_EOS.H:144-149
    return erfx_pow(p * erfx_pow(p_0, Gamma - 1), iGamma) * iR_d / (1.0 + R_v / R_d * qv);
}
inline double getRhogivenThetaPress(double theta, double p, double rdOcp, double qv = 0.0) {
    return erfx_pow(p_0, rdOcp) * erfx_pow(p, iGamma) / (R_d * theta * (1.0 + R_v / R_d * qv));
}
inline double getdPdRgivenConstantTheta(double rho, double theta, double qv = 0.0) {
    return Gamma * p_0 * erfx_pow((R_d * theta * (1.0 + R_v / R_d * qv) * ip_0), Gamma) *
           erfx_pow(rho, Gamma - 1.0);
}

// ---- saturation over liquid water (kessler.cpp) ------------------------------
double esatw(double t);     // mbar
double dtesatw(double t);   // mbar / K
double qsatw(double t, double p);
double dtqsatw(double t, double p);

// ---- column tridiagonal factor / solve (fast.cpp) ----------------------------
void column_factor(int n, const double* A, double* B, const double* C, long st);
void column_solve(int n, const double* A, const double* iB, const double* C, const double* rhs,
                  double* soln, long st);

// face value of the centred/upwind family (advection.cpp)
double face_interp(int type, double sp2, double sp1, double s, double sm1, double sm2, double sm3,
                   double upw_in, double upw_frac);

// ---- boundary-condition kinds (AMReX BCType + ERFBCType) -------------------
enum BCK { INT_DIR = 0, EXT_DIR, FOEXTRAP, REFLECT_EVEN, REFLECT_ODD, NEUMANN_INT };

// ---- array view -------------------------------------------------------------
struct Arr {
    double* p = nullptr;
    long sy = 0, sz = 0, ne = 0;
    int ng = 0;
    inline double& operator()(int i, int j, int k, int n = 0) const {
        return p[(long)(i + ng) + (long)(j + ng) * sy + (long)(k + ng) * sz + (long)n * ne];
    }
    explicit operator bool() const { return p != nullptr; }
};

struct Field {
    std::vector<double> v;
    int ncomp = 0;
    void alloc(const erfx_layout& L, int nc) {
        ncomp = nc;
        v.assign((size_t)nc * (size_t)L.nelem, std::nan(""));
    }
    Arr a(const erfx_layout& L) {
        Arr r; r.p = v.data(); r.sy = L.sy; r.sz = L.sz; r.ne = L.nelem; r.ng = L.ng; return r;
    }
};

struct Box {
    int lo[3], hi[3];
    Box() { lo[0]=lo[1]=lo[2]=0; hi[0]=hi[1]=hi[2]=-1; }
    Box(int l0,int l1,int l2,int h0,int h1,int h2){lo[0]=l0;lo[1]=l1;lo[2]=l2;hi[0]=h0;hi[1]=h1;hi[2]=h2;}
    bool ok() const { return lo[0]<=hi[0] && lo[1]<=hi[1] && lo[2]<=hi[2]; }
};

template <class F> inline void loop(const Box& b, F&& f) {
    if (!b.ok()) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i) f(i, j, k);
}
// serial variant for loops that must keep the reference's sequential order
template <class F> inline void loop_serial(const Box& b, F&& f) {
    if (!b.ok()) return;
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i) f(i, j, k);
}

struct StateA { Arr cons, xmom, ymom, zmom; };

// ---- terrain metrics (Source/Utils/ERF_TerrainMetrics.H:36-450) -------------
// z is the node height field z_phys_nd; di = 1/dx etc.
namespace met {
inline double h_xi_Iface(Arr z, int i, int j, int k, double dxi) {       // :105-116
    return 0.125 * dxi * (z(i + 1, j, k) + z(i + 1, j, k + 1) + z(i + 1, j + 1, k) + z(i + 1, j + 1, k + 1)
                          - z(i - 1, j, k) - z(i - 1, j, k + 1) - z(i - 1, j + 1, k) - z(i - 1, j + 1, k + 1));
}
inline double h_zeta_Iface(Arr z, int i, int j, int k, double dzi) {     // :91-100
    return 0.5 * dzi * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k) - z(i, j + 1, k));
}
inline double h_eta_Jface(Arr z, int i, int j, int k, double dyi) {      // :159-170
    return 0.125 * dyi * (z(i, j + 1, k) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k) + z(i + 1, j + 1, k + 1)
                          - z(i, j - 1, k) - z(i, j - 1, k + 1) - z(i + 1, j - 1, k) - z(i + 1, j - 1, k + 1));
}
inline double h_zeta_Jface(Arr z, int i, int j, int k, double dzi) {     // :135-144
    return 0.5 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k) - z(i + 1, j, k));
}
inline double h_zeta_Kface(Arr z, int i, int j, int k, double dzi) {     // :176-187
    return 0.125 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1)
                          - z(i, j, k - 1) - z(i + 1, j, k - 1) - z(i, j + 1, k - 1) - z(i + 1, j + 1, k - 1));
}
inline double h_zeta_EdgeK(Arr z, int i, int j, int k, double dzi) {     // :225-233
    return dzi * (z(i, j, k + 1) - z(i, j, k));
}
inline double h_zeta_EdgeJ(Arr z, int i, int j, int k, double dzi) {     // :268-277
    return 0.25 * dzi * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k - 1) - z(i, j + 1, k - 1));
}
inline double h_zeta_EdgeI(Arr z, int i, int j, int k, double dzi) {     // :307-316
    return 0.25 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k - 1) - z(i + 1, j, k - 1));
}
// cell-centre, k-face and edge-centre slopes for the stresses and the
// diffusive fluxes (:54-83, :192-219, :236-262, :279-303, :318-344)
inline double h_xi_Cell(Arr z, int i, int j, int k, double dxi) {
    return 0.25 * dxi * (z(i + 1, j, k) + z(i + 1, j + 1, k) + z(i + 1, j, k + 1) + z(i + 1, j + 1, k + 1)
                         - z(i, j, k) - z(i, j + 1, k) - z(i, j, k + 1) - z(i, j + 1, k + 1));
}
inline double h_eta_Cell(Arr z, int i, int j, int k, double dyi) {
    return 0.25 * dyi * (z(i, j + 1, k) + z(i + 1, j + 1, k) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1)
                         - z(i, j, k) - z(i + 1, j, k) - z(i, j, k + 1) - z(i + 1, j, k + 1));
}
inline double h_xi_Kface(Arr z, int i, int j, int k, double dxi) {
    return 0.5 * dxi * (z(i + 1, j, k) + z(i + 1, j + 1, k) - z(i, j, k) - z(i, j + 1, k));
}
inline double h_eta_Kface(Arr z, int i, int j, int k, double dyi) {
    return 0.5 * dyi * (z(i, j + 1, k) + z(i + 1, j + 1, k) - z(i, j, k) - z(i + 1, j, k));
}
inline double h_xi_EdgeK(Arr z, int i, int j, int k, double dxi) {
    return 0.25 * dxi * (z(i + 1, j, k) + z(i + 1, j, k + 1) - z(i - 1, j, k) - z(i - 1, j, k + 1));
}
inline double h_eta_EdgeK(Arr z, int i, int j, int k, double dyi) {
    return 0.25 * dyi * (z(i, j + 1, k) + z(i, j + 1, k + 1) - z(i, j - 1, k) - z(i, j - 1, k + 1));
}
inline double h_xi_EdgeJ(Arr z, int i, int j, int k, double dxi) {
    return 0.25 * dxi * (z(i + 1, j + 1, k) + z(i + 1, j, k) - z(i - 1, j + 1, k) - z(i - 1, j, k));
}
inline double h_eta_EdgeJ(Arr z, int i, int j, int k, double dyi) {
    return dyi * (z(i, j + 1, k) - z(i, j, k));
}
inline double h_xi_EdgeI(Arr z, int i, int j, int k, double dxi) {
    return dxi * (z(i + 1, j, k) - z(i, j, k));
}
inline double h_eta_EdgeI(Arr z, int i, int j, int k, double dyi) {
    return 0.25 * dyi * (z(i + 1, j + 1, k) + z(i, j + 1, k) - z(i + 1, j - 1, k) - z(i, j - 1, k));
}
// slopes of the z-face (i, j, k) used by OmegaFromW / WFromOmega (:368-381)
inline double zlo_xi(Arr z, int i, int j, int k, double dxi) {
    return 0.5 * dxi * (z(i + 1, j + 1, k) + z(i + 1, j, k) - z(i, j + 1, k) - z(i, j, k));
}
inline double zlo_eta(Arr z, int i, int j, int k, double dyi) {
    return 0.5 * dyi * (z(i + 1, j + 1, k) + z(i, j + 1, k) - z(i + 1, j, k) - z(i, j, k));
}
// u and v averaged to the z-face, extrapolated at k = 0 (:383-388)
inline double u_at_zface(Arr u, int i, int j, int k) {
    return (k == 0) ? 1.5 * (0.5 * (u(i, j, k) + u(i + 1, j, k))) - 0.5 * (0.5 * (u(i, j, k + 1) + u(i + 1, j, k + 1)))
                    : 0.25 * (u(i, j, k - 1) + u(i + 1, j, k - 1) + u(i, j, k) + u(i + 1, j, k));
}
inline double v_at_zface(Arr v, int i, int j, int k) {
    return (k == 0) ? 1.5 * (0.5 * (v(i, j, k) + v(i, j + 1, k))) - 0.5 * (0.5 * (v(i, j, k + 1) + v(i, j + 1, k + 1)))
                    : 0.25 * (v(i, j, k - 1) + v(i, j + 1, k - 1) + v(i, j, k) + v(i, j + 1, k));
}
// OmegaFromW (:359-392) and WFromOmega (:397-445)
inline double omega_from_w(int i, int j, int k, double w, Arr u, Arr v, Arr z, const double* di) {
    double mx = zlo_xi(z, i, j, k, di[0]);
    double my = zlo_eta(z, i, j, k, di[1]);
    double uu = u_at_zface(u, i, j, k);
    double vv = v_at_zface(v, i, j, k);
    return w - mx * uu - my * vv;
}
inline double w_from_omega(int i, int j, int k, double omega, Arr u, Arr v, Arr z, const double* di) {
    double uu = u_at_zface(u, i, j, k);
    double vv = v_at_zface(v, i, j, k);
    double mx = zlo_xi(z, i, j, k, di[0]);
    double my = zlo_eta(z, i, j, k, di[1]);
    return omega + mx * uu + my * vv;
}
}  // namespace met
struct StateF {
    Field cons, xmom, ymom, zmom;
    void alloc(const erfx_layout& L, int nc) { cons.alloc(L, nc); xmom.alloc(L,1); ymom.alloc(L,1); zmom.alloc(L,1); }
    StateA a(const erfx_layout& L) { return {cons.a(L), xmom.a(L), ymom.a(L), zmom.a(L)}; }
};

// EddyDiff component indices (Source/ERF_IndexDefines.H:152-169)
namespace EddyDiff {
enum { Mom_h = 0, Theta_h, KE_h, QKE_h, Scalar_h, Q_h, Mom_v, Theta_v, KE_v, QKE_v, Scalar_v, Q_v, PBL_lengthscale, NumDiffs };
}

// ---- the level: configuration, BCs and every buffer of the step ----------
struct Level {
    erfx_config cfg;
    erfx_layout L;
    int ng_cons = 0, ng_vel = 0;
    int ncons = 5;
    bool moist = false;        // a Kessler model is on: cons comps 5..7 are RhoQ1..RhoQ3
    bool terrain = false;      // terrain-following grid: z_phys_nd and its metrics below
    int n_qstate = 0;          // Kessler::Qstate_Size (ERF_Kessler.H:172)
    double rdOcp = 0, gravity = 0;
    double dxInv[3];
    // BC kinds: [var][ori], var 0..2 = x/y/z velocity, 3 = cons; ori 0..5 = xlo,ylo,zlo,xhi,yhi,zhi
    int bck[4][6];
    double vel_extdir[3][6];
    double cons_extdir[ERFX_NCONS_MOIST][6];   // ext_dir values of the cons components (inflow faces)
    bool periodic[3];

    // persistent state (vars_new) and base state
    Field cons, xvel, yvel, zvel;
    Field r0, p0, pi0;
    // advance scratch (vars_old aliases and momenta)
    Field cons_old_vars, xvel_old, yvel_old, zvel_old;
    // MRI buffers
    StateF S_old, S_new, S_sum, S_scratch, F_slow;
    Field prim, pi_stage, fast_coeffs;
    Field omega, eddy, hfx3;
    Field tau11, tau22, tau33, tau12, tau13, tau23;
    Field tau21, tau31, tau32;   // terrain grids: the non-symmetric stresses
    Field xmom_src, ymom_src, zmom_src, cc_src;

    void init(const erfx_config& c, const erfx_layout& lay);
    Arr A(Field& f) { return f.a(L); }
    StateA A(StateF& s) { return s.a(L); }

    // ---- hooks (restatements) --------------------------------------------
    void cons_to_prim(Arr cons_in, int ng);
    void fill_cons(Arr c, int icomp, int ncomp, int ng);
    void fill_xvel(Arr u, int ng);
    void fill_yvel(Arr v, int ng);
    void fill_zvel(Arr w, int ng, Arr u, Arr v);   // u, v: the filled velocities (terrain walls)
    void fill_boundary_periodic(Arr a, int ncomp, int ng, int nodal_dir);
    void momentum_to_velocity(Arr xv, Arr yv, Arr zv, Arr dens, Arr xm, Arr ym, Arr zm);
    void velocity_to_momentum(Arr xv, int ngu[3], Arr yv, int ngv[3], Arr zv, int ngw[3], Arr dens,
                              Arr xm, Arr ym, Arr zm);
    void fill_intermediate_patch(StateA S, Arr xv, Arr yv, Arr zv, int ng_cons_, int ng_vel_,
                                 bool cons_only, int icomp, int ncomp);
    void apply_bcs(StateA S, Arr xv, Arr yv, Arr zv, int ng_cons_, int ng_vel_, bool fast_only,
                   bool vel_and_mom_synced);
    void fill_patch(Arr c, Arr u, Arr v, Arr w);

    void make_fast_coeffs(Arr fc, StateA stage, Arr prim_, Arr pi_st, double dtau, double beta_s);
    void no_substep(StateA ssum, StateA sold, StateA fslow, double slow_dt);
    void fast_rhs_N(int step, int nrk, StateA slow_rhs, StateA prev, StateA stage, Arr prim_,
                    Arr pi_st, Arr fc, StateA data, StateA scratch, double dtau, double beta_s,
                    double facinv);

    void compute_strain_N(const Box& bxcc, Box tbxxy, Box tbxxz, Box tbxyz, Arr u, Arr v, Arr w,
                          Arr t11, Arr t22, Arr t33, Arr t12, Arr t13, Arr t23);
    void turbulent_viscosity(Arr cons_in, Arr u, Arr v, Arr w);
    void make_tau_terms(StateA data, Arr u, Arr v, Arr w);
    // terrain-following grid (molecular diffusion)
    void compute_strain_T(Box bxcc, Box tbxxy, Box tbxxz, Box tbxyz, Arr u, Arr v, Arr w, Arr t11, Arr t22,
                          Arr t33, Arr t12, Arr t13, Arr t21, Arr t23, Arr t31, Arr t32);
    void compute_stress_cons_T(Box bxcc, Box tbxxy, Box tbxxz, Box tbxyz, double mu_eff, bool use_rho, Arr cd,
                               Arr t11, Arr t22, Arr t33, Arr t12, Arr t13, Arr t21, Arr t23, Arr t31, Arr t32,
                               Arr er);
    void make_tau_terms_T(StateA data, Arr u, Arr v, Arr w);
    void diffusion_src_state_T(const Box& bx, int start_comp, int num_comp, Arr cell_data, Arr cell_prim, Arr rhs);
    void make_mom_sources(StateA data, Arr prim_);
    // PlaneAverage of ERF_make_buoyancy.cpp:216-255: per level k of the valid
    // box, the sums of value / (nx ny); rows rho, theta, qv, qc, qp
    void plane_averages(StateA data, Arr prim_, std::vector<double>& avg);
    // Kessler warm rain once per step (ERF_advance_microphysics.cpp:5-16)
    void kessler_advance(Arr c, double dt);
    std::vector<double> rain_accum;   // k = 0 plane of the layout (sz values)
    // terrain (use_terrain): node heights and the metrics of make_J, make_areas
    // and make_zcc (Source/Utils/ERF_TerrainMetrics.cpp:583-715); with a flat
    // grid they hold detJ = ax = ay = az = 1
    Field z_nd, detJ, ax, ay, az, z_cc;
    // smallest cell height: dz, times the smallest detJ of the valid cells on
    // a terrain grid (ERF.cpp:898-906), for the time-step estimate
    double dz_min = 0.0;
    void set_terrain(const double* z_nodes);
    void fast_rhs_T(int step, int nrk, StateA slow_rhs, StateA prev, StateA stage, Arr prim_, Arr pi_st,
                    Arr fc, StateA data, StateA scratch, double dtau, double beta_s, double facinv);
    void make_sources(StateA data, Arr prim_);
    void init_density_current(double x_c, double z_c, double x_r, double z_r, double T_pert);
    void init_density_current_T(double x_c, double z_c, double x_r, double z_r, double T_pert);
    void init_scalar_advdiff(const double* params);
    void init_taylor_green(double rho_0, double V_0, double M_0);
    void init_couette_poiseuille(double rho_0, double T_0, double u_0, double v_0, double w_0, int prob_type,
                                 bool uniform);
    void init_ekman(double rho_0, double T_0, double u_g, double rot_period, double mu);
    // ERF::estTimeStep (ERF_ComputeTimestep.cpp:49-233): returns dt, sets the MRI ratio
    double est_time_step(double cfl, double fixed_dt, double fixed_fast_dt, long& ratio);
    // Rayleigh damping reference profiles (nz+1 values each), set by orc_set_rayleigh
    std::vector<double> ray_ubar, ray_vbar, ray_wbar, ray_thetabar;
    void slow_rhs_pre(int nrk, double dt, StateA rhs, StateA data, Arr prim_, StateA scratch,
                      Arr u, Arr v, Arr w);
    void slow_rhs_post(int nrk, double dt, StateA rhs, StateA old, StateA snew, StateA data,
                       Arr prim_, StateA scratch, Arr u, Arr v);
    void advection_src_scalars(const Box& bx, int icomp, int ncomp, Arr ax, Arr ay, Arr az,
                               Arr cprim, Arr rhs, int hadv, int vadv, double hfrac, double vfrac);
    void diffusion_src_state(const Box& bx, int start_comp, int num_comp, Arr cell_data,
                             Arr cell_prim, Arr rhs);
    void advection_src_mom(const Box& tbx, const Box& tby, const Box& tbz, Arr ru_rhs, Arr rv_rhs,
                           Arr rw_rhs, Arr u, Arr v, Arr w, Arr rho_u, Arr rho_v, Arr Omega);
    void advection_src_mom_T(const Box& tbx, const Box& tby, const Box& tbz, Arr ru_rhs, Arr rv_rhs,
                             Arr rw_rhs, Arr u, Arr v, Arr w, Arr rho_u, Arr rho_v, Arr Omega);

    void advance_dycore(double old_time, double dt);
    void advance(double time, double dt);   // ERF::Advance (lev 0, dycore only)
};

}  // namespace orc
