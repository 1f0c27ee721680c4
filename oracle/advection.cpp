// advection.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  Flux-form advection of
// scalars and momentum on a flat grid without map factors.
#include "oracle.hpp"

namespace orc {

// Face interpolation of the centred/upwind family, lo face of cell i along a
// line (sm3..sp2 = values at i-3..i+2).  ERF_Interpolation_UPW.H: CENTERED2
// :9-87, UPWIND3 :92-201, CENTERED4 :207-295, UPWIND5 :300-395, CENTERED6 :400-480.
// WENO family (Source/Utils/ERF_Interpolation_WENO.H: WENO3 :9-127, WENO5 :132-262;
// ERF_Interpolation_WENO_Z.H: WENO_Z3 :9-126, WENO_MZQ3 :130-252, WENO_Z5 :256-389).
// Evaluate() takes the stencil ordered from the upwind side.
namespace weno {
constexpr double eps = 1.0e-6, tol = 1.0e-12;
static inline double w3(double sm1, double s, double sp1) {
    const double g1 = (1.0 / 3.0), g2 = (2.0 / 3.0);
    double b1 = (s - sm1) * (s - sm1);
    double b2 = (sp1 - s) * (sp1 - s);
    double w1 = g1 / ((eps + b1) * (eps + b1));
    double w2 = g2 / ((eps + b2) * (eps + b2));
    double wsum = w1 + w2;
    double v1 = -sm1 + 3.0 * s;
    double v2 = s + sp1;
    return ((w1 * v1 + w2 * v2) / (2.0 * wsum));
}
static inline double z3(double sm1, double s, double sp1) {
    const double g1 = (1.0 / 3.0), g2 = (2.0 / 3.0);
    double b1 = (s - sm1) * (s - sm1);
    double b2 = (sp1 - s) * (sp1 - s);
    double t5 = std::abs(b2 - b1);
    double w1 = g1 * (1.0 + (t5 * t5) / ((eps + b1) * (eps + b1)));
    double w2 = g2 * (1.0 + (t5 * t5) / ((eps + b2) * (eps + b2)));
    double wsum = w1 + w2;
    double v1 = -sm1 + 3.0 * s;
    double v2 = s + sp1;
    return ((w1 * v1 + w2 * v2) / (2.0 * wsum));
}
static inline double mzq3(double sm1, double s, double sp1) {
    const double g1 = (1.0 / 3.0), g2 = (1.0 / 3.0), g3 = (1.0 / 3.0);
    double b1 = (s - sm1) * (s - sm1);
    double b2 = (sp1 - s) * (sp1 - s);
    double b3 = ((13.0 / 12.0) * ((sm1 - 2.0 * s + sp1) * (sm1 - 2.0 * s + sp1))) + (((sp1 - sm1) * (sp1 - sm1)) / 4.0);
    double t5 = (std::abs(b3 - b1) + std::abs(b3 - b2)) / 32.0;
    double a1 = g1 * (1.0 + (t5 * t5) / ((eps + b1) * (eps + b1)));
    double a2 = g2 * (1.0 + (t5 * t5) / ((eps + b2) * (eps + b2)));
    double a3 = g3 * (1.0 + (t5 * t5) / ((eps + b3) * (eps + b3)));
    double asum = a1 + a2 + a3;
    double w1 = a1 / asum;
    double w2 = a2 / asum;
    double w3_ = a3 / asum;
    double v1 = (-sm1 + 3.0 * s) / 2.0;
    double v2 = (s + sp1) / 2.0;
    double v3 = (-sm1 + 5.0 * s + 2.0 * sp1) / 6.0;
    return ((w3_ / g3) * (v3 - g1 * v1 - g2 * v2) + w1 * v1 + w2 * v2);
}
static inline void beta5(double sm2, double sm1, double s, double sp1, double sp2, double& b1, double& b2, double& b3) {
    const double c1 = (13.0 / 12.0);
    b1 = c1 * (sm2 - 2.0 * sm1 + s) * (sm2 - 2.0 * sm1 + s) +
         0.25 * (sm2 - 4.0 * sm1 + 3.0 * s) * (sm2 - 4.0 * sm1 + 3.0 * s);
    b2 = c1 * (sm1 - 2.0 * s + sp1) * (sm1 - 2.0 * s + sp1) + 0.25 * (sm1 - sp1) * (sm1 - sp1);
    b3 = c1 * (s - 2.0 * sp1 + sp2) * (s - 2.0 * sp1 + sp2) +
         0.25 * (3.0 * s - 4.0 * sp1 + sp2) * (3.0 * s - 4.0 * sp1 + sp2);
}
static inline double w5(double sm2, double sm1, double s, double sp1, double sp2, bool z) {
    const double g1 = (1.0 / 10.0), g2 = (3.0 / 5.0), g3 = (3.0 / 10.0);
    double b1, b2, b3;
    beta5(sm2, sm1, s, sp1, sp2, b1, b2, b3);
    double w1, w2, w3_;
    if (z) {
        double t5 = std::abs(b3 - b1);
        w1 = g1 * (1.0 + (t5 * t5) / ((eps + b1) * (eps + b1)));
        w2 = g2 * (1.0 + (t5 * t5) / ((eps + b2) * (eps + b2)));
        w3_ = g3 * (1.0 + (t5 * t5) / ((eps + b3) * (eps + b3)));
    } else {
        w1 = g1 / ((eps + b1) * (eps + b1));
        w2 = g2 / ((eps + b2) * (eps + b2));
        w3_ = g3 / ((eps + b3) * (eps + b3));
    }
    double wsum = w1 + w2 + w3_;
    double v1 = 2.0 * sm2 - 7.0 * sm1 + 11.0 * s;
    double v2 = -sm1 + 5.0 * s + 2.0 * sp1;
    double v3 = 2.0 * s + 5.0 * sp1 - sp2;
    return ((w1 * v1 + w2 * v2 + w3_ * v3) / (6.0 * wsum));
}
// lo face of s; sm3..sp2 at offsets -3..+2
static inline double face(int type, double sp2, double sp1, double s, double sm1, double sm2, double sm3, double upw) {
    if (!(upw > tol) && !(upw < -tol)) return 0.5 * (s + sm1);
    const bool pos = upw > tol;
    switch (type) {
        case ERFX_ADV_WENO3: return pos ? w3(sm2, sm1, s) : w3(sp1, s, sm1);
        case ERFX_ADV_WENOZ3: return pos ? z3(sm2, sm1, s) : z3(sp1, s, sm1);
        case ERFX_ADV_WENOMZQ3: return pos ? mzq3(sm2, sm1, s) : mzq3(sp1, s, sm1);
        case ERFX_ADV_WENO5: return pos ? w5(sm3, sm2, sm1, s, sp1, false) : w5(sp2, sp1, s, sm1, sm2, false);
        case ERFX_ADV_WENOZ5: return pos ? w5(sm3, sm2, sm1, s, sp1, true) : w5(sp2, sp1, s, sm1, sm2, true);
    }
    throw std::runtime_error("bad WENO type");
}
}  // namespace weno

static inline bool is_weno(int t) { return t >= ERFX_ADV_WENO3 && t <= ERFX_ADV_WENOZ5; }

static inline double interp_line(int type, double sp2, double sp1, double s, double sm1,
                                 double sm2, double sm3, double upw_in, double upw_frac) {
    if (is_weno(type)) return weno::face(type, sp2, sp1, s, sm1, sm2, sm3, upw_in);
    switch (type) {
        case ERFX_ADV_CENTERED2: {
            double a1 = (s + sm1);
            return (0.5 * a1);
        }
        case ERFX_ADV_UPWIND3: {
            double upw = upw_in;
            if (upw != 0.) upw = (upw > 0) ? 1. : -1.;
            upw *= upw_frac;
            double a1 = (s + sm1), d1 = (s - sm1), a2 = (sp1 + sm2), d2 = (sp1 - sm2);
            const double g1 = (7.0 / 12.0), g2 = (1.0 / 12.0);
            return (g1 * a1 - g2 * a2 + upw * g2 * (d2 - 3.0 * d1));
        }
        case ERFX_ADV_CENTERED4: {
            double a1 = (s + sm1), a2 = (sp1 + sm2);
            const double g1 = (7.0 / 12.0), g2 = (1.0 / 12.0);
            return (g1 * a1 - g2 * a2);
        }
        case ERFX_ADV_UPWIND5: {
            double upw = upw_in;
            if (upw != 0.) upw = (upw > 0) ? 1. : -1.;
            upw *= upw_frac;
            double a1 = (s + sm1), a2 = (sp1 + sm2), a3 = (sp2 + sm3);
            double d1 = (s - sm1), d2 = (sp1 - sm2), d3 = (sp2 - sm3);
            const double g1 = (37.0 / 60.0), g2 = (2.0 / 15.0), g3 = (1.0 / 60.0);
            return (g1 * a1 - g2 * a2 + g3 * a3 - upw * g3 * (d3 - 5.0 * d2 + 10.0 * d1));
        }
        case ERFX_ADV_CENTERED6: {
            double a1 = (s + sm1), a2 = (sp1 + sm2), a3 = (sp2 + sm3);
            const double g1 = (37.0 / 60.0), g2 = (2.0 / 15.0), g3 = (1.0 / 60.0);
            return (g1 * a1 - g2 * a2 + g3 * a3);
        }
    }
    throw std::runtime_error("bad adv type");
}

double face_interp(int type, double sp2, double sp1, double s, double sm1, double sm2, double sm3,
                   double upw_in, double upw_frac) {
    return interp_line(type, sp2, sp1, s, sm1, sm2, sm3, upw_in, upw_frac);
}

static inline bool wide(int t) {
    return t == ERFX_ADV_UPWIND5 || t == ERFX_ADV_CENTERED6 || t == ERFX_ADV_WENO5 || t == ERFX_ADV_WENOZ5;
}
static inline bool narrow(int t) { return t == ERFX_ADV_CENTERED2; }

static inline double interp_x(int t, Arr q, int i, int j, int k, int n, double upw, double fr) {
    if (narrow(t)) return interp_line(t, 0, 0, q(i, j, k, n), q(i - 1, j, k, n), 0, 0, upw, fr);
    double sp2 = wide(t) ? q(i + 2, j, k, n) : 0., sm3 = wide(t) ? q(i - 3, j, k, n) : 0.;
    return interp_line(t, sp2, q(i + 1, j, k, n), q(i, j, k, n), q(i - 1, j, k, n), q(i - 2, j, k, n), sm3, upw, fr);
}
static inline double interp_y(int t, Arr q, int i, int j, int k, int n, double upw, double fr) {
    if (narrow(t)) return interp_line(t, 0, 0, q(i, j, k, n), q(i, j - 1, k, n), 0, 0, upw, fr);
    double sp2 = wide(t) ? q(i, j + 2, k, n) : 0., sm3 = wide(t) ? q(i, j - 3, k, n) : 0.;
    return interp_line(t, sp2, q(i, j + 1, k, n), q(i, j, k, n), q(i, j - 1, k, n), q(i, j - 2, k, n), sm3, upw, fr);
}
static inline double interp_z(int t, Arr q, int i, int j, int k, int n, double upw, double fr) {
    if (narrow(t)) return interp_line(t, 0, 0, q(i, j, k, n), q(i, j, k - 1, n), 0, 0, upw, fr);
    double sp2 = wide(t) ? q(i, j, k + 2, n) : 0., sm3 = wide(t) ? q(i, j, k - 3, n) : 0.;
    return interp_line(t, sp2, q(i, j, k + 1, n), q(i, j, k, n), q(i, j, k - 1, n), q(i, j, k - 2, n), sm3, upw, fr);
}

// AdvectionSrcForScalars (Source/Advection/ERF_AdvectionSrcForState.cpp:118-395,
// ERF_AdvectionSrcForScalars.H:7-110), unit map factors; 1/detJ on a
// terrain-following grid (:298, :361).
void Level::advection_src_scalars(const Box& bx, int icomp, int ncomp, Arr ax, Arr ay, Arr az,
                                  Arr cprim, Arr rhs, int hadv, int vadv, double hfrac, double vfrac) {
    const int nx = bx.hi[0], ny = bx.hi[1], nz = bx.hi[2];
    // a WENO horizontal type is used in all three directions (:228-262)
    if (is_weno(hadv)) vadv = hadv;
    Field fx, fy, fz;
    fx.alloc(L, 1); fy.alloc(L, 1); fz.alloc(L, 1);
    Arr Fx = fx.a(L), Fy = fy.a(L), Fz = fz.a(L);
    for (int n = 0; n < ncomp; ++n) {
        int ci = icomp + n, pi = ci - 1;
        if (hadv == ERFX_ADV_CENTERED2 && vadv == ERFX_ADV_CENTERED2) {
            loop(Box(0, 0, 0, nx + 1, ny, nz), [&](int i, int j, int k) {
                double pf = 0.5 * (cprim(i, j, k, pi) + cprim(i - 1, j, k, pi));
                Fx(i, j, k) = ax(i, j, k) * pf;
            });
            loop(Box(0, 0, 0, nx, ny + 1, nz), [&](int i, int j, int k) {
                double pf = 0.5 * (cprim(i, j, k, pi) + cprim(i, j - 1, k, pi));
                Fy(i, j, k) = ay(i, j, k) * pf;
            });
            loop(Box(0, 0, 0, nx, ny, nz + 1), [&](int i, int j, int k) {
                double pf = 0.5 * (cprim(i, j, k, pi) + cprim(i, j, k - 1, pi));
                Fz(i, j, k) = az(i, j, k) * pf;
            });
        } else {
            loop(Box(0, 0, 0, nx + 1, ny, nz), [&](int i, int j, int k) {
                double v = interp_x(hadv, cprim, i, j, k, pi, ax(i, j, k), hfrac);
                Fx(i, j, k) = ax(i, j, k) * v;
            });
            loop(Box(0, 0, 0, nx, ny + 1, nz), [&](int i, int j, int k) {
                double v = interp_y(hadv, cprim, i, j, k, pi, ay(i, j, k), hfrac);
                Fy(i, j, k) = ay(i, j, k) * v;
            });
            loop(Box(0, 0, 0, nx, ny, nz + 1), [&](int i, int j, int k) {
                double v = interp_z(vadv, cprim, i, j, k, pi, az(i, j, k), vfrac);
                Fz(i, j, k) = az(i, j, k) * v;
            });
        }
        const double dxi = dxInv[0], dyi = dxInv[1], dzi = dxInv[2];
        Arr Jc = A(detJ);
        loop(bx, [&](int i, int j, int k) {
            double invdetJ = terrain ? ((Jc(i, j, k) > 0.) ? 1. / Jc(i, j, k) : 1.) : 1.;
            double mfsq = 1.0 * 1.0;
            rhs(i, j, k, ci) = -invdetJ * mfsq * ((Fx(i + 1, j, k) - Fx(i, j, k)) * dxi +
                                                   (Fy(i, j + 1, k) - Fy(i, j, k)) * dyi +
                                                   (Fz(i, j, k + 1) - Fz(i, j, k)) * dzi);
        });
    }
}

// UPWINDALL::InterpolateInZ (ERF_Interpolation_UPW.H, struct UPWINDALL): the
// wall-reduced vertical interpolation used for z-momentum near k boundaries.
static inline double interp_z_all(Arr q, int i, int j, int k, double upw_lo, double upw_frac, int t) {
    if (t == ERFX_ADV_CENTERED2) return 0.5 * (q(i, j, k) + q(i, j, k - 1));
    double sp2 = wide(t) ? q(i, j, k + 2) : 0.;
    double sp1 = q(i, j, k + 1), s = q(i, j, k), sm1 = q(i, j, k - 1), sm2 = q(i, j, k - 2);
    double sm3 = wide(t) ? q(i, j, k - 3) : 0.;
    if (upw_lo != 0.) upw_lo = (upw_lo > 0) ? 1. : -1.;
    upw_lo *= upw_frac;
    double a1 = (s + sm1), a2 = (sp1 + sm2), d1 = (s - sm1), d2 = (sp1 - sm2), a3 = (sp2 + sm3), d3 = (sp2 - sm3);
    const double g1_3_4 = (7.0 / 12.0), g2_3_4 = (1.0 / 12.0);
    const double g1_5_6 = (37.0 / 60.0), g2_5_6 = (2.0 / 15.0), g3_5_6 = (1.0 / 60.0);
    double iv = 0.0;
    if (t == ERFX_ADV_UPWIND3) iv = g1_3_4 * a1 - g2_3_4 * a2 + upw_lo * g2_3_4 * (d2 - 3.0 * d1);
    else if (t == ERFX_ADV_CENTERED4) iv = g1_3_4 * a1 - g2_3_4 * a2;
    else if (t == ERFX_ADV_UPWIND5) iv = g1_5_6 * a1 - g2_5_6 * a2 + g3_5_6 * a3 - upw_lo * g3_5_6 * (d3 - 5.0 * d2 + 10.0 * d1);
    else if (t == ERFX_ADV_CENTERED6) iv = g1_5_6 * a1 - g2_5_6 * a2 + g3_5_6 * a3;
    return iv;
}

// AdvectionSrcForMom, flat branch (ERF_AdvectionSrcForMom.cpp:98-158 for
// C2/C2, ERF_AdvectionSrcForMom_N.H:17-407 otherwise), unit map factors.
void Level::advection_src_mom(const Box& tbx, const Box& tby, const Box& tbz, Arr ru_rhs,
                              Arr rv_rhs, Arr rw_rhs, Arr u, Arr v, Arr w, Arr rho_u, Arr rho_v,
                              Arr Omega) {
    const double dxInv_ = dxInv[0], dyInv_ = dxInv[1], dzInv_ = dxInv[2];
    const int hz = L.nz;   // hi_z_face
    const int lz = 0;      // lo_z_face
    const int ht = cfg.dycore_horiz_adv, vt = cfg.dycore_vert_adv;
    const double hf = cfg.dycore_horiz_upw_frac, vf = cfg.dycore_vert_upw_frac;
    if (ht == ERFX_ADV_CENTERED2 && vt == ERFX_ADV_CENTERED2) {
        loop(tbx, [&](int i, int j, int k) {
            double xflux_hi = 0.25 * (rho_u(i, j, k) * 1.0 + rho_u(i + 1, j, k) * 1.0) * (u(i + 1, j, k) + u(i, j, k));
            double xflux_lo = 0.25 * (rho_u(i, j, k) * 1.0 + rho_u(i - 1, j, k) * 1.0) * (u(i - 1, j, k) + u(i, j, k));
            double yflux_hi = 0.25 * (rho_v(i, j + 1, k) * 1.0 + rho_v(i - 1, j + 1, k) * 1.0) * (u(i, j + 1, k) + u(i, j, k));
            double yflux_lo = 0.25 * (rho_v(i, j, k) * 1.0 + rho_v(i - 1, j, k) * 1.0) * (u(i, j - 1, k) + u(i, j, k));
            double zflux_hi = 0.25 * (Omega(i, j, k + 1) + Omega(i - 1, j, k + 1)) * (u(i, j, k + 1) + u(i, j, k));
            double zflux_lo = 0.25 * (Omega(i, j, k) + Omega(i - 1, j, k)) * (u(i, j, k - 1) + u(i, j, k));
            double mfsq = 1.0 * 1.0;
            double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                       + (zflux_hi - zflux_lo) * dzInv_;
            ru_rhs(i, j, k) = -src;
        });
        loop(tby, [&](int i, int j, int k) {
            double xflux_hi = 0.25 * (rho_u(i + 1, j, k) * 1.0 + rho_u(i + 1, j - 1, k) * 1.0) * (v(i + 1, j, k) + v(i, j, k));
            double xflux_lo = 0.25 * (rho_u(i, j, k) * 1.0 + rho_u(i, j - 1, k) * 1.0) * (v(i - 1, j, k) + v(i, j, k));
            double yflux_hi = 0.25 * (rho_v(i, j + 1, k) * 1.0 + rho_v(i, j, k) * 1.0) * (v(i, j + 1, k) + v(i, j, k));
            double yflux_lo = 0.25 * (rho_v(i, j, k) * 1.0 + rho_v(i, j - 1, k) * 1.0) * (v(i, j - 1, k) + v(i, j, k));
            double zflux_hi = 0.25 * (Omega(i, j, k + 1) + Omega(i, j - 1, k + 1)) * (v(i, j, k + 1) + v(i, j, k));
            double zflux_lo = 0.25 * (Omega(i, j, k) + Omega(i, j - 1, k)) * (v(i, j, k - 1) + v(i, j, k));
            double mfsq = 1.0 * 1.0;
            double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                       + (zflux_hi - zflux_lo) * dzInv_;
            rv_rhs(i, j, k) = -src;
        });
        loop(tbz, [&](int i, int j, int k) {
            double xflux_hi = 0.25 * (rho_u(i + 1, j, k) + rho_u(i + 1, j, k - 1)) * 1.0 * (w(i + 1, j, k) + w(i, j, k));
            double xflux_lo = 0.25 * (rho_u(i, j, k) + rho_u(i, j, k - 1)) * 1.0 * (w(i - 1, j, k) + w(i, j, k));
            double yflux_hi = 0.25 * (rho_v(i, j + 1, k) + rho_v(i, j + 1, k - 1)) * 1.0 * (w(i, j + 1, k) + w(i, j, k));
            double yflux_lo = 0.25 * (rho_v(i, j, k) + rho_v(i, j, k - 1)) * 1.0 * (w(i, j - 1, k) + w(i, j, k));
            double zflux_lo = 0.25 * (Omega(i, j, k) + Omega(i, j, k - 1)) * (w(i, j, k) + w(i, j, k - 1));
            double zflux_hi = (k == hz) ? Omega(i, j, k) * w(i, j, k)
                                        : 0.25 * (Omega(i, j, k) + Omega(i, j, k + 1)) * (w(i, j, k) + w(i, j, k + 1));
            double mfsq = 1.0 * 1.0;
            double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                       + (zflux_hi - zflux_lo) * dzInv_;
            rw_rhs(i, j, k) = -src;
        });
        return;
    }
    // higher order: AdvectionSrcForXMom_N / YMom_N / ZMom_N
    loop(tbx, [&](int i, int j, int k) {
        double ih, il;
        double rho_u_avg_hi = 0.5 * (rho_u(i + 1, j, k) * 1.0 + rho_u(i, j, k) * 1.0);
        ih = interp_x(ht, u, i + 1, j, k, 0, rho_u_avg_hi, hf);
        double xflux_hi = rho_u_avg_hi * ih;
        double rho_u_avg_lo = 0.5 * (rho_u(i - 1, j, k) * 1.0 + rho_u(i, j, k) * 1.0);
        il = interp_x(ht, u, i, j, k, 0, rho_u_avg_lo, hf);
        double xflux_lo = rho_u_avg_lo * il;
        double rho_v_avg_hi = 0.5 * (rho_v(i, j + 1, k) * 1.0 + rho_v(i - 1, j + 1, k) * 1.0);
        ih = interp_y(ht, u, i, j + 1, k, 0, rho_v_avg_hi, hf);
        double yflux_hi = rho_v_avg_hi * ih;
        double rho_v_avg_lo = 0.5 * (rho_v(i, j, k) * 1.0 + rho_v(i - 1, j, k) * 1.0);
        il = interp_y(ht, u, i, j, k, 0, rho_v_avg_lo, hf);
        double yflux_lo = rho_v_avg_lo * il;
        double rho_w_avg_hi = 0.5 * (Omega(i, j, k + 1) + Omega(i - 1, j, k + 1));
        ih = interp_z(vt, u, i, j, k + 1, 0, rho_w_avg_hi, vf);
        double zflux_hi = rho_w_avg_hi * ih;
        double rho_w_avg_lo = 0.5 * (Omega(i, j, k) + Omega(i - 1, j, k));
        il = interp_z(vt, u, i, j, k, 0, rho_w_avg_lo, vf);
        double zflux_lo = rho_w_avg_lo * il;
        double mfsq = 1 / (1.0 * 1.0);
        double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                   + (zflux_hi - zflux_lo) * dzInv_;
        ru_rhs(i, j, k) = -src;
    });
    loop(tby, [&](int i, int j, int k) {
        double ih, il;
        double rho_u_avg_hi = 0.5 * (rho_u(i + 1, j, k) * 1.0 + rho_u(i + 1, j - 1, k) * 1.0);
        ih = interp_x(ht, v, i + 1, j, k, 0, rho_u_avg_hi, hf);
        double xflux_hi = rho_u_avg_hi * ih;
        double rho_u_avg_lo = 0.5 * (rho_u(i, j, k) * 1.0 + rho_u(i, j - 1, k) * 1.0);
        il = interp_x(ht, v, i, j, k, 0, rho_u_avg_lo, hf);
        double xflux_lo = rho_u_avg_lo * il;
        double rho_v_avg_hi = 0.5 * (rho_v(i, j, k) * 1.0 + rho_v(i, j + 1, k) * 1.0);
        ih = interp_y(ht, v, i, j + 1, k, 0, rho_v_avg_hi, hf);
        double yflux_hi = rho_v_avg_hi * ih;
        double rho_v_avg_lo = 0.5 * (rho_v(i, j, k) * 1.0 + rho_v(i, j - 1, k) * 1.0);
        il = interp_y(ht, v, i, j, k, 0, rho_v_avg_lo, hf);
        double yflux_lo = rho_v_avg_lo * il;
        double rho_w_avg_hi = 0.5 * (Omega(i, j, k + 1) + Omega(i, j - 1, k + 1));
        ih = interp_z(vt, v, i, j, k + 1, 0, rho_w_avg_hi, vf);
        double zflux_hi = rho_w_avg_hi * ih;
        double rho_w_avg_lo = 0.5 * (Omega(i, j, k) + Omega(i, j - 1, k));
        il = interp_z(vt, v, i, j, k, 0, rho_w_avg_lo, vf);
        double zflux_lo = rho_w_avg_lo * il;
        double mfsq = 1 / (1.0 * 1.0);
        double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                   + (zflux_hi - zflux_lo) * dzInv_;
        rv_rhs(i, j, k) = -src;
    });
    loop(tbz, [&](int i, int j, int k) {
        double ih = 0., il = 0.;
        double rho_u_avg_hi = 0.5 * (rho_u(i + 1, j, k) + rho_u(i + 1, j, k - 1)) * 1.0;
        ih = interp_x(ht, w, i + 1, j, k, 0, rho_u_avg_hi, hf);
        double xflux_hi = rho_u_avg_hi * ih;
        double rho_u_avg_lo = 0.5 * (rho_u(i, j, k) + rho_u(i, j, k - 1)) * 1.0;
        il = interp_x(ht, w, i, j, k, 0, rho_u_avg_lo, hf);
        double xflux_lo = rho_u_avg_lo * il;
        double rho_v_avg_hi = 0.5 * (rho_v(i, j + 1, k) + rho_v(i, j + 1, k - 1)) * 1.0;
        ih = interp_y(ht, w, i, j + 1, k, 0, rho_v_avg_hi, hf);
        double yflux_hi = rho_v_avg_hi * ih;
        double rho_v_avg_lo = 0.5 * (rho_v(i, j, k) + rho_v(i, j, k - 1)) * 1.0;
        il = interp_y(ht, w, i, j, k, 0, rho_v_avg_lo, hf);
        double yflux_lo = rho_v_avg_lo * il;
        double zflux_hi, zflux_lo;
        if (k == hz) {
            zflux_hi = Omega(i, j, k) * w(i, j, k);
        } else {
            double rho_w_avg_hi = 0.5 * (Omega(i, j, k) + Omega(i, j, k + 1));
            if (k == hz - 1) {
                ih = interp_z_all(w, i, j, k + 1, rho_w_avg_hi, 0, ERFX_ADV_CENTERED2);
            } else if (k == hz - 2 || k == lz + 1) {
                if (vt != ERFX_ADV_CENTERED2 && vt != ERFX_ADV_UPWIND3)
                    ih = interp_z_all(w, i, j, k + 1, rho_w_avg_hi, 0, ERFX_ADV_CENTERED4);
                else
                    ih = interp_z_all(w, i, j, k + 1, rho_w_avg_hi, vf, vt);
            } else {
                ih = interp_z(vt, w, i, j, k + 1, 0, rho_w_avg_hi, vf);
            }
            zflux_hi = rho_w_avg_hi * ih;
        }
        if (k == lz) {
            zflux_lo = Omega(i, j, k) * w(i, j, k);
        } else {
            double rho_w_avg_lo = 0.5 * (Omega(i, j, k) + Omega(i, j, k - 1));
            if (k == lz + 1) {
                il = interp_z_all(w, i, j, k, rho_w_avg_lo, 0, ERFX_ADV_CENTERED2);
            } else if (k == lz + 2 || k == hz - 1) {
                if (vt != ERFX_ADV_CENTERED2 && vt != ERFX_ADV_UPWIND3)
                    il = interp_z_all(w, i, j, k, rho_w_avg_lo, 0, ERFX_ADV_CENTERED4);
                else
                    il = interp_z_all(w, i, j, k, rho_w_avg_lo, vf, vt);
            } else {
                il = interp_z(vt, w, i, j, k, 0, rho_w_avg_lo, vf);
            }
            zflux_lo = rho_w_avg_lo * il;
        }
        double mfsq = 1.0 * 1.0;
        double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                   + (zflux_hi - zflux_lo) * dzInv_;
        rw_rhs(i, j, k) = -src;
    });
}

}  // namespace orc

namespace orc {

// AdvectionSrcForMom with terrain metrics (ERF_AdvectionSrcForMom.cpp:204-288
// for C2/C2, ERF_AdvectionSrcForMom_T.H:22-330 otherwise), static terrain,
// unit map factors: the flat fluxes weighted by the face areas (ax, ay, az),
// the edge heights h_zeta, and divided by detJ averaged to the face.
void Level::advection_src_mom_T(const Box& tbx, const Box& tby, const Box& tbz, Arr ru_rhs, Arr rv_rhs,
                                Arr rw_rhs, Arr u, Arr v, Arr w, Arr rho_u, Arr rho_v, Arr Omega) {
    const double dxInv_ = dxInv[0], dyInv_ = dxInv[1], dzInv_ = dxInv[2];
    const int hz = L.nz, lz = 0;
    const int ht = cfg.dycore_horiz_adv, vt = cfg.dycore_vert_adv;
    const double hf = cfg.dycore_horiz_upw_frac, vf = cfg.dycore_vert_upw_frac;
    Arr z = A(z_nd), Ax = A(ax), Ay = A(ay), Az = A(az), J = A(detJ);
    const double dzi = dzInv_;
    if (ht == ERFX_ADV_CENTERED2 && vt == ERFX_ADV_CENTERED2) {
        loop(tbx, [&](int i, int j, int k) {
            double xflux_hi = 0.25 * (rho_u(i, j, k) * 1.0 + rho_u(i + 1, j, k) * 1.0) * (u(i + 1, j, k) + u(i, j, k)) *
                              0.5 * (Ax(i, j, k) + Ax(i + 1, j, k));
            double xflux_lo = 0.25 * (rho_u(i, j, k) * 1.0 + rho_u(i - 1, j, k) * 1.0) * (u(i - 1, j, k) + u(i, j, k)) *
                              0.5 * (Ax(i, j, k) + Ax(i - 1, j, k));
            double met_h_zeta_yhi = met::h_zeta_EdgeK(z, i, j + 1, k, dzi);
            double yflux_hi = 0.25 * (rho_v(i, j + 1, k) * 1.0 + rho_v(i - 1, j + 1, k) * 1.0) *
                              (u(i, j + 1, k) + u(i, j, k)) * met_h_zeta_yhi;
            double met_h_zeta_ylo = met::h_zeta_EdgeK(z, i, j, k, dzi);
            double yflux_lo = 0.25 * (rho_v(i, j, k) * 1.0 + rho_v(i - 1, j, k) * 1.0) *
                              (u(i, j - 1, k) + u(i, j, k)) * met_h_zeta_ylo;
            double zflux_hi = 0.25 * (Omega(i, j, k + 1) + Omega(i - 1, j, k + 1)) * (u(i, j, k + 1) + u(i, j, k)) *
                              0.5 * (Az(i, j, k + 1) + Az(i - 1, j, k + 1));
            double zflux_lo = 0.25 * (Omega(i, j, k) + Omega(i - 1, j, k)) * (u(i, j, k - 1) + u(i, j, k)) *
                              0.5 * (Az(i, j, k) + Az(i - 1, j, k));
            double mfsq = 1.0 * 1.0;
            double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                       + (zflux_hi - zflux_lo) * dzInv_;
            ru_rhs(i, j, k) = -src / (0.5 * (J(i, j, k) + J(i - 1, j, k)));
        });
        loop(tby, [&](int i, int j, int k) {
            double met_h_zeta_xhi = met::h_zeta_EdgeK(z, i + 1, j, k, dzi);
            double xflux_hi = 0.25 * (rho_u(i + 1, j, k) * 1.0 + rho_u(i + 1, j - 1, k) * 1.0) *
                              (v(i + 1, j, k) + v(i, j, k)) * met_h_zeta_xhi;
            double met_h_zeta_xlo = met::h_zeta_EdgeK(z, i, j, k, dzi);
            double xflux_lo = 0.25 * (rho_u(i, j, k) * 1.0 + rho_u(i, j - 1, k) * 1.0) *
                              (v(i - 1, j, k) + v(i, j, k)) * met_h_zeta_xlo;
            double yflux_hi = 0.25 * (rho_v(i, j + 1, k) * 1.0 + rho_v(i, j, k) * 1.0) * (v(i, j + 1, k) + v(i, j, k)) *
                              0.5 * (Ay(i, j, k) + Ay(i, j + 1, k));
            double yflux_lo = 0.25 * (rho_v(i, j, k) * 1.0 + rho_v(i, j - 1, k) * 1.0) * (v(i, j - 1, k) + v(i, j, k)) *
                              0.5 * (Ay(i, j, k) + Ay(i, j - 1, k));
            double zflux_hi = 0.25 * (Omega(i, j, k + 1) + Omega(i, j - 1, k + 1)) * (v(i, j, k + 1) + v(i, j, k)) *
                              0.5 * (Az(i, j, k + 1) + Az(i, j - 1, k + 1));
            double zflux_lo = 0.25 * (Omega(i, j, k) + Omega(i, j - 1, k)) * (v(i, j, k - 1) + v(i, j, k)) *
                              0.5 * (Az(i, j, k) + Az(i, j - 1, k));
            double mfsq = 1.0 * 1.0;
            double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                       + (zflux_hi - zflux_lo) * dzInv_;
            rv_rhs(i, j, k) = -src / (0.5 * (J(i, j, k) + J(i, j - 1, k)));
        });
        loop(tbz, [&](int i, int j, int k) {
            double met_h_zeta_xhi = met::h_zeta_EdgeJ(z, i + 1, j, k, dzi);
            double xflux_hi = 0.25 * (rho_u(i + 1, j, k) + rho_u(i + 1, j, k - 1)) * 1.0 *
                              (w(i + 1, j, k) + w(i, j, k)) * met_h_zeta_xhi;
            double met_h_zeta_xlo = met::h_zeta_EdgeJ(z, i, j, k, dzi);
            double xflux_lo = 0.25 * (rho_u(i, j, k) + rho_u(i, j, k - 1)) * 1.0 *
                              (w(i - 1, j, k) + w(i, j, k)) * met_h_zeta_xlo;
            double met_h_zeta_yhi = met::h_zeta_EdgeI(z, i, j + 1, k, dzi);
            double yflux_hi = 0.25 * (rho_v(i, j + 1, k) + rho_v(i, j + 1, k - 1)) * 1.0 *
                              (w(i, j + 1, k) + w(i, j, k)) * met_h_zeta_yhi;
            double met_h_zeta_ylo = met::h_zeta_EdgeI(z, i, j, k, dzi);
            double yflux_lo = 0.25 * (rho_v(i, j, k) + rho_v(i, j, k - 1)) * 1.0 *
                              (w(i, j - 1, k) + w(i, j, k)) * met_h_zeta_ylo;
            double zflux_lo = 0.25 * (Omega(i, j, k) + Omega(i, j, k - 1)) * (w(i, j, k) + w(i, j, k - 1));
            double zflux_hi = (k == hz) ? Omega(i, j, k) * w(i, j, k) * Az(i, j, k)
                                        : 0.25 * (Omega(i, j, k) + Omega(i, j, k + 1)) * (w(i, j, k) + w(i, j, k + 1)) *
                                              0.5 * (Az(i, j, k) + Az(i, j, k + 1));
            double mfsq = 1.0 * 1.0;
            double src = (xflux_hi - xflux_lo) * dxInv_ * mfsq + (yflux_hi - yflux_lo) * dyInv_ * mfsq
                       + (zflux_hi - zflux_lo) * dzInv_;
            rw_rhs(i, j, k) = -src / (0.5 * (J(i, j, k) + J(i, j, k - 1)));
        });
        return;
    }
    // higher order: AdvectionSrcForXMom / YMom / ZMom with metrics
    loop(tbx, [&](int i, int j, int k) {
        double rho_u_avg_hi = 0.5 * (rho_u(i + 1, j, k) * 1.0 + rho_u(i, j, k) * 1.0);
        double interp_hi = interp_x(ht, u, i + 1, j, k, 0, rho_u_avg_hi, hf);
        double rho_u_avg_lo = 0.5 * (rho_u(i - 1, j, k) * 1.0 + rho_u(i, j, k) * 1.0);
        double interp_lo = interp_x(ht, u, i, j, k, 0, rho_u_avg_lo, hf);
        double centFluxXXNext = rho_u_avg_hi * interp_hi * 0.5 * (Ax(i, j, k) + Ax(i + 1, j, k));
        double centFluxXXPrev = rho_u_avg_lo * interp_lo * 0.5 * (Ax(i, j, k) + Ax(i - 1, j, k));
        double rho_v_avg_hi = 0.5 * (rho_v(i, j + 1, k) * 1.0 + rho_v(i - 1, j + 1, k) * 1.0);
        interp_hi = interp_y(ht, u, i, j + 1, k, 0, rho_v_avg_hi, hf);
        double rho_v_avg_lo = 0.5 * (rho_v(i, j, k) * 1.0 + rho_v(i - 1, j, k) * 1.0);
        interp_lo = interp_y(ht, u, i, j, k, 0, rho_v_avg_lo, hf);
        double edgeFluxXYNext = rho_v_avg_hi * interp_hi * met::h_zeta_EdgeK(z, i, j + 1, k, dzi);
        double edgeFluxXYPrev = rho_v_avg_lo * interp_lo * met::h_zeta_EdgeK(z, i, j, k, dzi);
        double Omega_avg_hi = 0.5 * (Omega(i, j, k + 1) + Omega(i - 1, j, k + 1)) * 0.5 * (Az(i, j, k + 1) + Az(i - 1, j, k + 1));
        double Omega_avg_lo = 0.5 * (Omega(i, j, k) + Omega(i - 1, j, k)) * 0.5 * (Az(i, j, k) + Az(i - 1, j, k));
        interp_hi = interp_z(vt, u, i, j, k + 1, 0, Omega_avg_hi, vf);
        interp_lo = interp_z(vt, u, i, j, k, 0, Omega_avg_lo, vf);
        double edgeFluxXZNext = Omega_avg_hi * interp_hi;
        double edgeFluxXZPrev = Omega_avg_lo * interp_lo;
        double mfsq = 1 / (1.0 * 1.0);
        double src = (centFluxXXNext - centFluxXXPrev) * dxInv_ * mfsq + (edgeFluxXYNext - edgeFluxXYPrev) * dyInv_ * mfsq
                   + (edgeFluxXZNext - edgeFluxXZPrev) * dzInv_;
        src /= 0.5 * (J(i, j, k) + J(i - 1, j, k));
        ru_rhs(i, j, k) = -src;
    });
    loop(tby, [&](int i, int j, int k) {
        double rho_u_avg_hi = 0.5 * (rho_u(i + 1, j, k) * 1.0 + rho_u(i + 1, j - 1, k) * 1.0);
        double rho_u_avg_lo = 0.5 * (rho_u(i, j, k) * 1.0 + rho_u(i, j - 1, k) * 1.0);
        double interp_hi = interp_x(ht, v, i + 1, j, k, 0, rho_u_avg_hi, hf);
        double interp_lo = interp_x(ht, v, i, j, k, 0, rho_u_avg_lo, hf);
        double edgeFluxYXNext = rho_u_avg_hi * interp_hi * met::h_zeta_EdgeK(z, i + 1, j, k, dzi);
        double edgeFluxYXPrev = rho_u_avg_lo * interp_lo * met::h_zeta_EdgeK(z, i, j, k, dzi);
        double rho_v_avg_hi = 0.5 * (rho_v(i, j, k) * 1.0 + rho_v(i, j + 1, k) * 1.0);
        double rho_v_avg_lo = 0.5 * (rho_v(i, j, k) * 1.0 + rho_v(i, j - 1, k) * 1.0);
        interp_hi = interp_y(ht, v, i, j + 1, k, 0, rho_v_avg_hi, hf);
        interp_lo = interp_y(ht, v, i, j, k, 0, rho_v_avg_lo, hf);
        double centFluxYYNext = rho_v_avg_hi * 0.5 * (Ay(i, j, k) + Ay(i, j + 1, k)) * interp_hi;
        double centFluxYYPrev = rho_v_avg_lo * 0.5 * (Ay(i, j, k) + Ay(i, j - 1, k)) * interp_lo;
        double Omega_avg_hi = 0.5 * (Omega(i, j, k + 1) + Omega(i, j - 1, k + 1)) * 0.5 * (Az(i, j, k + 1) + Az(i, j - 1, k + 1));
        double Omega_avg_lo = 0.5 * (Omega(i, j, k) + Omega(i, j - 1, k)) * 0.5 * (Az(i, j, k) + Az(i, j - 1, k));
        interp_hi = interp_z(vt, v, i, j, k + 1, 0, Omega_avg_hi, vf);
        interp_lo = interp_z(vt, v, i, j, k, 0, Omega_avg_lo, vf);
        double edgeFluxYZNext = Omega_avg_hi * interp_hi;
        double edgeFluxYZPrev = Omega_avg_lo * interp_lo;
        double mfsq = 1 / (1.0 * 1.0);
        double src = (edgeFluxYXNext - edgeFluxYXPrev) * dxInv_ * mfsq + (centFluxYYNext - centFluxYYPrev) * dyInv_ * mfsq
                   + (edgeFluxYZNext - edgeFluxYZPrev) * dzInv_;
        src /= 0.5 * (J(i, j, k) + J(i, j - 1, k));
        rv_rhs(i, j, k) = -src;
    });
    loop(tbz, [&](int i, int j, int k) {
        double rho_u_avg_hi = 0.5 * (rho_u(i + 1, j, k) + rho_u(i + 1, j, k - 1)) * 1.0;
        double rho_u_avg_lo = 0.5 * (rho_u(i, j, k) + rho_u(i, j, k - 1)) * 1.0;
        double interp_hi = interp_x(ht, w, i + 1, j, k, 0, rho_u_avg_hi, hf);
        double interp_lo = interp_x(ht, w, i, j, k, 0, rho_u_avg_lo, hf);
        double edgeFluxZXNext = rho_u_avg_hi * interp_hi * met::h_zeta_EdgeJ(z, i + 1, j, k, dzi);
        double edgeFluxZXPrev = rho_u_avg_lo * interp_lo * met::h_zeta_EdgeJ(z, i, j, k, dzi);
        double rho_v_avg_hi = 0.5 * (rho_v(i, j + 1, k) + rho_v(i, j + 1, k - 1)) * 1.0;
        interp_hi = interp_y(ht, w, i, j + 1, k, 0, rho_v_avg_hi, hf);
        double rho_v_avg_lo = 0.5 * (rho_v(i, j, k) + rho_v(i, j, k - 1)) * 1.0;
        interp_lo = interp_y(ht, w, i, j, k, 0, rho_v_avg_lo, hf);
        double edgeFluxZYNext = rho_v_avg_hi * interp_hi * met::h_zeta_EdgeI(z, i, j + 1, k, dzi);
        double edgeFluxZYPrev = rho_v_avg_lo * interp_lo * met::h_zeta_EdgeI(z, i, j, k, dzi);
        double Omega_avg_hi = (k == hz) ? Omega(i, j, k) * Az(i, j, k)
                                        : 0.25 * (Omega(i, j, k) + Omega(i, j, k + 1)) * (Az(i, j, k) + Az(i, j, k + 1));
        double centFluxZZNext = Omega_avg_hi;
        if (k == hz) {
            centFluxZZNext *= w(i, j, k);
        } else {
            if (k == hz - 1) {
                interp_hi = interp_z_all(w, i, j, k + 1, Omega_avg_hi, 0, ERFX_ADV_CENTERED2);
            } else if (k == hz - 2 || k == lz + 1) {
                if (vt != ERFX_ADV_CENTERED2 && vt != ERFX_ADV_UPWIND3)
                    interp_hi = interp_z_all(w, i, j, k + 1, Omega_avg_hi, 0, ERFX_ADV_CENTERED4);
                else
                    interp_hi = interp_z_all(w, i, j, k + 1, Omega_avg_hi, vf, vt);
            } else {
                interp_hi = interp_z(vt, w, i, j, k + 1, 0, Omega_avg_hi, vf);
            }
            centFluxZZNext *= interp_hi;
        }
        double Omega_avg_lo = (k == 0) ? Omega(i, j, k) * Az(i, j, k)
                                       : 0.25 * (Omega(i, j, k) + Omega(i, j, k - 1)) * (Az(i, j, k) + Az(i, j, k - 1));
        double centFluxZZPrev = Omega_avg_lo;
        if (k == 0) {
            centFluxZZPrev *= w(i, j, k);
        } else {
            if (k == lz + 1) {
                interp_lo = interp_z_all(w, i, j, k, Omega_avg_lo, 0, ERFX_ADV_CENTERED2);
            } else if (k == lz + 2 || k == hz - 1) {
                if (vt != ERFX_ADV_CENTERED2 && vt != ERFX_ADV_UPWIND3)
                    interp_lo = interp_z_all(w, i, j, k, Omega_avg_lo, 0, ERFX_ADV_CENTERED4);
                else
                    interp_lo = interp_z_all(w, i, j, k, Omega_avg_lo, vf, vt);
            } else {
                interp_lo = interp_z(vt, w, i, j, k, 0, Omega_avg_lo, vf);
            }
            centFluxZZPrev *= interp_lo;
        }
        double mfsq = 1.0 * 1.0;
        double src = (edgeFluxZXNext - edgeFluxZXPrev) * dxInv_ * mfsq + (edgeFluxZYNext - edgeFluxZYPrev) * dyInv_ * mfsq
                   + (centFluxZZNext - centFluxZZPrev) * dzInv_;
        double denom = 0.5 * (J(i, j, k) + J(i, j, k - 1));
        src /= denom;
        rw_rhs(i, j, k) = -src;
    });
}

}  // namespace orc
