// kessler.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  Kessler warm-rain
// microphysics, once per step after the dycore, and the plane averages of the
// moist buoyancy.
//
// ERF::advance_microphysics (Source/TimeIntegration/ERF_advance_microphysics.cpp:5-16)
// calls, on the new state:
//   Copy_State_to_Micro (Source/Microphysics/Kessler/ERF_Init_Kessler.cpp:64-99),
//   Kessler::AdvanceKessler (ERF_Kessler.cpp:11-250) with the step's dt,
//   Copy_Micro_to_State (ERF_Update_Kessler.cpp:14-40), which ends with a
//   periodic FillBoundary of every cons component.
#include "oracle.hpp"

namespace orc {

// Saturation vapour pressure over liquid water and its temperature
// derivative [mbar, mbar/K]: the Flatau et al. (1992, Table 4) polynomial fits
// of ERF_Microphysics_Utils.H:51-71 and :104-132, evaluated in Horner form in
// the reference's order.  ERF aborts below -85 C; above 70 C the pressure is
// held constant, so its derivative is 0.
static double horner8(const double* a, double x) {
    double r = a[8];
    for (int n = 7; n >= 0; --n) r = a[n] + x * r;
    return r;
}
double esatw(double t) {
    static const double a[9] = {6.11239921,       0.443987641,     0.142986287e-1,
                                0.264847430e-3,   0.302950461e-5,  0.206739458e-7,
                                0.640689451e-10, -0.952447341e-13, -0.976195544e-15};
    double dtt_in = t - 273.16;
    if (!(dtt_in > -85)) throw std::runtime_error("esatw: temperature below -85 C");
    double dtt = std::min(dtt_in, 70.);
    return horner8(a, dtt);
}
double dtesatw(double t) {
    static const double a[9] = {0.443956472,      0.285976452e-1,  0.794747212e-3,
                                0.121167162e-4,   0.103167413e-6,  0.385208005e-9,
                               -0.604119582e-12, -0.792933209e-14, -0.599634321e-17};
    double dtt = t - 273.16;
    if (!(dtt > -85)) throw std::runtime_error("dtesatw: temperature below -85 C");
    if (dtt >= 70.) return 0.0;
    return horner8(a, dtt);
}
// erf_qsatw / erf_dtqsatw (ERF_Microphysics_Utils.H:141-156); p in mbar
double qsatw(double t, double p) {
    double e = esatw(t);
    return Rd_on_Rv * e / std::max(e, p - e);
}
double dtqsatw(double t, double p) { return Rd_on_Rv * dtesatw(t) / p; }

void Level::kessler_advance(Arr c, double dt) {
    if (!moist) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const double dz = cfg.dz;
    const double fac_cond = lcond / cfg.c_p;       // Kessler::Define (ERF_Kessler.H:65)
    const bool rain = cfg.moisture_type == ERFX_MOIST_KESSLER;
    const Box vb(0, 0, 0, nx - 1, ny - 1, nz - 1);
    // Copy_State_to_Micro: one micro field per quantity on the valid box
    Field f_rho, f_th, f_qv, f_qc, f_qp, f_T, f_p, f_fz;
    for (Field* f : {&f_rho, &f_th, &f_qv, &f_qc, &f_qp, &f_T, &f_p, &f_fz}) f->alloc(L, 1);
    Arr rho = f_rho.a(L), th = f_th.a(L), qv = f_qv.a(L), qc = f_qc.a(L), qp = f_qp.a(L);
    Arr tabs = f_T.a(L), pres = f_p.a(L), fz = f_fz.a(L);
    loop(vb, [&](int i, int j, int k) {
        double r = c(i, j, k, ERFX_RHO), rt = c(i, j, k, ERFX_RHOTHETA);
        rho(i, j, k) = r;
        th(i, j, k) = rt / r;
        qv(i, j, k) = c(i, j, k, ERFX_RHOQ1) / r;
        qc(i, j, k) = c(i, j, k, ERFX_RHOQ2) / r;
        qp(i, j, k) = c(i, j, k, ERFX_RHOQ3) / r;
        tabs(i, j, k) = getTgivenRandRTh(r, rt, qv(i, j, k));
        pres(i, j, k) = getPgivenRTh(rt, qv(i, j, k)) / 100.;
    });
    const double LvOcp = L_v / Cp_d;
    if (rain) {
        // sedimentation flux on the z-faces of the valid box (:27-69)
        loop(Box(0, 0, 0, nx - 1, ny - 1, nz), [&](int i, int j, int k) {
            double rho_avg, qp_avg;
            if (k == 0) { rho_avg = rho(i, j, k); qp_avg = qp(i, j, k); }
            else if (k == nz) { rho_avg = rho(i, j, k - 1); qp_avg = qp(i, j, k - 1); }
            else {
                rho_avg = 0.5 * (rho(i, j, k - 1) + rho(i, j, k));
                qp_avg = 0.5 * (qp(i, j, k - 1) + qp(i, j, k));
            }
            qp_avg = std::max(0.0, qp_avg);
            double V_terminal = 36.34 * erfx_pow(rho_avg * 0.001 * qp_avg, 0.1346) * erfx_pow(rho_avg / 1.16, -0.5);
            fz(i, j, k) = rho_avg * V_terminal * qp_avg;
            if (k == 0) {
                double& ra = rain_accum[(size_t)(i + L.ng) + (size_t)(j + L.ng) * L.sy];
                ra = ra + rho_avg * qp_avg * V_terminal * dt / 1000.0 * 1000.0;
            }
        });
        // the faces' tiny-flux cut of :162-163 depends only on the face value,
        // so applying it to every face first gives what each cell reads
        loop(Box(0, 0, 0, nx - 1, ny - 1, nz), [&](int i, int j, int k) {
            if (std::fabs(fz(i, j, k)) < 1e-14) fz(i, j, k) = 0.0;
        });
        // cell update (:71-184)
        loop(vb, [&](int i, int j, int k) {
            double dJinv = 1.0;
            double& v = qv(i, j, k); double& cw = qc(i, j, k); double& r = qp(i, j, k);
            v = std::max(0.0, v);
            cw = std::max(0.0, cw);
            r = std::max(0.0, r);
            double pressure = pres(i, j, k);
            double qsat = qsatw(tabs(i, j, k), pressure);
            double dtqsat = dtqsatw(tabs(i, j, k), pressure);
            if (qsat <= 0.0) qsat = 0.0;
            double dq_clwater_to_rain = 0.0, dq_rain_to_vapor = 0.0;
            double dq_vapor_to_clwater = 0.0, dq_clwater_to_vapor = 0.0;
            double fac = 1.0 + LvOcp * dtqsat;
            if (v > qsat) dq_vapor_to_clwater = std::min(v, (v - qsat) / (1.0 + fac));
            if (v < qsat && cw > 0.0) dq_clwater_to_vapor = std::min(cw, (qsat - v) / (1.0 + fac));
            const double rh = rho(i, j, k);
            if (r > 0.0 && v < qsat) {
                double C = 1.6 + 124.9 * erfx_pow(0.001 * rh * r, 0.2046);
                dq_rain_to_vapor = 1.0 / (0.001 * rh) * (1.0 - v / qsat) * C * erfx_pow(0.001 * rh * r, 0.525) /
                                   (5.4e5 + 2.55e6 / (pressure * qsat)) * dt;
                dq_rain_to_vapor = std::min(r, dq_rain_to_vapor);
            }
            if (cw > 0.0) {
                double qcc = cw;
                double auto_r = 0.0;
                if (qcc > qcw0) auto_r = alphaelq;
                double accrr = 2.2 * erfx_pow(r, 0.875);
                dq_clwater_to_rain = dt * (accrr * qcc + auto_r * (qcc - qcw0));
                dq_clwater_to_rain = std::min(dq_clwater_to_rain, cw);
            }
            if (terrain) dJinv = 1.0 / A(detJ)(i, j, k);
            double dq_sed = dt * dJinv * (1.0 / rh) * (fz(i, j, k + 1) - fz(i, j, k)) / dz;
            if (std::fabs(dq_sed) < 1e-14) dq_sed = 0.0;
            v += -dq_vapor_to_clwater + dq_clwater_to_vapor + dq_rain_to_vapor;
            cw += dq_vapor_to_clwater - dq_clwater_to_vapor - dq_clwater_to_rain;
            r += dq_sed + dq_clwater_to_rain - dq_rain_to_vapor;
            double theta_over_T = th(i, j, k) / tabs(i, j, k);
            th(i, j, k) += theta_over_T * fac_cond * (dq_vapor_to_clwater - dq_clwater_to_vapor - dq_rain_to_vapor);
            v = std::max(0.0, v);
            cw = std::max(0.0, cw);
            r = std::max(0.0, r);
        });
    } else {
        // Kessler_NoRain (:187-248): saturation adjustment only
        loop(vb, [&](int i, int j, int k) {
            double& v = qv(i, j, k); double& cw = qc(i, j, k);
            cw = std::max(0.0, cw);
            double pressure = pres(i, j, k);
            double qsat = qsatw(tabs(i, j, k), pressure);
            double dtqsat = dtqsatw(tabs(i, j, k), pressure);
            double dq_vapor_to_clwater = 0.0, dq_clwater_to_vapor = 0.0;
            double fac = 1.0 + LvOcp * dtqsat;
            if (v > qsat) dq_vapor_to_clwater = std::min(v, (v - qsat) / (1.0 + fac));
            if (v < qsat && cw > 0.0) dq_clwater_to_vapor = std::min(cw, (qsat - v) / (1.0 + fac));
            v += -dq_vapor_to_clwater + dq_clwater_to_vapor;
            cw += dq_vapor_to_clwater - dq_clwater_to_vapor;
            double theta_over_T = th(i, j, k) / tabs(i, j, k);
            th(i, j, k) += theta_over_T * fac_cond * (dq_vapor_to_clwater - dq_clwater_to_vapor);
            v = std::max(0.0, v);
            cw = std::max(0.0, cw);
        });
    }
    // Copy_Micro_to_State, then FillBoundary of cons
    loop(vb, [&](int i, int j, int k) {
        double r = rho(i, j, k);
        c(i, j, k, ERFX_RHOTHETA) = r * th(i, j, k);
        c(i, j, k, ERFX_RHOQ1) = r * qv(i, j, k);
        c(i, j, k, ERFX_RHOQ2) = r * qc(i, j, k);
        c(i, j, k, ERFX_RHOQ3) = r * qp(i, j, k);
    });
    fill_boundary_periodic(c, ncons, ng_cons, -1);
}

// Error-free double-double accumulation: two-sum of the high parts, the low
// parts added to the error term, renormalised.
struct DDs { double hi = 0.0, lo = 0.0; };
static inline void dd_acc(DDs& a, double bh, double bl) {
    double s = a.hi + bh;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (bh - bb);
    e += a.lo + bl;
    double h = s + e;
    a.lo = e - (h - s);
    a.hi = h;
}

// PlaneAverage::compute_averages (Source/Utils/ERF_PlaneAverage.H:184-239) of
// the rows the moist buoyancy reads: each cell adds value * (1 / ncell_plane).
// ERF's reduction order is the device's; here each level's terms are summed in
// double-double, so the result is their exactly rounded sum in any order.
void Level::plane_averages(StateA data, Arr prim_, std::vector<double>& avg) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const double denom = 1.0 / (double)((long)nx * ny);
    avg.assign(5 * (size_t)nz, 0.0);
    for (int k = 0; k < nz; ++k) {
        DDs acc[5];
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double fac = 1.0;
                double v[5] = {data.cons(i, j, k, ERFX_RHO), prim_(i, j, k, 0), 0.0, 0.0, 0.0};
                for (int q = 0; q < n_qstate && q < 3; ++q) v[2 + q] = prim_(i, j, k, ERFX_RHOQ1 - 1 + q);
                for (int r = 0; r < 5; ++r) dd_acc(acc[r], v[r] * denom * fac, 0.0);
            }
        for (int r = 0; r < 5; ++r) avg[(size_t)r * nz + k] = acc[r].hi;
    }
}

}  // namespace orc
