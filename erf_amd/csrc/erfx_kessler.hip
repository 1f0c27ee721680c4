// erfx_kessler.hip -- Kessler warm-rain microphysics, once per step after the
// dycore (ERF::advance_microphysics, Source/TimeIntegration/ERF_advance_microphysics.cpp:5-16).
//
// ERF copies the state into nine micro fields (Copy_State_to_Micro,
// Source/Microphysics/Kessler/ERF_Init_Kessler.cpp:64-99), advances them
// (Kessler::AdvanceKessler, ERF_Kessler.cpp:11-250) and copies back
// (Copy_Micro_to_State, ERF_Update_Kessler.cpp:14-40).  Here the micro values
// are recomputed in registers from cons where they are needed, so the step
// is two launches on cons alone:
//   KS1  z-faces: sedimentation flux of rain (ERF_Kessler.cpp:27-69), with
//        the tiny-flux cut that the cell loop applies to each face (:162-163)
//   KS2  cells: saturation adjustment, rain evaporation, autoconversion,
//        accretion, sedimentation, theta update (:71-184; no-rain :187-248),
//        and the copy back to rho*theta and rho*q
// followed by the periodic FillBoundary of every cons component.
// Every expression keeps the reference's operation order.
#include "erfx_common.h"
#include "erfx_internal.h"

namespace erfx {

// Flatau et al. (1992, Table 4) fits over liquid water in Horner form
// (ERF_Microphysics_Utils.H:51-71, 104-132); ERF aborts below -85 C, which
// is reported through *bad.
__device__ __forceinline__ double horner8(const double* a, double x) {
    double r = a[8];
#pragma unroll
    for (int n = 7; n >= 0; --n) r = a[n] + x * r;
    return r;
}
__device__ __forceinline__ double esatw_dev(double t, int* bad) {
    const double a[9] = {6.11239921,       0.443987641,     0.142986287e-1,
                         0.264847430e-3,   0.302950461e-5,  0.206739458e-7,
                         0.640689451e-10, -0.952447341e-13, -0.976195544e-15};
    double dtt_in = t - 273.16;
    if (!(dtt_in > -85)) *bad = 1;
    double dtt = (70. < dtt_in) ? 70. : dtt_in;     // std::min(dtt_in, 70.)
    return horner8(a, dtt);
}
__device__ __forceinline__ double dtesatw_dev(double t, int* bad) {
    const double a[9] = {0.443956472,      0.285976452e-1,  0.794747212e-3,
                         0.121167162e-4,   0.103167413e-6,  0.385208005e-9,
                        -0.604119582e-12, -0.792933209e-14, -0.599634321e-17};
    double dtt = t - 273.16;
    if (!(dtt > -85)) *bad = 1;
    if (dtt >= 70.) return 0.0;
    return horner8(a, dtt);
}
// std::max / std::min of the reference, NaN and signed-zero behaviour included
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

void kessler_advance(const Params& P, double* cons, double* fz, double* rain_accum, double dt, hipStream_t s,
                     int* bad) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sz = L.sz;
    const double dz = P.cfg.dz;
    const double fac_cond = lcond / P.cfg.c_p;    // Kessler::Define (ERF_Kessler.H:65)
    const bool rain = P.cfg.moisture_type == ERFX_MOIST_KESSLER;
    double* rho = cons;
    double* rt = cons + ne;
    double* q1 = cons + ERFX_RHOQ1 * ne;
    double* q2 = cons + ERFX_RHOQ2 * ne;
    double* q3 = cons + ERFX_RHOQ3 * ne;
    if (rain) {
        // KS1: faces 0..nz of the valid columns
        launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz), s, [=] __device__(int i, int j, int k) {
            const long q = L.ix(i, j, k);
            double rho_avg, qp_avg;
            if (k == 0) {
                rho_avg = rho[q]; qp_avg = q3[q] / rho[q];
            } else if (k == nz) {
                rho_avg = rho[q - sz]; qp_avg = q3[q - sz] / rho[q - sz];
            } else {
                double rl = rho[q - sz], rh = rho[q];
                rho_avg = 0.5 * (rl + rh);
                qp_avg = 0.5 * (q3[q - sz] / rl + q3[q] / rh);
            }
            qp_avg = smax(0.0, qp_avg);
            double V_terminal = 36.34 * erfx_pow(rho_avg * 0.001 * qp_avg, 0.1346) * erfx_pow(rho_avg / 1.16, -0.5);
            double f = rho_avg * V_terminal * qp_avg;
            if (k == 0 && rain_accum) {
                const long p2 = (long)(i + L.ng) + (long)(j + L.ng) * L.sy;
                rain_accum[p2] = rain_accum[p2] + rho_avg * qp_avg * V_terminal * dt / 1000.0 * 1000.0;
            }
            if (fabs(f) < 1e-14) f = 0.0;
            fz[q] = f;
        });
    }
    const double LvOcp = L_v / Cp_d;
    // KS2: cells
    launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        const double r = rho[q], rtv = rt[q];
        double th = rtv / r;
        double qv = q1[q] / r;
        double qc = q2[q] / r;
        double qp = q3[q] / r;
        const double tabs = eos_T(r, rtv, qv);
        const double pressure = eos_p_of_rt_qv(rtv, qv) / 100.;
        int flag = 0;
        if (rain) {
            double dJinv = 1.0;
            qv = smax(0.0, qv);
            qc = smax(0.0, qc);
            qp = smax(0.0, qp);
            double esat = esatw_dev(tabs, &flag);
            double qsat = Rd_on_Rv * esat / smax(esat, pressure - esat);
            double dtqsat = Rd_on_Rv * dtesatw_dev(tabs, &flag) / pressure;
            if (qsat <= 0.0) qsat = 0.0;
            double dq_clwater_to_rain = 0.0, dq_rain_to_vapor = 0.0;
            double dq_vapor_to_clwater = 0.0, dq_clwater_to_vapor = 0.0;
            double fac = 1.0 + LvOcp * dtqsat;
            if (qv > qsat) dq_vapor_to_clwater = smin(qv, (qv - qsat) / (1.0 + fac));
            if (qv < qsat && qc > 0.0) dq_clwater_to_vapor = smin(qc, (qsat - qv) / (1.0 + fac));
            if (qp > 0.0 && qv < qsat) {
                double C = 1.6 + 124.9 * erfx_pow(0.001 * r * qp, 0.2046);
                dq_rain_to_vapor = 1.0 / (0.001 * r) * (1.0 - qv / qsat) * C * erfx_pow(0.001 * r * qp, 0.525) /
                                   (5.4e5 + 2.55e6 / (pressure * qsat)) * dt;
                dq_rain_to_vapor = smin(qp, dq_rain_to_vapor);
            }
            if (qc > 0.0) {
                double qcc = qc;
                double auto_r = 0.0;
                if (qcc > qcw0) auto_r = alphaelq;
                double accrr = 2.2 * erfx_pow(qp, 0.875);
                dq_clwater_to_rain = dt * (accrr * qcc + auto_r * (qcc - qcw0));
                dq_clwater_to_rain = smin(dq_clwater_to_rain, qc);
            }
            double dq_sed = dt * dJinv * (1.0 / r) * (fz[q + sz] - fz[q]) / dz;
            if (fabs(dq_sed) < 1e-14) dq_sed = 0.0;
            qv += -dq_vapor_to_clwater + dq_clwater_to_vapor + dq_rain_to_vapor;
            qc += dq_vapor_to_clwater - dq_clwater_to_vapor - dq_clwater_to_rain;
            qp += dq_sed + dq_clwater_to_rain - dq_rain_to_vapor;
            double theta_over_T = th / tabs;
            th += theta_over_T * fac_cond * (dq_vapor_to_clwater - dq_clwater_to_vapor - dq_rain_to_vapor);
            qv = smax(0.0, qv);
            qc = smax(0.0, qc);
            qp = smax(0.0, qp);
        } else {
            qc = smax(0.0, qc);
            double esat = esatw_dev(tabs, &flag);
            double qsat = Rd_on_Rv * esat / smax(esat, pressure - esat);
            double dtqsat = Rd_on_Rv * dtesatw_dev(tabs, &flag) / pressure;
            double dq_vapor_to_clwater = 0.0, dq_clwater_to_vapor = 0.0;
            double fac = 1.0 + LvOcp * dtqsat;
            if (qv > qsat) dq_vapor_to_clwater = smin(qv, (qv - qsat) / (1.0 + fac));
            if (qv < qsat && qc > 0.0) dq_clwater_to_vapor = smin(qc, (qsat - qv) / (1.0 + fac));
            qv += -dq_vapor_to_clwater + dq_clwater_to_vapor;
            qc += dq_vapor_to_clwater - dq_clwater_to_vapor;
            double theta_over_T = th / tabs;
            th += theta_over_T * fac_cond * (dq_vapor_to_clwater - dq_clwater_to_vapor);
            qv = smax(0.0, qv);
            qc = smax(0.0, qc);
        }
        if (flag) *bad = 1;
        rt[q] = r * th;
        q1[q] = r * qv;
        q2[q] = r * qc;
        q3[q] = r * qp;
    });
    fill_boundary_periodic(P, cons, P.cfg.ncons, P.ng_cons, -1, s);
}

}  // namespace erfx
