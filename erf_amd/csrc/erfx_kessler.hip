// erfx_kessler.hip -- Kessler warm-rain microphysics, once per step after the
// dycore (ERF::advance_microphysics, Source/TimeIntegration/ERF_advance_microphysics.cpp:5-16),
// and the same step with the learned tendencies (erfx_mlp.hip) in place of
// Kessler's local processes.
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
#include "erfx_mlp.h"

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

// the micro variables of one cell (Copy_State_to_Micro, ERF_Init_Kessler.cpp:84-97)
struct Micro { double r, th, qv, qc, qp, tabs, pres; };
__device__ __forceinline__ Micro micro_of(const double* cons, long q, long ne) {
    Micro m;
    m.r = cons[q];
    const double rtv = cons[q + ne];
    m.th = rtv / m.r;
    m.qv = cons[q + ERFX_RHOQ1 * ne] / m.r;
    m.qc = cons[q + ERFX_RHOQ2 * ne] / m.r;
    m.qp = cons[q + ERFX_RHOQ3 * ne] / m.r;
    m.tabs = eos_T(m.r, rtv, m.qv);
    m.pres = eos_p_of_rt_qv(rtv, m.qv) / 100.;
    return m;
}

// saturation and its temperature derivative (erf_qsatw / erf_dtqsatw,
// ERF_Microphysics_Utils.H:141-156)
__device__ __forceinline__ void qsat_of(double tabs, double pressure, double& qsat, double& dtqsat, int* bad) {
    double esat = esatw_dev(tabs, bad);
    qsat = Rd_on_Rv * esat / smax(esat, pressure - esat);
    dtqsat = Rd_on_Rv * dtesatw_dev(tabs, bad) / pressure;
}

// Kessler's local process amounts of one step, rain model (ERF_Kessler.cpp:95-158),
// on the clipped qv, qc, qp
struct Rates { double v2c, c2v, r2v, c2r; };
__device__ __forceinline__ Rates rates_rain(const Micro& m, double qv, double qc, double qp, double dt, int* bad) {
    double qsat, dtqsat;
    qsat_of(m.tabs, m.pres, qsat, dtqsat, bad);
    if (qsat <= 0.0) qsat = 0.0;
    Rates R{0.0, 0.0, 0.0, 0.0};
    const double fac = 1.0 + (L_v / Cp_d) * dtqsat;
    if (qv > qsat) R.v2c = smin(qv, (qv - qsat) / (1.0 + fac));
    if (qv < qsat && qc > 0.0) R.c2v = smin(qc, (qsat - qv) / (1.0 + fac));
    const double r = m.r;
    const double pressure = m.pres;
    if (qp > 0.0 && qv < qsat) {
        double C = 1.6 + 124.9 * erfx_pow(0.001 * r * qp, 0.2046);
        double d = 1.0 / (0.001 * r) * (1.0 - qv / qsat) * C * erfx_pow(0.001 * r * qp, 0.525) /
                   (5.4e5 + 2.55e6 / (pressure * qsat)) * dt;
        R.r2v = smin(qp, d);
    }
    if (qc > 0.0) {
        double qcc = qc;
        double auto_r = 0.0;
        if (qcc > qcw0) auto_r = alphaelq;
        double accrr = 2.2 * erfx_pow(qp, 0.875);
        double d = dt * (accrr * qcc + auto_r * (qcc - qcw0));
        R.c2r = smin(d, qc);
    }
    return R;
}

// sedimentation tendency of rain in a cell from the face fluxes (:160-165)
__device__ __forceinline__ double sed_of(const double* fz, long q, long sz, double r, double dt, double dz,
                                         const double* J) {
    double dJinv = J ? 1.0 / J[q] : 1.0;   // 1/detJ on a terrain grid (ERF_Kessler.cpp:82-94)
    double dq_sed = dt * dJinv * (1.0 / r) * (fz[q + sz] - fz[q]) / dz;
    if (fabs(dq_sed) < 1e-14) dq_sed = 0.0;
    return dq_sed;
}

// KS1: sedimentation flux on faces 0..nz of the valid columns (ERF_Kessler.cpp:27-69)
static void sed_flux(const Params& P, const double* cons, double* fz, double* rain_accum, double dt, hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sz = L.sz;
    const double* rho = cons;
    const double* q3 = cons + ERFX_RHOQ3 * ne;
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

void kessler_advance(const Params& P, double* cons, double* fz, double* rain_accum, double dt, hipStream_t s,
                     int* bad) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sz = L.sz;
    const double dz = P.cfg.dz;
    const double* Jt = P.terrain ? P.detJ : nullptr;
    const double fac_cond = lcond / P.cfg.c_p;    // Kessler::Define (ERF_Kessler.H:65)
    const bool rain = P.cfg.moisture_type == ERFX_MOIST_KESSLER;
    if (rain) sed_flux(P, cons, fz, rain_accum, dt, s);
    const double LvOcp = L_v / Cp_d;
    // KS2: cells
    launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        const Micro m = micro_of(cons, q, ne);
        double th = m.th, qv = m.qv, qc = m.qc, qp = m.qp;
        int flag = 0;
        if (rain) {
            qv = smax(0.0, qv);
            qc = smax(0.0, qc);
            qp = smax(0.0, qp);
            const Rates R = rates_rain(m, qv, qc, qp, dt, &flag);
            const double dq_sed = sed_of(fz, q, sz, m.r, dt, dz, Jt);
            qv += -R.v2c + R.c2v + R.r2v;
            qc += R.v2c - R.c2v - R.c2r;
            qp += dq_sed + R.c2r - R.r2v;
            double theta_over_T = th / m.tabs;
            th += theta_over_T * fac_cond * (R.v2c - R.c2v - R.r2v);
            qv = smax(0.0, qv);
            qc = smax(0.0, qc);
            qp = smax(0.0, qp);
        } else {
            // Kessler_NoRain (:187-248): saturation adjustment only
            qc = smax(0.0, qc);
            double qsat, dtqsat;
            qsat_of(m.tabs, m.pres, qsat, dtqsat, &flag);
            double dq_vapor_to_clwater = 0.0, dq_clwater_to_vapor = 0.0;
            double fac = 1.0 + LvOcp * dtqsat;
            if (qv > qsat) dq_vapor_to_clwater = smin(qv, (qv - qsat) / (1.0 + fac));
            if (qv < qsat && qc > 0.0) dq_clwater_to_vapor = smin(qc, (qsat - qv) / (1.0 + fac));
            qv += -dq_vapor_to_clwater + dq_clwater_to_vapor;
            qc += dq_vapor_to_clwater - dq_clwater_to_vapor;
            double theta_over_T = th / m.tabs;
            th += theta_over_T * fac_cond * (dq_vapor_to_clwater - dq_clwater_to_vapor);
            qv = smax(0.0, qv);
            qc = smax(0.0, qc);
        }
        if (flag) *bad = 1;
        const double r = m.r;
        cons[q + ne] = r * th;
        cons[q + ERFX_RHOQ1 * ne] = r * qv;
        cons[q + ERFX_RHOQ2 * ne] = r * qc;
        cons[q + ERFX_RHOQ3 * ne] = r * qp;
    });
    fill_boundary_periodic(P, cons, P.cfg.ncons, P.ng_cons, -1, s);
}

// ---- samples and the learned step ---------------------------------------------------
// A cell's local processes can be non-zero only when it is supersaturated,
// holds cloud water, or holds rain in subsaturated air (the conditions of
// ERF_Kessler.cpp:121-147); elsewhere all four amounts are exactly 0.
__device__ __forceinline__ bool active_cell(double qv, double qc, double qp, double qsat) {
    return qv > qsat || qc > 0.0 || (qp > 0.0 && qv < qsat);
}

// Kessler's local tendencies of every active cell as (T, p, rho, qv, qc, qr,
// w, dt, dqv, dqc, dqr, dtheta) rows, appended at the device counter.
void kessler_samples(const Params& P, const double* cons, const double* w, double dt, float* out, long max_rows,
                     unsigned long long* counter, int* bad, hipStream_t s) {
    if (rec_refuse("training-row harvest")) return;
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sz = L.sz;
    const double fac_cond = lcond / P.cfg.c_p;
    launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        const Micro m = micro_of(cons, q, ne);
        const double qv = smax(0.0, m.qv), qc = smax(0.0, m.qc), qp = smax(0.0, m.qp);
        int flag = 0;
        double qsat, dtqsat;
        qsat_of(m.tabs, m.pres, qsat, dtqsat, &flag);
        if (qsat <= 0.0) qsat = 0.0;
        if (flag) { *bad = 1; return; }
        if (!active_cell(qv, qc, qp, qsat)) return;
        const Rates R = rates_rain(m, qv, qc, qp, dt, &flag);
        const unsigned long long row = atomicAdd(counter, 1ull);
        if ((long)row >= max_rows) return;
        float* o = out + row * 12;
        const double wc = 0.5 * (w[q] + w[q + sz]);
        o[0] = (float)m.tabs; o[1] = (float)m.pres; o[2] = (float)m.r;
        o[3] = (float)qv; o[4] = (float)qc; o[5] = (float)qp; o[6] = (float)wc; o[7] = (float)dt;
        o[8] = (float)(-R.v2c + R.c2v + R.r2v);
        o[9] = (float)(R.v2c - R.c2v - R.c2r);
        o[10] = (float)(R.c2r - R.r2v);
        o[11] = (float)(m.th / m.tabs * fac_cond * (R.v2c - R.c2v - R.r2v));
    });
}

// The learned step: KS1, then
//   E1  cells: the gate above; active cells inside the training range get a
//       row of scaled inputs for the network, the others fall back to Kessler;
//   E2  the network on the rows (erfx_mlp.hip);
//   E3  cells: network tendencies projected in FP64 so that the vapour change
//       has the sign and bound of Kessler's (toward saturation, never past
//       it), total water is conserved, no species goes negative and theta
//       follows the latent heat of the vapour change; fallback cells run
//       Kessler's own update.
// slot: per-cell int (-1 inactive, -2 fallback, else the row); x, y: rows of
// network inputs and outputs; rows_cell: the cell of each row.
void kessler_mlp_advance(const Params& P, const Net& N, const MlpScaling& S, double* cons, const double* w,
                         double* fz, double* rain_accum, double dt, int* slot, float* x, float* y, long* rows_cell,
                         unsigned long long* counter, unsigned long long* host_counts, int* bad, hipStream_t s) {
    if (rec_refuse("learned microphysics")) return;
    using namespace mlp;
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sz = L.sz;
    const double dz = P.cfg.dz;
    const double* Jt = P.terrain ? P.detJ : nullptr;
    const double fac_cond = lcond / P.cfg.c_p;
    sed_flux(P, cons, fz, rain_accum, dt, s);
    (void)hipMemsetAsync(counter, 0, 2 * sizeof(unsigned long long), s);
    // E1
    launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        const Micro m = micro_of(cons, q, ne);
        const double qv = smax(0.0, m.qv), qc = smax(0.0, m.qc), qp = smax(0.0, m.qp);
        int flag = 0;
        double qsat, dtqsat;
        qsat_of(m.tabs, m.pres, qsat, dtqsat, &flag);
        if (qsat <= 0.0) qsat = 0.0;
        if (flag) *bad = 1;
        if (!active_cell(qv, qc, qp, qsat)) { slot[q] = -1; return; }
        const double f[NI] = {m.tabs, m.pres, m.r, qv, qc, qp, 0.5 * (w[q] + w[q + sz]), dt};
        bool inside = true;
        for (int c = 0; c < NI; ++c) inside = inside && f[c] >= S.x_lo[c] && f[c] <= S.x_hi[c];
        if (!inside) {
            slot[q] = -2;
            atomicAdd(counter + 1, 1ull);
            return;
        }
        const unsigned long long row = atomicAdd(counter, 1ull);
        slot[q] = (int)row;
        rows_cell[row] = q;
        for (int c = 0; c < NI; ++c) x[row * NI + c] = (float)((f[c] - S.x_mean[c]) / S.x_scale[c]);
    });
    (void)hipMemcpyAsync(host_counts, counter, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    const int nrows = (int)host_counts[0];
    // E2
    mlp_forward(N, x, y, nrows, s);
    // E3
    launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        const Micro m = micro_of(cons, q, ne);
        double th = m.th;
        double qv = smax(0.0, m.qv), qc = smax(0.0, m.qc), qp = smax(0.0, m.qp);
        int flag = 0;
        const double dq_sed = sed_of(fz, q, sz, m.r, dt, dz, Jt);
        const int sl = slot[q];
        if (sl == -2) {                       // Kessler's own update
            const Rates R = rates_rain(m, qv, qc, qp, dt, &flag);
            qv += -R.v2c + R.c2v + R.r2v;
            qc += R.v2c - R.c2v - R.c2r;
            qp += dq_sed + R.c2r - R.r2v;
            double theta_over_T = th / m.tabs;
            th += theta_over_T * fac_cond * (R.v2c - R.c2v - R.r2v);
        } else {
            double d[3] = {0.0, 0.0, 0.0};
            if (sl >= 0) {
                for (int c = 0; c < 3; ++c) d[c] = (double)y[(long)sl * NO + c] * S.y_scale[c];
                // Kessler's local processes fix the direction of the vapour
                // exchange (ERF_Kessler.cpp:126-150): in subsaturated air vapour
                // only gains, up to saturation, and cloud water only loses; in
                // supersaturated air vapour only condenses, down to saturation
                // at most, and rain only grows.  The network's vapour change is
                // held to that range; cloud and rain take the opposite change
                // (so total water is conserved), split as the network splits it.
                double qsat, dtqsat;
                qsat_of(m.tabs, m.pres, qsat, dtqsat, &flag);
                if (qsat <= 0.0) qsat = 0.0;
                double dv = d[0];
                if (qv < qsat) dv = smin(smax(dv, 0.0), qsat - qv);
                else dv = smax(smin(dv, 0.0), qsat - qv);
                const double cond = -dv;                     // change of cloud + rain
                double dr = (qv < qsat) ? d[2] : smax(d[2], 0.0);
                double dc = cond - dr;
                if (qv < qsat && dc > 0.0) { dc = 0.0; dr = cond; }
                d[0] = dv;
                d[1] = dc;
                d[2] = dr;
                // non-negativity: raise negative results to 0 and take the
                // deficit from the others in proportion to what they hold
                const double qs[3] = {qv, qc, qp};
                double deficit = 0.0, avail = 0.0;
                for (int c = 0; c < 3; ++c) {
                    const double post = qs[c] + d[c];
                    if (post < 0.0) deficit -= post; else avail += post;
                }
                if (deficit > 0.0) {
                    for (int c = 0; c < 3; ++c) {
                        const double post = qs[c] + d[c];
                        if (post < 0.0) d[c] = -qs[c];
                        else d[c] -= deficit * (post / avail);
                    }
                }
            }
            qv += d[0];
            qc += d[1];
            qp += dq_sed + d[2];
            // latent heat of the vapour change (ERF_Kessler.cpp:172-173)
            th += th / m.tabs * fac_cond * (-d[0]);
        }
        if (flag) *bad = 1;
        qv = smax(0.0, qv);
        qc = smax(0.0, qc);
        qp = smax(0.0, qp);
        const double r = m.r;
        cons[q + ne] = r * th;
        cons[q + ERFX_RHOQ1 * ne] = r * qv;
        cons[q + ERFX_RHOQ2 * ne] = r * qc;
        cons[q + ERFX_RHOQ3 * ne] = r * qp;
    });
    fill_boundary_periodic(P, cons, P.cfg.ncons, P.ng_cons, -1, s);
}

}  // namespace erfx
